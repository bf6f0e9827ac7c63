// Host-side candidate minima of many chi2 rows (no device code).
//
// The fine-fit candidate loop (fitz.py:198-210) walks find_minima(zchi2)
// (fitz.py:46-67): the local minima of a row, edges included, with ties
// between neighbours counted as minima (x[i] <= x[i-1] and x[i] <= x[i+1]),
// sorted by value with np.argsort.  The loop needs only the first few of
// them for most targets, so this pass keeps the k smallest by insertion,
// O(nz) per row without temporaries.  For distinct finite values the order
// is the only one a sort can give.  Where ties or non-finite values could
// leave the order to np.argsort's tie handling (two equal values among the
// first k, a value equal to the k-th one outside them, a NaN or an infinity
// among the candidates) the row is flagged and the caller takes
// find_minima itself for it.
#include <cmath>
#include <cstdint>

#include "common.hpp"

namespace {

constexpr int kBlock = 16;     // samples per block of the first pass
constexpr int kMaxHead = 64;   // head sizes above this are left to the caller

// First pass over the interior samples 1 .. nz-2 (nz >= 3), in loops without
// branches that the compiler vectorises: per block of kBlock samples the
// smallest candidate value (+inf for none), the number of candidates, and
// whether a candidate is NaN or infinite.
void block_minima(const double *__restrict r, long long nz, double *__restrict bmin,
                  long long &n, int &odd) {
    long long cnt = 0;
    int bad = 0;
    const long long ni = nz - 2;
    const long long nfull = ni / kBlock;
    for (long long b = 0; b < nfull; ++b) {
        const double *p = r + 1 + b * kBlock;
        double mn = INFINITY;
        int c = 0, bd = 0;
#pragma unroll
        for (int i = 0; i < kBlock; ++i) {   // fixed length: unrolled and vectorised
            const double v = p[i];
            const int m = (v <= p[i - 1]) & (v <= p[i + 1]);
            c += m;
            bd |= m & !(v - v == 0.0);
            const double w = m ? v : INFINITY;
            mn = w < mn ? w : mn;
        }
        bmin[b] = mn;
        cnt += c;
        bad |= bd;
    }
    if (nfull * kBlock < ni) {               // the last, partial block
        double mn = INFINITY;
        for (long long j = 1 + nfull * kBlock; j < nz - 1; ++j) {
            const double v = r[j];
            const int m = (v <= r[j - 1]) & (v <= r[j + 1]);
            cnt += m;
            bad |= m & !(v - v == 0.0);
            const double w = m ? v : INFINITY;
            mn = w < mn ? w : mn;
        }
        bmin[nfull] = mn;
    }
    n += cnt;
    odd |= bad;
}

// number of interior candidates of row r with value <= c in block b
long long count_le_block(const double *__restrict r, long long nz, long long b, double c) {
    const long long j0 = 1 + b * kBlock;
    const long long j1 = j0 + kBlock < nz - 1 ? j0 + kBlock : nz - 1;
    long long le = 0;
    for (long long j = j0; j < j1; ++j) le += (r[j] <= r[j - 1]) & (r[j] <= r[j + 1]) & (r[j] <= c);
    return le;
}

}  // namespace

extern "C" int rr_minima_head(const double *x, long long nt, long long nz, long long ldx, int k,
                              long long *head, long long *ncand, unsigned char *flag) {
    if (nt < 0 || nz < 0 || k < 0 || (nt > 0 && (!x || !ncand || !flag || (k > 0 && !head))) ||
        ldx < nz) {
        rr::set_error("rr_minima_head: bad arguments (nt %lld, nz %lld, ldx %lld, k %d)", nt, nz, ldx, k);
        return RR_EINVAL;
    }
    const long long nblk = nz > 2 ? (nz - 2 + kBlock - 1) / kBlock : 0;
    double *bmin = nblk ? new double[nblk] : nullptr;
    for (long long t = 0; t < nt; ++t) {
        const double *r = x + t * ldx;
        long long *h = k > 0 ? head + t * k : nullptr;
        double hv[kMaxHead];
        long long walked[kMaxHead];   // blocks walked for the head (-1: none)
        const int kk = k < kMaxHead ? k : kMaxHead;
        int nh = 0;
        long long n = 0;
        int odd = k > kMaxHead;
        // the head: the kk smallest candidates in value order (a strictly
        // smaller value moves up, so equal values keep index order; rows with
        // equal values are flagged anyway)
        auto offer = [&](long long j) {
            const double v = r[j];
            if (kk == 0 || !(v - v == 0.0) || (nh == kk && !(v < hv[kk - 1]))) return;
            int p = nh < kk ? nh++ : kk - 1;
            while (p > 0 && v < hv[p - 1]) {
                hv[p] = hv[p - 1];
                h[p] = h[p - 1];
                --p;
            }
            hv[p] = v;
            h[p] = j;
        };
        auto edge = [&](long long j, int m) {
            n += m;
            odd |= m & !(r[j] - r[j] == 0.0);
            if (m) offer(j);
        };
        if (nz == 1) {
            edge(0, 1);
        } else if (nz == 2) {
            edge(0, r[0] <= r[1]);
            edge(1, r[1] <= r[0]);
        } else if (nz > 2) {
            edge(0, r[0] <= r[1]);
            edge(nz - 1, r[nz - 1] <= r[nz - 2]);
            block_minima(r, nz, bmin, n, odd);
            // the kk smallest candidates lie in the kk blocks with the
            // smallest block minima: walk those blocks (ascending minima)
            for (int i = 0; i < kk; ++i) walked[i] = -1;
            for (int pick = 0; pick < kk; ++pick) {
                long long best = -1;
                double bv = INFINITY;
                for (long long b = 0; b < nblk; ++b)
                    if (bmin[b] < bv) {
                        bv = bmin[b];
                        best = b;
                    }
                if (best < 0 || (nh == kk && !(bv < hv[kk - 1]))) break;
                bmin[best] = INFINITY;
                walked[pick] = best;
                const long long j0 = 1 + best * kBlock;
                const long long j1 = j0 + kBlock < nz - 1 ? j0 + kBlock : nz - 1;
                for (long long j = j0; j < j1; ++j)
                    if (r[j] <= r[j - 1] && r[j] <= r[j + 1]) offer(j);
            }
        }
        for (int i = nh; i < k; ++i) h[i] = -1;
        for (int i = 1; i < nh; ++i) odd |= hv[i] == hv[i - 1];
        // a candidate outside the head with the k-th value: a tie at the cut.
        // Candidates <= the k-th value lie in the walked blocks, the edges, or
        // a block whose minimum (walked blocks now read +inf) is that small.
        if (!odd && nh == kk && n > kk && kk > 0) {
            const double c = hv[kk - 1];
            long long le = ((r[0] <= r[1]) & (r[0] <= c)) + ((r[nz - 1] <= r[nz - 2]) & (r[nz - 1] <= c));
            for (int i = 0; i < kk && nz > 2; ++i)
                if (walked[i] >= 0) le += count_le_block(r, nz, walked[i], c);
            for (long long b = 0; b < nblk; ++b)
                if (bmin[b] <= c) le += count_le_block(r, nz, b, c);
            odd = le > kk;
        }
        ncand[t] = n;
        flag[t] = odd ? 1 : 0;
    }
    delete[] bmin;
    return RR_OK;
}
