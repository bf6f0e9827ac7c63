// FP64 polish of a mixed-precision scan (rr_zscan_polish).
//
// A mixed-precision scan gives chi2 with an absolute error that scales with
// the target's f'Wf, the sum chi2 = f'Wf - 2 c'y + c'Mc cancels against:
// |chi2_mixed - chi2_f64| <= e f'Wf with e about 1e-6 for split fp16, 1e-4
// for fp16, 1e-3 for bf16 (relative to chi2 the error grows with the
// signal-to-noise ratio).  Where two redshifts' chi2 differ by less than
// that, the scan's best redshift can differ from the FP64 scan's.  The
// polish re-evaluates in FP64 every redshift whose mixed chi2, or chi2 +
// penalty, lies within tol * max(f'Wf, |minimum|) of the item's minimum, and
// overwrites those rows (chi2, coefficients, penalty, status).  If the error
// is at most e f'Wf everywhere and tol >= 2e, the FP64 minimum is among the
// re-evaluated rows and every other row stays above it, so the item's argmin
// is the FP64 scan's.
//
// polish_select_kernel (one wavefront per item): per aligned window of 4
//   redshifts the minimum of chi2 and of chi2 + penalty (in LDS), the item's
//   minima, then the windows holding a value under the thresholds; if more
//   than max_windows qualify the tolerance is halved until they fit (so the
//   windows kept are the ones nearest the minimum), and they are written in
//   increasing redshift order.  An item with a single row under the full
//   thresholds gets no window: that row is the FP64 argmin already.
// polish_eval_kernel (4 waves per (item, window), persistent over a work
//   queue): thread (zl, chunk) accumulates the FP64 normal equations of
//   redshift 4 w + zl over the 2-pixel blocks 2 chunk + 128 j (the per-pixel
//   arithmetic of zscan_f64.hip, on the FP64 cache [b][row][z], so the 4
//   lanes of one pixel read one 32-byte run and the workgroup's 64 chunks read
//   one 128-pixel span at a time), the 64 chunks are summed in a fixed order
//   (lane shuffles, then the four waves through LDS), and 4 lanes solve: LU
//   with
//   partial pivoting, chi2 = f'Wf - 2 c'y + c'Mc compensated, HUGE_CHI2 rules
//   as zscan_finish.hip.  One workgroup per CU walks the queue, so windows
//   spread over every XCD and items without a window cost nothing.  A
//   window's time is latency (one pass over the pixels at one wave per
//   SIMD), hence narrow windows and many pixel chunks.
//
// No counterpart in the reference (its scan is FP64 throughout).
#include "common.hpp"
#include "solve.hpp"
#include "zscan_args.hpp"

namespace {

constexpr double HUGE_CHI2 = 9999999.0;   // zscan.py:18
constexpr int WZ = RR_POLISH_WINDOW;      // redshifts per window

__device__ __forceinline__ double wave_min(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o, 64));
    return v;
}

__device__ __forceinline__ int wave_sum(int v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

__global__ __launch_bounds__(64) void polish_select_kernel(rr::ZArgs a, rr::PolishArgs p) {
    extern __shared__ double wmin[];     // [2][nw_all]: window minima of chi2, chi2 + penalty
    const int item = blockIdx.x;
    const int lane = threadIdx.x;
    const int nzi = a.item_nz[item];
    const long long row0 = a.item_out[item];
    const int nw = (nzi + WZ - 1) / WZ;
    double *wm1 = wmin, *wm2 = wmin + p.nw_all;
    const double inf = __builtin_huge_val();
    double m1 = inf, m2 = inf;
    // rows that hold the HUGE_CHI2 convention (singular solve: status != OK)
    // take no part; a target without weight is HUGE_CHI2 everywhere.  A real
    // chi2 may exceed HUGE_CHI2 (bright targets), so the value itself is no test.
    const double sumw = static_cast<const double *>(a.work)[2 * item + 1];
    for (int k = lane; k < nw; k += 64) {
        double w1 = inf, w2 = inf;
        for (int i = 0; i < WZ; ++i) {
            const int iz = k * WZ + i;
            if (iz < nzi && (!a.status || a.status[row0 + iz] == RR_STATUS_OK)) {
                const double c = a.zchi2[row0 + iz];
                const double s = c + (a.penalty ? a.penalty[row0 + iz] : 0.0);
                w1 = fmin(w1, c);
                w2 = fmin(w2, s);
            }
        }
        wm1[k] = w1;
        wm2[k] = w2;
        m1 = fmin(m1, w1);
        m2 = fmin(m2, w2);
    }
    m1 = wave_min(m1);
    m2 = wave_min(m2);
    __syncthreads();
    int nsel = 0;
    if (sumw > 0.0 && m1 < inf) {
        // error scale of a mixed-precision chi2: the item's f'Wf (the sum the
        // chi2 cancels against), at least |minimum|
        const double fwf = static_cast<const double *>(a.work)[2 * item];
        double d1 = p.tol * fmax(fwf, fabs(m1)), d2 = p.tol * fmax(fwf, fabs(m2));
        double t1 = m1 + d1, t2 = m2 + d2;
        int it = 0;
        for (;; ++it) {
            t1 = m1 + d1;
            t2 = m2 + d2;
            int cnt = 0;
            for (int k = lane; k < nw; k += 64) cnt += (wm1[k] <= t1 || wm2[k] <= t2) ? 1 : 0;
            cnt = wave_sum(cnt);
            if (cnt <= p.max_windows || (d1 == 0.0 && d2 == 0.0)) break;
            if (it >= 60) {
                d1 = 0.0;
                d2 = 0.0;
            } else {
                d1 *= 0.5;
                d2 *= 0.5;
            }
        }
        // An isolated minimum (one row under the full tolerance) is already the
        // FP64 argmin: nothing to re-evaluate.
        int rows = 0;
        if (it == 0) {
            for (int iz = lane; iz < nzi; iz += 64) {
                if (a.status && a.status[row0 + iz] != RR_STATUS_OK) continue;
                const double c = a.zchi2[row0 + iz];
                const double s = c + (a.penalty ? a.penalty[row0 + iz] : 0.0);
                rows += (c <= t1 || s <= t2) ? 1 : 0;
            }
            rows = wave_sum(rows);
        }
        if (it == 0 && rows <= 1) t1 = t2 = -__builtin_huge_val();   // selects no window
        int base = 0;
        for (int k0 = 0; k0 < nw; k0 += 64) {
            const int k = k0 + lane;
            const bool sel = k < nw && (wm1[k] <= t1 || wm2[k] <= t2);
            const unsigned long long b = __ballot(sel);
            const int pos = base + __popcll(b & ((1ull << lane) - 1ull));
            if (sel && pos < p.max_windows) p.windows[(long long)item * p.max_windows + pos] = k;
            base += __popcll(b);
        }
        nsel = base < p.max_windows ? base : p.max_windows;
    }
    if (lane == 0) p.nwindows[item] = nsel;
    // queue entries item * RR_POLISH_MAX_WINDOWS + slot for the eval kernel
    int q0 = 0;
    if (lane == 0 && nsel > 0) q0 = atomicAdd(p.queue, nsel);
    q0 = __shfl(q0, 0, 64);
    if (lane < nsel) p.queue[1 + q0 + lane] = item * RR_POLISH_MAX_WINDOWS + lane;
}

__global__ void polish_reset_kernel(int *queue) { queue[0] = 0; }

constexpr int kEvalWaves = 4;                   // 8 would cap a wave at 256 registers and spill
constexpr int kChunks = kEvalWaves * 64 / WZ;   // 64 pixel chunks per redshift

template <int NB>
__device__ __forceinline__ void eval_window(const rr::ZArgs &a, const rr::PolishArgs &p, int item,
                                            int slot, double (&red)[kEvalWaves][WZ][NB * (NB + 1) / 2 + NB]);

// Persistent over the work queue that polish_select_kernel filled (one
// workgroup per CU): no workgroup is launched for items without a window.
template <int NB>
__global__ __launch_bounds__(kEvalWaves * 64) void polish_eval_kernel(rr::ZArgs a, rr::PolishArgs p) {
    constexpr int NV = NB * (NB + 1) / 2 + NB;
    __shared__ double red[kEvalWaves][WZ][NV];

    const int nwork = __builtin_amdgcn_readfirstlane(*(volatile const int *)p.queue);
    for (int wk = blockIdx.x; wk < nwork; wk += gridDim.x) {   // uniform over the workgroup
        const int entry = p.queue[1 + wk];
        eval_window<NB>(a, p, entry / RR_POLISH_MAX_WINDOWS, entry % RR_POLISH_MAX_WINDOWS, red);
        __syncthreads();                     // red is reused by the next window
    }
}

// One (item, window): FP64 normal equations of its 16 redshifts, solve, chi2.
template <int NB>
__device__ __forceinline__ void eval_window(const rr::ZArgs &a, const rr::PolishArgs &p, int item,
                                            int slot, double (&red)[kEvalWaves][WZ][NB * (NB + 1) / 2 + NB]) {
    constexpr int NM = NB * (NB + 1) / 2;
    constexpr int NV = NM + NB;
    constexpr int P = 2;
    const int tid = threadIdx.x;
    const int zl = tid & (WZ - 1);
    const int chunk = tid / WZ;              // 0..kChunks-1
    const int wave = tid >> 6;
    const int nzi = a.item_nz[item];
    const int iz = p.windows[(long long)item * p.max_windows + slot] * WZ + zl;
    const int zc = a.item_z0[item] + (iz < nzi ? iz : nzi - 1);
    const int t = a.item_target[item];

    const long long npp = a.npix_pad;
    const double *__restrict__ Rt = a.rband + (long long)t * npp * 11;
    const double *__restrict__ Wt = a.w + (long long)t * npp;
    const double *__restrict__ WFt = a.wf + (long long)t * npp;
    const long long ldz = p.ldz;
    const long long bstride = a.cache_rows * ldz;
    const double *Tz = p.tcache64 + (long long)(a.halo - 5) * ldz + zc;

    double v[NV];
#pragma unroll
    for (int i = 0; i < NV; ++i) v[i] = 0.0;

    // chunk c takes the 2-pixel blocks 2c + 128j: at any time the workgroup's
    // chunks cover one 128-pixel span and share their window rows in L1
    for (int p0 = P * chunk; p0 < a.npix_pad; p0 += P * kChunks) {
        double r[P][11];
#pragma unroll
        for (int i = 0; i < P; ++i)
#pragma unroll
            for (int k = 0; k < 11; ++k) r[i][k] = Rt[(long long)(p0 + i) * 11 + k];
        double A[P][NB];
#pragma unroll
        for (int b = 0; b < NB; ++b) {
            const double *col = Tz + b * bstride + (long long)p0 * ldz;
            double win[P + 10];
#pragma unroll
            for (int i = 0; i < P + 10; ++i) win[i] = col[i * ldz];
#pragma unroll
            for (int i = 0; i < P; ++i) {
                double s = r[i][0] * win[i];
#pragma unroll
                for (int k = 1; k < 11; ++k) s = fma(r[i][k], win[i + k], s);
                A[i][b] = s;
            }
        }
#pragma unroll
        for (int i = 0; i < P; ++i) {
            const double w = Wt[p0 + i];
            const double wf = WFt[p0 + i];
            int idx = 0;
#pragma unroll
            for (int b1 = 0; b1 < NB; ++b1) {
                const double wa = w * A[i][b1];
                v[NM + b1] = fma(A[i][b1], wf, v[NM + b1]);
#pragma unroll
                for (int b2 = b1; b2 < NB; ++b2) {
                    v[idx] = fma(wa, A[i][b2], v[idx]);
                    ++idx;
                }
            }
        }
    }

    // this wave's 64 / WZ chunks sit in lanes zl + WZ c: butterfly over c
#pragma unroll
    for (int i = 0; i < NV; ++i) {
#pragma unroll
        for (int o = WZ; o < 64; o <<= 1) v[i] += __shfl_xor(v[i], o, 64);
    }
    if ((tid & 63) < WZ) {
#pragma unroll
        for (int i = 0; i < NV; ++i) red[wave][zl][i] = v[i];
    }
    __syncthreads();
    if (tid >= WZ || iz >= nzi) return;      // only the end of this window's work

#pragma unroll
    for (int i = 0; i < NV; ++i) {
        double s = red[0][zl][i];
#pragma unroll
        for (int w = 1; w < kEvalWaves; ++w) s += red[w][zl][i];
        v[i] = s;
    }

    const long long row = a.item_out[item] + iz;
    const double *st = static_cast<const double *>(a.work);
    const double fwf = st[2 * item];
    const double sumw = st[2 * item + 1];
    double M[NB][NB], y[NB], c[NB];
    {
        int idx = 0;
#pragma unroll
        for (int b1 = 0; b1 < NB; ++b1) {
#pragma unroll
            for (int b2 = b1; b2 < NB; ++b2) {
                M[b1][b2] = v[idx];
                M[b2][b1] = v[idx];
                ++idx;
            }
            y[b1] = v[NM + b1];
        }
    }
    if (a.prior) {
#pragma unroll
        for (int i = 0; i < NB; ++i)
#pragma unroll
            for (int j = 0; j < NB; ++j) M[i][j] += a.prior[i * NB + j];
    }
#pragma unroll
    for (int i = 0; i < NB; ++i) c[i] = y[i];
    const bool ok = rr::lu_solve<NB>(M, c);
    unsigned char status = RR_STATUS_OK;
    double chi2;
    if (sumw == 0.0) {
        chi2 = HUGE_CHI2;
#pragma unroll
        for (int b = 0; b < NB; ++b) c[b] = 0.0;
    } else if (!ok) {
        chi2 = HUGE_CHI2;
        status = RR_STATUS_SINGULAR;
#pragma unroll
        for (int b = 0; b < NB; ++b) c[b] = 0.0;
    } else {
        // t = M c with the unregularised M
        double tv[NB];
#pragma unroll
        for (int i = 0; i < NB; ++i) {
            double s = 0.0;
#pragma unroll
            for (int j = 0; j < NB; ++j) {
                const int lo = i <= j ? i : j, hi = i <= j ? j : i;
                s = fma(v[lo * NB - lo * (lo - 1) / 2 + (hi - lo)], c[j], s);
            }
            tv[i] = s;
        }
        chi2 = rr::compensated_chi2<NB>(fwf, c, y, tv);
    }
    a.zchi2[row] = chi2;
#pragma unroll
    for (int b = 0; b < NB; ++b) a.zcoeff[row * NB + b] = c[b];
    if (a.penalty) {
        double oii = 0.0;
        if (a.oiisum) {
#pragma unroll
            for (int b = 0; b < NB; ++b) oii = fma(c[b], a.oiisum[b], oii);
        }
        a.penalty[row] = (oii < 0.0) ? -oii : 0.0;
    }
    if (a.status) a.status[row] = status;
}

template <int NB>
int launch_eval(const rr::ZArgs &a, const rr::PolishArgs &p, hipStream_t s) {
    static int cus = 0;
    if (cus == 0) {
        int dev = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus < 1)
            cus = 256;
    }
    hipLaunchKernelGGL((polish_eval_kernel<NB>), dim3(cus), dim3(kEvalWaves * 64), 0, s, a, p);
    return rr::check_launch("rr_zscan_polish(eval)");
}

}  // namespace

namespace rr {

int zscan_polish(const ZArgs &a, const PolishArgs &p, hipStream_t s) {
    hipLaunchKernelGGL(polish_reset_kernel, dim3(1), dim3(1), 0, s, p.queue);
    int rc = check_launch("rr_zscan_polish(reset)");
    if (rc) return rc;
    const long long lds = 2LL * p.nw_all * 8;
    hipLaunchKernelGGL(polish_select_kernel, dim3(a.nitems), dim3(64), (size_t)lds, s, a, p);
    rc = check_launch("rr_zscan_polish(select)");
    if (rc) return rc;
    switch (a.nb) {
        case 1: return launch_eval<1>(a, p, s);
        case 2: return launch_eval<2>(a, p, s);
        case 3: return launch_eval<3>(a, p, s);
        case 4: return launch_eval<4>(a, p, s);
        case 5: return launch_eval<5>(a, p, s);
        case 6: return launch_eval<6>(a, p, s);
        case 7: return launch_eval<7>(a, p, s);
        case 8: return launch_eval<8>(a, p, s);
        case 9: return launch_eval<9>(a, p, s);
        case 10: return launch_eval<10>(a, p, s);
        case 11: return launch_eval<11>(a, p, s);
        case 12: return launch_eval<12>(a, p, s);
        default: return fail(RR_EUNSUPPORTED, "rr_zscan_polish: nb=%d unsupported", a.nb);
    }
}

}  // namespace rr
