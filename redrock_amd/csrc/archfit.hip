// Archetype fit: redrock's archetype mode, the final fit at every minimum.
//
// For one target at one redshift and one archetype a (or na of them, the
// nearest-neighbour form) the design matrix is
//   A = [R a_1 .. R a_na | R L_1 .. R L_nleg]
// with a_j the archetype rebinned to the pixel grid and multiplied by the
// IGM factor, L_l the Legendre columns (per camera or shared) and R the
// target's 11-diagonal resolution band.  M = A'WA + prior, y = A'Wf; the
// coefficients solve min |M c - y|^2 with c_j >= 0 for the archetype columns
// (scipy's lsq_linear 'bvls' applied to (M, y), zscan.py:1175-1201), and
// chi2 = sum w (f - A c)^2 (zscan.py:777-780).
//
// Reference path replaced: Archetype.get_best_archetype /
// nearest_neighbour_model (archetypes.py:241-400) over
// per_camera_coeff_with_least_square_batch (zscan.py:245-348).
//
// Kernels:
//   arch_prep_kernel   per target, z-independent: w (R L), (RL)'W(RL),
//                      (RL)'Wf, f'Wf.  One workgroup per target.
//   arch_geom_kernel   (fused rebin) per (pair, pixel): the sample index j of
//                      the bin's lower edge and the last interior sample jend
//                      (rebin.py:142,154), the same for every archetype.
//   arch_accum_kernel  per (pair, problem): the archetype rows of M and y.
//                      A wavefront is one pair (one target, wave-uniform R, w,
//                      w f and w R L: scalar loads) and 64 problems, one per
//                      lane; with one archetype per problem a lane's column
//                      is archetype `lane`, so every cache row is one
//                      coalesced 64 x 8-byte run.  FP64 vector FMA, 13 + nleg
//                      FMAs per problem and pixel; the cache is read once.
//                      With the fused rebin the window rows are bin means
//                      computed here from the archetype fluxes stored
//                      wavelength-major (a row of samples is again one
//                      coalesced run across the lanes), each new row once,
//                      the window sliding by 4 pixels; no cache is written.
//   arch_solve_kernel  per problem: assemble M and y, bounded least squares
//                      by active sets (Householder QR of the free columns),
//                      chi2 as c'M0c - 2c'y + f'Wf with M0 = M - prior.
#include "common.hpp"
#include "solve.hpp"

// the bin-mean expressions match rr_rebin's double for double; every other
// product-sum here is an explicit fma()
#pragma clang fp contract(off)

namespace {

constexpr int kMaxNA = RR_ARCH_MAX_NA;
constexpr int kMaxNL = RR_ARCH_MAX_NLEG;

struct ArchArgs {
    int npair, nprob, nbound, npix_pad, halo, narch;
    long long ld, nproblems;
    const int *pair_target;
    const double *acache;
    const int *columns;
    const double *rband, *w, *wf;
    const double *wrl, *gll, *yl, *fwf;
    const double *prior;
    const double *wave;
    const double *zigm;
    rr_lyman_line lines[RR_MAX_LYMAN_LINES];
    int nlines;
    double *coeff, *chi2;
    double *ne;   // (K, nproblems), structure of arrays
    // fused rebin
    const double *x, *yt, *pe_lo, *pe_hi, *zrebin;
    long long ldy;
    int nx, npix;
    // fused rebin, per (pair, row) for rows 0 .. npix_pad + kGeomPad - 1:
    //   geom (k0, n >= 0): the bin mean is sum_k wts[k] y[k0 + k], k < kNW
    //   geom (j, -3 - jend): more than kNW samples, the exact expressions
    int2 *geom;
    double *wts;  // (npair, npix_pad + kGeomPad, kNW)
};

constexpr int kNW = 6;          // sample weights per bin (4 interior samples)
constexpr int kGeomPad = 8;     // zero rows after npix_pad: the sliding window reads ahead

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// ---------------------------------------------------------------- prep
template <int NL>
__global__ __launch_bounds__(256) void arch_prep_kernel(const double *__restrict__ rband,
                                                        const double *__restrict__ w,
                                                        const double *__restrict__ f, int npp,
                                                        const double *__restrict__ leg,
                                                        double *__restrict__ wrl, double *__restrict__ gll,
                                                        double *__restrict__ yl, double *__restrict__ fwf) {
    constexpr int NG = NL * (NL + 1) / 2;
    constexpr int K = NG + NL + 1;
    __shared__ double part[4][K];
    const int t = blockIdx.x;
    const double *Rt = rband + (long long)t * npp * 11;
    const double *Wt = w + (long long)t * npp;
    const double *Ft = f + (long long)t * npp;
    double acc[K];
#pragma unroll
    for (int k = 0; k < K; ++k) acc[k] = 0.0;
    for (int p = threadIdx.x; p < npp; p += 256) {
        const double wv = Wt[p], fv = Ft[p];
        double rl[NL > 0 ? NL : 1];
#pragma unroll
        for (int l = 0; l < NL; ++l) {
            double s = 0.0;
#pragma unroll
            for (int k = 0; k < 11; ++k) {
                const int q = p + k - 5;
                const double lv = (q >= 0 && q < npp) ? leg[(long long)l * npp + q] : 0.0;
                s = fma(Rt[(long long)p * 11 + k], lv, s);
            }
            rl[l] = s;
            wrl[((long long)t * NL + l) * npp + p] = wv * s;
        }
        int g = 0;
#pragma unroll
        for (int l1 = 0; l1 < NL; ++l1)
#pragma unroll
            for (int l2 = l1; l2 < NL; ++l2) acc[g++] += wv * rl[l1] * rl[l2];
#pragma unroll
        for (int l = 0; l < NL; ++l) acc[NG + l] += wv * fv * rl[l];
        acc[K - 1] += wv * fv * fv;
    }
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const double s = wave_sum(acc[k]);
        if (lane == 0) part[wid][k] = s;
    }
    __syncthreads();
    for (int k = threadIdx.x; k < K; k += 256) {
        const double s = (part[0][k] + part[1][k]) + (part[2][k] + part[3][k]);
        if (k == K - 1) {
            fwf[t] = s;
        } else if (k >= NG) {
            yl[(long long)t * NL + (k - NG)] = s;
        } else {
            // k -> (l1, l2), l1 <= l2, row-major upper triangle
            int l1 = 0, rem = k;
            while (rem >= NL - l1) { rem -= NL - l1; ++l1; }
            const int l2 = l1 + rem;
            gll[((long long)t * NL + l1) * NL + l2] = s;
            gll[((long long)t * NL + l2) * NL + l1] = s;
        }
    }
}

// ---------------------------------------------------------------- fused rebin
__global__ __launch_bounds__(256) void arch_geom_kernel(ArchArgs a) {
    const int rowsg = a.npix_pad + kGeomPad;
    const long long idx = (long long)blockIdx.x * 256 + threadIdx.x;
    const long long n = (long long)a.npair * rowsg;
    if (idx >= n) return;
    const int pair = (int)(idx / rowsg);
    const int r = (int)(idx - (long long)pair * rowsg);
    int2 g = make_int2(0, 0);                 // zero row: no weights
    double w[kNW];
#pragma unroll
    for (int k = 0; k < kNW; ++k) w[k] = 0.0;
    if (r < a.npix) {
        const double *x = a.x;
        const int nx = a.nx;
        const double opz = 1. + a.zrebin[pair];
        const double elo = a.pe_lo[r], ehi = a.pe_hi[r];
        int lo = 0, hi = nx;                  // first x[j]*(1+z) above the lower edge
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (x[mid] * opz <= elo) lo = mid + 1;
            else hi = mid;
        }
        const int j = lo < 1 ? 1 : (lo > nx - 1 ? nx - 1 : lo);
        int jend = -1;
        if (j < nx - 1 && x[j] * opz < ehi) {     // x[jend + 1] stays in bounds
            jend = j;
            while (jend + 1 < nx - 1 && x[jend + 1] * opz < ehi) ++jend;
        }
        // the trapezoid of rebin.py:144-176 as weights of the samples
        // j-1 .. jend+1 (a linear map of y), divided by the bin width
        const double xj = x[j] * opz, xjm1 = x[j - 1] * opz;
        const double width = ehi - elo;
        if (jend < 0) {
            const double u = (0.5 * (elo + ehi) - xj) / (xj - xjm1);
            w[0] = -u;
            w[1] = 1.0 + u;
            g = make_int2(j - 1, 2);
        } else if (jend - j + 3 <= kNW) {
            const double t0 = (elo - xjm1) / (xj - xjm1), d0 = xj - elo;
            double c[kNW];
#pragma unroll
            for (int k = 0; k < kNW; ++k) c[k] = 0.0;
            c[0] = 0.5 * d0 * (1.0 - t0);
            c[1] = 0.5 * d0 * (1.0 + t0);
            double xprev = xj;
            for (int jj = j + 1; jj <= jend; ++jj) {
                const double xc = x[jj] * opz, h = 0.5 * (xc - xprev);
                const int m = jj - j;              // weight slot of sample jj - 1
#pragma unroll
                for (int k = 0; k < kNW - 1; ++k)
                    if (k == m) {
                        c[k] += h;
                        c[k + 1] += h;
                    }
                xprev = xc;
            }
            const double xe1 = x[jend + 1] * opz, d1 = ehi - xprev;
            const double t1 = d1 / (xe1 - xprev);
            const int m = jend - j + 1;            // weight slot of sample jend
#pragma unroll
            for (int k = 0; k < kNW - 1; ++k)
                if (k == m) {
                    c[k] += 0.5 * d1 * (2.0 - t1);
                    c[k + 1] += 0.5 * d1 * t1;
                }
#pragma unroll
            for (int k = 0; k < kNW; ++k) w[k] = c[k] / width;
            g = make_int2(j - 1, jend - j + 3);
        } else {
            g = make_int2(j, -3 - jend);          // exact expressions in the fit kernel
        }
    }
    a.geom[idx] = g;
#pragma unroll
    for (int k = 0; k < kNW; ++k) a.wts[idx * kNW + k] = w[k];
}

// Bin mean of archetype column `col` for geometry g at 1+z = opz (the
// expressions of rr_rebin, rebin.py:144-176).
__device__ __forceinline__ double bin_mean(const ArchArgs &a, int2 g, double opz, double elo, double ehi,
                                           int col) {
    if (g.y == -2) return 0.0;
    const double *x = a.x;
    const double *Y = a.yt + col;
    const long long ld = a.ldy;
    const int j = g.x;
    const double xj = x[j] * opz;
    const double xjm1 = x[j - 1] * opz;
    const double width = ehi - elo;
    const double yj = Y[j * ld], yjm1 = Y[(j - 1) * ld];
    double r = 0.0;
    if (g.y >= 0) {
        const int jend = g.y;
        const double yedge = yjm1 + (elo - xjm1) * (yj - yjm1) / (xj - xjm1);
        r += 0.5 * (yj + yedge) * (xj - elo);
        double xprev = xj, yprev = yj;
        for (int jj = j + 1; jj <= jend; ++jj) {
            const double xc = x[jj] * opz;
            const double yc = Y[jj * ld];
            r += 0.5 * (yc + yprev) * (xc - xprev);
            xprev = xc;
            yprev = yc;
        }
        const double xe1 = x[jend + 1] * opz;
        const double ye1 = Y[(jend + 1) * ld];
        const double yedge2 = yprev + (ehi - xprev) * (ye1 - yprev) / (xe1 - xprev);
        r += 0.5 * (yedge2 + yprev) * (ehi - xprev);
    } else {
        const double ylo = yj + (elo - xj) * (yj - yjm1) / (xj - xjm1);
        const double yhi = yj + (ehi - xj) * (yj - yjm1) / (xj - xjm1);
        r += 0.5 * (ylo + yhi) * width;
    }
    return r / width;
}

// ---------------------------------------------------------------- accumulate
__device__ __forceinline__ double igm_factor(const ArchArgs &a, double lobs, double opz) {
    double T = 1.0;
    const double lrest = (1.0 / opz) * lobs;
    for (int l = 0; l < a.nlines; ++l) {
        const double line = a.lines[l].line;
        if (lrest < line) {
            const double zpix = lobs / line - 1.;
            T *= exp(-(a.lines[l].A * pow(1. + zpix, a.lines[l].B)));
        }
    }
    return T;
}

template <int NA, int NL, bool FUSED>
__global__ __launch_bounds__(256) void arch_accum_kernel(ArchArgs a) {
    constexpr int P = 4;
    constexpr int NG = NA * (NA + 1) / 2;
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int pair = blockIdx.y * 4 + wave;
    if (pair >= a.npair) return;
    const int q = blockIdx.x * 64 + lane;
    const bool active = q < a.nprob;
    const int qq = active ? q : a.nprob - 1;
    const long long prob = (long long)pair * a.nprob + qq;
    int col[NA];
#pragma unroll
    for (int j = 0; j < NA; ++j) {
        int c = a.columns ? a.columns[prob * NA + j] : qq;
        col[j] = c < 0 ? 0 : (c >= a.narch ? a.narch - 1 : c);
    }
    const int t = a.pair_target[pair];
    const int npp = a.npix_pad;
    const double *__restrict__ Rt = a.rband + (long long)t * npp * 11;
    const double *__restrict__ Wt = a.w + (long long)t * npp;
    const double *__restrict__ WFt = a.wf + (long long)t * npp;
    const double *__restrict__ WRL = a.wrl + (long long)t * NL * npp;
    const long long rows = npp + 2LL * a.halo;
    const double *__restrict__ Ap = FUSED ? nullptr : a.acache + (long long)pair * rows * a.ld;
    const double zi = a.zigm ? a.zigm[pair] : 0.0;
    const bool igm = a.nlines > 0 && zi == zi;
    const double opz = 1.0 + zi;
    const double opzr = FUSED ? 1.0 + a.zrebin[pair] : 1.0;
    const long long growb = (long long)pair * (npp + kGeomPad);
    const int2 *G = FUSED ? a.geom + growb : nullptr;
    const double *Wg = FUSED ? a.wts + growb * kNW : nullptr;
    // fused: the window of rows p0-5 .. p0+P+4 slides by P; row i enters once.
    // A row is sum_k w_k y[k0 + k]: the weights and k0 are the same for every
    // lane (scalar loads), the samples one coalesced load each.
    double fw[NA][P + 10];
    auto fused_row = [&](int pix, int j) -> double {
        if (pix < 0) return 0.0;
        const int2 gg = G[pix];
        double v = 0.0;
        if (gg.y >= 0) {
            const double *Yc = a.yt + col[j];
#pragma unroll
            for (int k = 0; k < kNW; ++k) {
                const int sidx = gg.x + k < a.nx ? gg.x + k : a.nx - 1;
                v = fma(Wg[(long long)pix * kNW + k], Yc[(long long)sidx * a.ldy], v);
            }
        } else {
            v = bin_mean(a, make_int2(gg.x, -3 - gg.y), opzr, a.pe_lo[pix], a.pe_hi[pix], col[j]);
        }
        return (igm && pix < a.npix) ? v * igm_factor(a, a.wave[pix], opz) : v;
    };
    if (FUSED) {
#pragma unroll
        for (int j = 0; j < NA; ++j)
#pragma unroll
            for (int i = 0; i < 10; ++i) fw[j][i + P] = fused_row(i - 5, j);
    }

    double gaa[NG], ya[NA], gal[NA][NL > 0 ? NL : 1];
#pragma unroll
    for (int k = 0; k < NG; ++k) gaa[k] = 0.0;
#pragma unroll
    for (int j = 0; j < NA; ++j) {
        ya[j] = 0.0;
#pragma unroll
        for (int l = 0; l < NL; ++l) gal[j][l] = 0.0;
    }

    for (int p0 = 0; p0 < npp; p0 += P) {
        double r[P][11];
#pragma unroll
        for (int i = 0; i < P; ++i)
#pragma unroll
            for (int k = 0; k < 11; ++k) r[i][k] = Rt[(long long)(p0 + i) * 11 + k];
        double T[P + 10];
#pragma unroll
        for (int i = 0; i < P + 10; ++i) T[i] = 1.0;
        if (!FUSED && igm) {
#pragma unroll
            for (int i = 0; i < P + 10; ++i) {
                const int pix = p0 - 5 + i;
                if (pix >= 0 && pix < npp) T[i] = igm_factor(a, a.wave[pix], opz);
            }
        }
        if (FUSED) {
#pragma unroll
            for (int j = 0; j < NA; ++j) {
#pragma unroll
                for (int i = 0; i < 10; ++i) fw[j][i] = fw[j][i + P];
#pragma unroll
                for (int i = 0; i < P; ++i) fw[j][10 + i] = fused_row(p0 + 5 + i, j);
            }
        }
        double ra[P][NA];
#pragma unroll
        for (int j = 0; j < NA; ++j) {
            const double *colp = FUSED ? nullptr : Ap + (long long)(a.halo + p0 - 5) * a.ld + col[j];
            double win[P + 10];
#pragma unroll
            for (int i = 0; i < P + 10; ++i) win[i] = FUSED ? fw[j][i] : colp[(long long)i * a.ld] * T[i];
#pragma unroll
            for (int i = 0; i < P; ++i) {
                double s = r[i][0] * win[i];
#pragma unroll
                for (int k = 1; k < 11; ++k) s = fma(r[i][k], win[i + k], s);
                ra[i][j] = s;
            }
        }
#pragma unroll
        for (int i = 0; i < P; ++i) {
            const double wv = Wt[p0 + i], wfv = WFt[p0 + i];
            int g = 0;
#pragma unroll
            for (int j1 = 0; j1 < NA; ++j1) {
                const double wr = wv * ra[i][j1];
#pragma unroll
                for (int j2 = j1; j2 < NA; ++j2, ++g) gaa[g] = fma(wr, ra[i][j2], gaa[g]);
                ya[j1] = fma(wfv, ra[i][j1], ya[j1]);
            }
#pragma unroll
            for (int l = 0; l < NL; ++l) {
                const double v = WRL[(long long)l * npp + p0 + i];
#pragma unroll
                for (int j = 0; j < NA; ++j) gal[j][l] = fma(ra[i][j], v, gal[j][l]);
            }
        }
    }
    if (!active) return;
    double *ne = a.ne + prob;
    const long long S = a.nproblems;
    int k = 0;
#pragma unroll
    for (int g = 0; g < NG; ++g) ne[(k++) * S] = gaa[g];
#pragma unroll
    for (int j = 0; j < NA; ++j)
#pragma unroll
        for (int l = 0; l < NL; ++l) ne[(k++) * S] = gal[j][l];
#pragma unroll
    for (int j = 0; j < NA; ++j) ne[(k++) * S] = ya[j];
}

// ---------------------------------------------------------------- solve
template <int NA, int NL>
__global__ __launch_bounds__(64) void arch_solve_kernel(ArchArgs a) {
    constexpr int NB = NA + NL;
    constexpr int NG = NA * (NA + 1) / 2;
    const long long prob = (long long)blockIdx.x * 64 + threadIdx.x;
    if (prob >= a.nproblems) return;
    const int pair = (int)(prob / a.nprob);
    const int t = a.pair_target[pair];
    const long long S = a.nproblems;
    const double *ne = a.ne + prob;

    double M[NB][NB], y[NB];
    int k = 0;
#pragma unroll
    for (int j1 = 0; j1 < NA; ++j1)
#pragma unroll
        for (int j2 = j1; j2 < NA; ++j2) {
            const double v = ne[(k++) * S];
            M[j1][j2] = v;
            M[j2][j1] = v;
        }
#pragma unroll
    for (int j = 0; j < NA; ++j)
#pragma unroll
        for (int l = 0; l < NL; ++l) {
            const double v = ne[(k++) * S];
            M[j][NA + l] = v;
            M[NA + l][j] = v;
        }
#pragma unroll
    for (int j = 0; j < NA; ++j) y[j] = ne[(k++) * S];
#pragma unroll
    for (int l1 = 0; l1 < NL; ++l1) {
        y[NA + l1] = a.yl[(long long)t * NL + l1];
#pragma unroll
        for (int l2 = 0; l2 < NL; ++l2) M[NA + l1][NA + l2] = a.gll[((long long)t * NL + l1) * NL + l2];
    }
    if (a.prior) {
#pragma unroll
        for (int i = 0; i < NB; ++i)
#pragma unroll
            for (int j = 0; j < NB; ++j) M[i][j] += a.prior[i * NB + j];
    }
    (void)NG;

    // columns of M that are exactly zero: coefficient 0 (minimum-norm solution)
    unsigned live = 0;
#pragma unroll
    for (int j = 0; j < NB; ++j) {
        bool nz = false;
#pragma unroll
        for (int i = 0; i < NB; ++i) nz = nz || (M[i][j] != 0.0);
        if (nz) live |= 1u << j;
    }

    // active sets over the bounded columns: the first free set whose solution
    // is feasible and whose fixed columns have a non-negative gradient of
    // |Mc - y|^2 is the (unique) minimiser; the best feasible one otherwise
    double c[NB], best[NB];
    double best_obj = -1.0;
    bool done = false;
    const int nbound = a.nbound;
    for (unsigned S0 = 0; S0 < (1u << nbound) && !done; ++S0) {
        const unsigned Fm = live & ~S0;
        rr::ls_free<NB>(M, y, Fm, c);
        bool feasible = true;
#pragma unroll
        for (int j = 0; j < NA; ++j)
            if (j < nbound && ((Fm >> j) & 1u) && !(c[j] >= 0.0)) feasible = false;
        if (!feasible) continue;
        double res[NB], scale[NB];
        double obj = 0.0;
#pragma unroll
        for (int i = 0; i < NB; ++i) {
            double s = -y[i], sa = fabs(y[i]);
#pragma unroll
            for (int j = 0; j < NB; ++j) {
                s = fma(M[i][j], c[j], s);
                sa += fabs(M[i][j] * c[j]);
            }
            res[i] = s;
            scale[i] = sa;
            obj = fma(s, s, obj);
        }
        bool kkt = true;
#pragma unroll
        for (int j = 0; j < NA; ++j) {
            if (j >= nbound || !((S0 >> j) & 1u) || !((live >> j) & 1u)) continue;
            double g = 0.0, gs = 0.0;
#pragma unroll
            for (int i = 0; i < NB; ++i) {
                g = fma(M[i][j], res[i], g);
                gs += fabs(M[i][j]) * scale[i];
            }
            if (g < -1e-12 * gs) kkt = false;
        }
        if (kkt) {
#pragma unroll
            for (int j = 0; j < NB; ++j) best[j] = c[j];
            done = true;
        } else if (best_obj < 0.0 || obj < best_obj) {
            best_obj = obj;
#pragma unroll
            for (int j = 0; j < NB; ++j) best[j] = c[j];
        }
    }
    if (!done && best_obj < 0.0) {
#pragma unroll
        for (int j = 0; j < NB; ++j) best[j] = 0.0;
    }
    // chi2 = f'Wf - 2 c'y + c'(M - prior) c
    double quad = 0.0, lin = 0.0;
#pragma unroll
    for (int i = 0; i < NB; ++i) {
        double s = 0.0;
#pragma unroll
        for (int j = 0; j < NB; ++j) {
            const double m0 = a.prior ? M[i][j] - a.prior[i * NB + j] : M[i][j];
            s = fma(m0, best[j], s);
        }
        quad = fma(best[i], s, quad);
        lin = fma(best[i], y[i], lin);
    }
    a.chi2[prob] = a.fwf[t] - 2.0 * lin + quad;
#pragma unroll
    for (int j = 0; j < NB; ++j) a.coeff[prob * NB + j] = best[j];
}

// ---------------------------------------------------------------- dispatch
template <int NA, int NL>
int launch_fit(const ArchArgs &a, hipStream_t s) {
    dim3 g1((a.nprob + 63) / 64, (a.npair + 3) / 4);
    if (a.geom) {
        const long long n = (long long)a.npair * (a.npix_pad + kGeomPad);
        hipLaunchKernelGGL(arch_geom_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, a);
        int rc = rr::check_launch("rr_arch_fit(geometry)");
        if (rc) return rc;
        hipLaunchKernelGGL((arch_accum_kernel<NA, NL, true>), g1, dim3(256), 0, s, a);
    } else {
        hipLaunchKernelGGL((arch_accum_kernel<NA, NL, false>), g1, dim3(256), 0, s, a);
    }
    int rc = rr::check_launch("rr_arch_fit(accumulate)");
    if (rc) return rc;
    const unsigned nblk = (unsigned)((a.nproblems + 63) / 64);
    hipLaunchKernelGGL((arch_solve_kernel<NA, NL>), dim3(nblk), dim3(64), 0, s, a);
    return rr::check_launch("rr_arch_fit(solve)");
}

template <int NA>
int launch_na(const ArchArgs &a, int nleg, hipStream_t s) {
    switch (nleg) {
        case 0: return launch_fit<NA, 0>(a, s);
        case 1: return launch_fit<NA, 1>(a, s);
        case 2: return launch_fit<NA, 2>(a, s);
        case 3: return launch_fit<NA, 3>(a, s);
        case 4: return launch_fit<NA, 4>(a, s);
        case 5: return launch_fit<NA, 5>(a, s);
        case 6: return launch_fit<NA, 6>(a, s);
        case 7: return launch_fit<NA, 7>(a, s);
        case 8: return launch_fit<NA, 8>(a, s);
        case 9: return launch_fit<NA, 9>(a, s);
        case 10: return launch_fit<NA, 10>(a, s);
        case 11: return launch_fit<NA, 11>(a, s);
        case 12: return launch_fit<NA, 12>(a, s);
        default: return rr::fail(RR_EUNSUPPORTED, "rr_arch_fit: nleg=%d unsupported", nleg);
    }
}

template <int NL>
void launch_prep(const double *rband, const double *w, const double *f, int nt, int npp,
                 const double *leg, double *wrl, double *gll, double *yl, double *fwf, hipStream_t s) {
    hipLaunchKernelGGL((arch_prep_kernel<NL>), dim3(nt), dim3(256), 0, s, rband, w, f, npp, leg, wrl, gll,
                       yl, fwf);
}

long long ne_count(int na, int nleg) { return (long long)na * (na + 1) / 2 + (long long)na * nleg + na; }

// normal-equation part of the workspace, 256-byte aligned; the fused rebin's
// (npair, npix_pad) int2 geometry follows it
long long ne_bytes(const rr_arch_desc *d) {
    return (ne_count(d->na, d->nleg) * (long long)d->npair * d->nprob * 8 + 255) / 256 * 256;
}

}  // namespace

extern "C" int rr_arch_prep(const double *d_rband, const double *d_w, const double *d_f, int ntarget,
                            int npix_pad, const double *d_leg, int nleg, double *d_wrl, double *d_gll,
                            double *d_yl, double *d_fwf, void *stream) {
    if (ntarget < 0 || npix_pad <= 0 || npix_pad % 4 != 0 || nleg < 0 || nleg > kMaxNL)
        return rr::fail(RR_EINVAL, "rr_arch_prep: bad sizes ntarget=%d npix_pad=%d nleg=%d", ntarget,
                        npix_pad, nleg);
    if (ntarget == 0) return RR_OK;
    if (!d_rband || !d_w || !d_f || !d_fwf || (nleg > 0 && (!d_leg || !d_wrl || !d_gll || !d_yl)))
        return rr::fail(RR_EINVAL, "rr_arch_prep: null required pointer");
    hipStream_t s = static_cast<hipStream_t>(stream);
    switch (nleg) {
        case 0: launch_prep<0>(d_rband, d_w, d_f, ntarget, npix_pad, d_leg, d_wrl, d_gll, d_yl, d_fwf, s); break;
        case 1: launch_prep<1>(d_rband, d_w, d_f, ntarget, npix_pad, d_leg, d_wrl, d_gll, d_yl, d_fwf, s); break;
        case 2: launch_prep<2>(d_rband, d_w, d_f, ntarget, npix_pad, d_leg, d_wrl, d_gll, d_yl, d_fwf, s); break;
        case 3: launch_prep<3>(d_rband, d_w, d_f, ntarget, npix_pad, d_leg, d_wrl, d_gll, d_yl, d_fwf, s); break;
        case 4: launch_prep<4>(d_rband, d_w, d_f, ntarget, npix_pad, d_leg, d_wrl, d_gll, d_yl, d_fwf, s); break;
        case 5: launch_prep<5>(d_rband, d_w, d_f, ntarget, npix_pad, d_leg, d_wrl, d_gll, d_yl, d_fwf, s); break;
        case 6: launch_prep<6>(d_rband, d_w, d_f, ntarget, npix_pad, d_leg, d_wrl, d_gll, d_yl, d_fwf, s); break;
        case 7: launch_prep<7>(d_rband, d_w, d_f, ntarget, npix_pad, d_leg, d_wrl, d_gll, d_yl, d_fwf, s); break;
        case 8: launch_prep<8>(d_rband, d_w, d_f, ntarget, npix_pad, d_leg, d_wrl, d_gll, d_yl, d_fwf, s); break;
        case 9: launch_prep<9>(d_rband, d_w, d_f, ntarget, npix_pad, d_leg, d_wrl, d_gll, d_yl, d_fwf, s); break;
        case 10: launch_prep<10>(d_rband, d_w, d_f, ntarget, npix_pad, d_leg, d_wrl, d_gll, d_yl, d_fwf, s); break;
        case 11: launch_prep<11>(d_rband, d_w, d_f, ntarget, npix_pad, d_leg, d_wrl, d_gll, d_yl, d_fwf, s); break;
        default: launch_prep<12>(d_rband, d_w, d_f, ntarget, npix_pad, d_leg, d_wrl, d_gll, d_yl, d_fwf, s); break;
    }
    return rr::check_launch("rr_arch_prep");
}

extern "C" long long rr_arch_workspace_bytes(const rr_arch_desc *d) {
    if (!d || d->npair < 0 || d->nprob < 0 || d->na < 1 || d->na > kMaxNA || d->nleg < 0 || d->nleg > kMaxNL)
        return -1;
    const long long rows = d->d_acache ? 0 : (long long)d->npair * (d->npix_pad + kGeomPad);
    return ne_bytes(d) + rows * 8 + rows * kNW * 8;
}

extern "C" int rr_arch_fit(const rr_arch_desc *d, void *stream) {
    if (!d) return rr::fail(RR_EINVAL, "rr_arch_fit: null descriptor");
    if (d->npair < 0 || d->nprob < 0 || d->na < 1 || d->na > kMaxNA || d->nleg < 0 || d->nleg > kMaxNL ||
        d->nbound < 0 || d->nbound > d->na || d->npix_pad <= 0 || d->npix_pad % 4 != 0 || d->halo < 5 ||
        d->narch < 1 || d->ld < d->narch || d->nlines < 0 || d->nlines > RR_MAX_LYMAN_LINES)
        return rr::fail(RR_EINVAL,
                        "rr_arch_fit: bad sizes npair=%d nprob=%d na=%d nleg=%d nbound=%d npix_pad=%d "
                        "halo=%d narch=%d ld=%lld nlines=%d",
                        d->npair, d->nprob, d->na, d->nleg, d->nbound, d->npix_pad, d->halo, d->narch, d->ld,
                        d->nlines);
    if (d->npair == 0 || d->nprob == 0) return RR_OK;
    if (!d->d_columns && (d->na != 1 || d->nprob > d->narch))
        return rr::fail(RR_EINVAL, "rr_arch_fit: without d_columns na must be 1 and nprob <= narch");
    const bool fused = d->d_acache == nullptr;
    if (fused && (!d->d_x || d->nx < 2 || !d->d_yt || d->ldy < d->narch || !d->d_pe_lo || !d->d_pe_hi ||
                  d->npix < 1 || d->npix > d->npix_pad || !d->d_zrebin))
        return rr::fail(RR_EINVAL, "rr_arch_fit: fused rebin needs d_x (nx >= 2), d_yt (ldy >= narch), "
                                   "pixel edges, 0 < npix <= npix_pad and d_zrebin");
    if (!d->d_pair_target || !d->d_rband || !d->d_w || !d->d_wf || !d->d_fwf ||
        !d->d_coeff || !d->d_chi2 || !d->d_work ||
        (d->nleg > 0 && (!d->d_wrl || !d->d_gll || !d->d_yl)) ||
        (d->nlines > 0 && (!d->d_wave || !d->d_zigm || !d->lines)))
        return rr::fail(RR_EINVAL, "rr_arch_fit: null required pointer");
    if (d->work_bytes < rr_arch_workspace_bytes(d))
        return rr::fail(RR_EINVAL, "rr_arch_fit: workspace of %lld bytes, %lld needed", d->work_bytes,
                        rr_arch_workspace_bytes(d));
    ArchArgs a;
    a.npair = d->npair;
    a.nprob = d->nprob;
    a.nbound = d->nbound;
    a.npix_pad = d->npix_pad;
    a.halo = d->halo;
    a.narch = d->narch;
    a.ld = d->ld;
    a.nproblems = (long long)d->npair * d->nprob;
    a.pair_target = d->d_pair_target;
    a.acache = d->d_acache;
    a.columns = d->d_columns;
    a.rband = d->d_rband;
    a.w = d->d_w;
    a.wf = d->d_wf;
    a.wrl = d->d_wrl;
    a.gll = d->d_gll;
    a.yl = d->d_yl;
    a.fwf = d->d_fwf;
    a.prior = d->d_prior;
    a.wave = d->d_wave;
    a.zigm = d->d_zigm;
    a.nlines = d->nlines;
    for (int l = 0; l < d->nlines; ++l) a.lines[l] = d->lines[l];
    a.coeff = d->d_coeff;
    a.chi2 = d->d_chi2;
    a.ne = static_cast<double *>(d->d_work);
    a.x = d->d_x;
    a.yt = d->d_yt;
    a.ldy = d->ldy;
    a.nx = d->nx;
    a.pe_lo = d->d_pe_lo;
    a.pe_hi = d->d_pe_hi;
    a.npix = fused ? d->npix : d->npix_pad;
    a.zrebin = d->d_zrebin;
    a.geom = fused ? reinterpret_cast<int2 *>(static_cast<char *>(d->d_work) + ne_bytes(d)) : nullptr;
    a.wts = fused ? reinterpret_cast<double *>(static_cast<char *>(d->d_work) + ne_bytes(d) +
                                               (long long)d->npair * (d->npix_pad + kGeomPad) * 8)
                  : nullptr;
    hipStream_t s = static_cast<hipStream_t>(stream);
    switch (d->na) {
        case 1: return launch_na<1>(a, d->nleg, s);
        case 2: return launch_na<2>(a, d->nleg, s);
        case 3: return launch_na<3>(a, d->nleg, s);
        default: return launch_na<4>(a, d->nleg, s);
    }
}
