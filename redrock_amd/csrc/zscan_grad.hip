// Adjoint of the chi2 scan (rr_zscan_grad): d chi2 / d z and d chi2 / d(template
// flux) at given (target, z, coefficients), FP64, recomputing the rebin.
//
// chi2(z) = min_c |W^1/2 (f - R T(z) c)|^2.  With c at the minimum (envelope
// theorem) and m = T(z) c the unconvolved model,
//     r = f - R m,   g = R^T (W r),
//     d chi2/dz        = -2 g . dm/dz
//     d chi2/d y_b[j]  = -2 c_b sum_q g_q t_q K_qj
// where K is the rebin operator (bin means of the piecewise-linear template),
// t_q the IGM transmission of pixel q (piecewise constant in z), and, with
// s = 1 + z and bin edges e_lo, e_hi,
//     dT_q/dz = (T_q - (e_hi y(e_hi/s) - e_lo y(e_lo/s)) / (e_hi - e_lo)) / s,
// y(u) the template interpolated at rest wavelength u (continuous in z).
// Since T is linear in the template, m is the rebin of the single spectrum
// sum_b c_b y_b, evaluated on the fly: nothing from the forward pass is
// stored except the coefficients.
//
// One workgroup per item walks the target's pixel axis in chunks of 256
// pixels: (1) m and dm/dz on the chunk plus a 10-pixel halo, (2) residual and
// W r on the chunk plus 5, (3) g = R^T W r, the dz and chi2 sums and, if
// requested, v_q = -2 g_q t_q (item scale) per pixel into the workspace.
// The template-flux gradient is then assembled without atomics, in a fixed
// order (bit-identical from run to run):
//   flux_rows_kernel: per item, every template sample k of the item's span
//     sums v_q K_qk over the pixels q whose bin reaches it, in pixel order
//     (one thread writes each sample), into the item's row P_i;
//   flux_partial_kernel, flux_sum_kernel: dchi2_dy[b][k] += sum_i c_ib P_i[k],
//     items in 64 chunks summed in item order, the chunk sums in chunk order.
#include "common.hpp"
#include "zscan_args.hpp"

#pragma clang fp contract(off)

namespace {

constexpr int CH = 256;     // output pixels per chunk
constexpr int HB = 5;       // band half width

struct GradArgs {
    int nitems, npix, nb, nx, nlines;
    const double *x, *y, *pe_lo, *pe_hi, *pwave;
    const double *rband, *w, *f;
    int npix_pad;
    const int *item_target;
    const double *item_z, *item_coeff, *item_scale;
    double *dchi2_dz, *chi2, *dchi2_dy;
    double *v;        // workspace: (nitems, npix) scatter coefficients v_q
    double *P;        // workspace: (nitems, nx) per-item template-flux rows
    rr_lyman_line lines[RR_MAX_LYMAN_LINES];
};

// Weight of template sample ks in the bin mean of [elo, ehi] at 1+z = s:
// the coefficient bin_value's linear map gives y[ks] (the same expressions),
// 0 when the bin does not reach it.  The caller passes only bins with
// elo < x[min(ks + 1, nx - 1)] s, so the first sample beyond the lower edge
// (rebin.py:142) is at most that index and is found walking down from it
// (a bin holds a few samples) — the index a binary search returns.
__device__ double bin_weight(const GradArgs &a, double s, double elo, double ehi, int ks) {
    const double *x = a.x;
    const int nx = a.nx;
    int lo = ks + 1 < nx - 1 ? ks + 1 : nx - 1;
    while (lo > 0 && x[lo - 1] * s > elo) --lo;
    const int j = lo < 1 ? 1 : (lo > nx - 1 ? nx - 1 : lo);
    const double xj = x[j] * s, xjm1 = x[j - 1] * s;
    const double width = ehi - elo;
    double wsum = 0.0;
    if (xj < ehi) {
        int jend = j;
        while (jend + 1 < nx - 1 && x[jend + 1] * s < ehi) ++jend;
        if (ks < j - 1 || ks > jend + 1) return 0.0;
        const double t = (elo - xjm1) / (xj - xjm1);
        if (ks == j - 1) wsum += 0.5 * (xj - elo) * (1.0 - t);
        if (ks == j) wsum += 0.5 * (xj - elo) * (1.0 + t);
        double xprev = xj;
        for (int jj = j + 1; jj <= jend; ++jj) {
            const double xc = x[jj] * s;
            if (ks == jj - 1) wsum += 0.5 * (xc - xprev);
            if (ks == jj) wsum += 0.5 * (xc - xprev);
            xprev = xc;
        }
        const double xe1 = x[jend + 1] * s;
        const double t2 = (ehi - xprev) / (xe1 - xprev);
        if (ks == jend) wsum += 0.5 * (ehi - xprev) * (2.0 - t2);
        if (ks == jend + 1) wsum += 0.5 * (ehi - xprev) * t2;
    } else {
        if (ks != j - 1 && ks != j) return 0.0;
        const double tm = (0.5 * (elo + ehi) - xjm1) / (xj - xjm1);
        wsum = ks == j - 1 ? width * (1.0 - tm) : width * tm;
    }
    return wsum / width;
}

__device__ __forceinline__ double ycomb(const GradArgs &a, const double *c, int j) {
    double v = 0.0;
    for (int b = 0; b < a.nb; ++b) v += c[b] * a.y[(long long)b * a.nx + j];
    return v;
}

// Rebin of ycomb over [elo, ehi] at 1+z = s (same expressions as rr_rebin),
// also returning ycomb interpolated at both edges.  If `scatter` is set the
// sample weights times `coef` are added to dchi2_dy (weights of the linear map
// y -> bin mean), for every basis b with factor c_b.
__device__ void bin_value(const GradArgs &a, const double *c, double s, double elo, double ehi,
                          double &mean, double &ylo_e, double &yhi_e, bool scatter, double coef) {
    const double *x = a.x;
    const int nx = a.nx;
    int lo = 0, hi = nx;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (x[mid] * s <= elo) lo = mid + 1;
        else hi = mid;
    }
    int j = lo < 1 ? 1 : (lo > nx - 1 ? nx - 1 : lo);
    const double xj = x[j] * s, xjm1 = x[j - 1] * s;
    const double width = ehi - elo;
    double r = 0.0;
    // weights w_k of samples k contributing to the bin (before / width): the
    // linear map y -> bin mean whose transpose bin_weight() evaluates
    auto add = [&](int k, double wk) { (void)k; (void)wk; (void)scatter; (void)coef; };
    if (xj < ehi) {
        int jend = j;
        while (jend + 1 < nx - 1 && x[jend + 1] * s < ehi) ++jend;
        const double yj = ycomb(a, c, j), yjm1 = ycomb(a, c, j - 1);
        const double t = (elo - xjm1) / (xj - xjm1);
        const double yedge = yjm1 + (elo - xjm1) * (yj - yjm1) / (xj - xjm1);
        ylo_e = yedge;
        r += 0.5 * (yj + yedge) * (xj - elo);
        add(j - 1, 0.5 * (xj - elo) * (1.0 - t));
        add(j, 0.5 * (xj - elo) * (1.0 + t));
        double xprev = xj, yprev = yj;
        for (int jj = j + 1; jj <= jend; ++jj) {
            const double xc = x[jj] * s;
            const double yc = ycomb(a, c, jj);
            r += 0.5 * (yc + yprev) * (xc - xprev);
            add(jj - 1, 0.5 * (xc - xprev));
            add(jj, 0.5 * (xc - xprev));
            xprev = xc;
            yprev = yc;
        }
        const double xe1 = x[jend + 1] * s;
        const double ye1 = ycomb(a, c, jend + 1);
        const double t2 = (ehi - xprev) / (xe1 - xprev);
        const double yedge2 = yprev + (ehi - xprev) * (ye1 - yprev) / (xe1 - xprev);
        yhi_e = yedge2;
        r += 0.5 * (yedge2 + yprev) * (ehi - xprev);
        add(jend, 0.5 * (ehi - xprev) * (2.0 - t2));
        add(jend + 1, 0.5 * (ehi - xprev) * t2);
    } else {
        const double yj = ycomb(a, c, j), yjm1 = ycomb(a, c, j - 1);
        const double slope = (yj - yjm1) / (xj - xjm1);
        const double ylo = yj + (elo - xj) * slope;
        const double yhi = yj + (ehi - xj) * slope;
        ylo_e = ylo;
        yhi_e = yhi;
        r += 0.5 * (ylo + yhi) * width;
        // mean = y_j + slope*((elo+ehi)/2 - x_j): weights on y_j, y_{j-1}
        const double tm = (0.5 * (elo + ehi) - xjm1) / (xj - xjm1);
        add(j - 1, width * (1.0 - tm));
        add(j, width * tm);
    }
    mean = r / width;
}

__device__ __forceinline__ double igm_trans(const GradArgs &a, double s, double lobs) {
    double T = 1.0;
    const double lrest = (1.0 / s) * lobs;
    for (int l = 0; l < a.nlines; ++l) {
        const double line = a.lines[l].line;
        if (lrest < line) {
            const double zpix = lobs / line - 1.;
            T *= exp(-a.lines[l].A * pow(1. + zpix, a.lines[l].B));
        }
    }
    return T;
}

__device__ __forceinline__ double block_sum(double v, double *red) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    double t = 0.0;
    for (int i = 0; i < (int)(blockDim.x >> 6); ++i) t += red[i];
    return t;
}

__global__ __launch_bounds__(256) void grad_kernel(GradArgs a) {
    __shared__ double mb[CH + 4 * HB];
    __shared__ double dmb[CH + 4 * HB];
    __shared__ double ub[CH + 2 * HB];
    __shared__ double rb[CH + 2 * HB][11];
    __shared__ double cc[RR_MAX_NBASIS];
    __shared__ double red[8];
    const int item = blockIdx.x;
    if (item >= a.nitems) return;
    const int tid = threadIdx.x;
    const int t = a.item_target[item];
    const double z = a.item_z[item];
    const double s = 1. + z;
    if (tid < a.nb) cc[tid] = a.item_coeff[(long long)item * a.nb + tid];
    __syncthreads();
    const long long npp = a.npix_pad;
    const double *R = a.rband + (long long)t * npp * 11;
    const double *W = a.w + (long long)t * npp;
    const double *F = a.f + (long long)t * npp;
    const bool want_dy = a.dchi2_dy != nullptr;
    const double iscale = a.item_scale ? a.item_scale[item] : 1.0;

    double dz = 0.0, chi2 = 0.0;
    for (int p0 = 0; p0 < a.npix; p0 += CH) {
        // (1) m, dm/dz on pixels [p0 - 10, p0 + CH + 10)
        for (int i = tid; i < CH + 4 * HB; i += blockDim.x) {
            const int q = p0 - 2 * HB + i;
            double m = 0.0, dm = 0.0;
            if (q >= 0 && q < a.npix) {
                const double elo = a.pe_lo[q], ehi = a.pe_hi[q];
                double mean, ye_lo, ye_hi;
                bin_value(a, cc, s, elo, ehi, mean, ye_lo, ye_hi, false, 0.0);
                const double tr = a.nlines ? igm_trans(a, s, a.pwave[q]) : 1.0;
                m = mean * tr;
                dm = tr * (mean - (ehi * ye_hi - elo * ye_lo) / (ehi - elo)) / s;
            }
            mb[i] = m;
            dmb[i] = dm;
        }
        // R rows of pixels [p0 - 5, p0 + CH + 5)
        for (int i = tid; i < (CH + 2 * HB) * 11; i += blockDim.x) {
            const int r = i / 11, k = i % 11;
            const int p = p0 - HB + r;
            rb[r][k] = (p >= 0 && p < a.npix) ? R[(long long)p * 11 + k] : 0.0;
        }
        __syncthreads();
        // (2) residual and W r on pixels [p0 - 5, p0 + CH + 5)
        for (int i = tid; i < CH + 2 * HB; i += blockDim.x) {
            const int p = p0 - HB + i;
            double u = 0.0;
            if (p >= 0 && p < a.npix) {
                double am = 0.0;
#pragma unroll
                for (int k = 0; k < 11; ++k) am += rb[i][k] * mb[i + k];
                const double r = F[p] - am;
                u = W[p] * r;
                if (i >= HB && i < CH + HB) chi2 += u * r;
            }
            ub[i] = u;
        }
        __syncthreads();
        // (3) g = R^T W r on the chunk; dz; template-flux scatter
        {
            const int q = p0 + tid;
            if (tid < CH && q < a.npix) {
                double gq = 0.0;
                // row p = q - k + 5 has column q at band index k
#pragma unroll
                for (int k = 0; k < 11; ++k) {
                    const int ip = tid + HB - k + HB;   // index of p = q - k + 5 in rb/ub
                    gq += rb[ip][k] * ub[ip];
                }
                dz += gq * dmb[tid + 2 * HB];
                if (want_dy) {
                    const double tr = a.nlines ? igm_trans(a, s, a.pwave[q]) : 1.0;
                    a.v[(long long)item * a.npix + q] = -2.0 * gq * tr * iscale;
                }
            }
        }
        __syncthreads();
    }
    dz = block_sum(dz, red);
    chi2 = block_sum(chi2, red);
    if (tid == 0) {
        a.dchi2_dz[item] = -2.0 * dz;
        if (a.chi2) a.chi2[item] = chi2;
    }
}

// One workgroup per item: P_i[k] = sum_q v_q K_qk for every template sample
// k, in pixel order, written once (zero where no pixel reaches k).  The pixel
// axis is the cameras (and exposures) one after another, each with increasing
// bin edges: within a segment the pixels reaching k are those with
// ehi_q > x_{k-1} s and elo_q < x_{k+1} s, found by binary search.
constexpr int kMaxSeg = 64;

__global__ __launch_bounds__(256) void flux_rows_kernel(GradArgs a) {
    __shared__ int seg[kMaxSeg + 1];
    __shared__ int nseg;
    const int item = blockIdx.x;
    const double s = 1. + a.item_z[item];
    const double *v = a.v + (long long)item * a.npix;
    double *P = a.P + (long long)item * a.nx;
    if (threadIdx.x == 0) {                  // segment starts: where the lower edge falls back
        int n = 0;
        seg[n++] = 0;
        for (int q = 1; q < a.npix && n < kMaxSeg; ++q)
            if (a.pe_lo[q] < a.pe_lo[q - 1]) seg[n++] = q;
        seg[n] = a.npix;
        nseg = n;
    }
    __syncthreads();
    for (int k = threadIdx.x; k < a.nx; k += blockDim.x) {
        const double xl = a.x[k > 0 ? k - 1 : 0] * s, xr = a.x[k + 1 < a.nx ? k + 1 : a.nx - 1] * s;
        double acc = 0.0;
        for (int g = 0; g < nseg; ++g) {
            const int q0 = seg[g], q1 = seg[g + 1];
            if (a.pe_hi[q1 - 1] <= xl || a.pe_lo[q0] >= xr) continue;
            int lo = q0, hi = q1;            // first q of the segment with ehi_q > xl
            while (lo < hi) {
                const int mid = (lo + hi) >> 1;
                if (a.pe_hi[mid] > xl) hi = mid;
                else lo = mid + 1;
            }
            for (int q = lo; q < q1 && a.pe_lo[q] < xr; ++q)
                acc += v[q] * bin_weight(a, s, a.pe_lo[q], a.pe_hi[q], k);
        }
        P[k] = acc;
    }
}

// dchi2_dy[b][k] += sum_i c_ib P_i[k] in two passes with a fixed order: the
// items in kRedChunks consecutive chunks, each summed in item order by its own
// workgroups (one thread per sample), then the chunks' sums added in chunk
// order.  (One pass over all items had ~80 workgroups for the chip, each
// thread a chain of nitems dependent loads.)
constexpr int kRedChunks = 64;

__host__ __device__ inline int red_chunks(int nitems) {
    return nitems < kRedChunks ? (nitems > 0 ? nitems : 1) : kRedChunks;
}

__global__ __launch_bounds__(256) void flux_partial_kernel(GradArgs a, double *part) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int S = red_chunks(a.nitems), sc = blockIdx.y;
    if (k >= a.nx) return;
    const int i0 = (int)((long long)sc * a.nitems / S), i1 = (int)((long long)(sc + 1) * a.nitems / S);
    double acc[RR_MAX_NBASIS];
    for (int b = 0; b < a.nb; ++b) acc[b] = 0.0;
    for (int i = i0; i < i1; ++i) {
        const double p = a.P[(long long)i * a.nx + k];
        const double *c = a.item_coeff + (long long)i * a.nb;
        for (int b = 0; b < a.nb; ++b) acc[b] = fma(c[b], p, acc[b]);
    }
    for (int b = 0; b < a.nb; ++b) part[((long long)sc * a.nb + b) * a.nx + k] = acc[b];
}

__global__ __launch_bounds__(256) void flux_sum_kernel(GradArgs a, const double *part) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int b = blockIdx.y;
    if (k >= a.nx) return;
    const int S = red_chunks(a.nitems);
    double acc = 0.0;
    for (int sc = 0; sc < S; ++sc) acc += part[((long long)sc * a.nb + b) * a.nx + k];
    a.dchi2_dy[(long long)b * a.nx + k] += acc;
}

}  // namespace

extern "C" long long rr_grad_workspace_bytes(const rr_grad_desc *d) {
    if (!d || d->nitems < 0 || d->npix < 1 || d->nx < 2)
        return rr::fail(RR_EINVAL, "rr_grad_workspace_bytes: bad descriptor");
    if (!d->d_dchi2_dy) return 0;
    return rr::align256((long long)d->nitems * d->npix * 8) + rr::align256((long long)d->nitems * d->nx * 8) +
           rr::align256((long long)red_chunks(d->nitems) * d->nb * d->nx * 8);
}

extern "C" int rr_zscan_grad(const rr_grad_desc *d, void *stream) {
    if (!d) return rr::fail(RR_EINVAL, "rr_zscan_grad: null descriptor");
    if (d->nitems < 0 || d->npix < 1 || d->npix_pad < d->npix || d->nb < 1 ||
        d->nb > RR_MAX_NBASIS || d->nx < 2 || d->nlines < 0 || d->nlines > RR_MAX_LYMAN_LINES)
        return rr::fail(RR_EINVAL, "rr_zscan_grad: bad sizes");
    if (!d->d_x || !d->d_y || !d->d_pe_lo || !d->d_pe_hi || !d->d_rband || !d->d_w || !d->d_f ||
        !d->d_item_target || !d->d_item_z || !d->d_item_coeff || !d->d_dchi2_dz ||
        (d->nlines > 0 && (!d->d_pwave || !d->lines)))
        return rr::fail(RR_EINVAL, "rr_zscan_grad: null required pointer");
    if (d->nitems == 0) return RR_OK;
    GradArgs a;
    a.nitems = d->nitems; a.npix = d->npix; a.nb = d->nb; a.nx = d->nx; a.nlines = d->nlines;
    a.x = d->d_x; a.y = d->d_y; a.pe_lo = d->d_pe_lo; a.pe_hi = d->d_pe_hi; a.pwave = d->d_pwave;
    a.rband = d->d_rband; a.w = d->d_w; a.f = d->d_f; a.npix_pad = d->npix_pad;
    a.item_target = d->d_item_target; a.item_z = d->d_item_z; a.item_coeff = d->d_item_coeff;
    a.item_scale = d->d_item_scale;
    a.dchi2_dz = d->d_dchi2_dz; a.chi2 = d->d_chi2; a.dchi2_dy = d->d_dchi2_dy;
    a.v = nullptr;
    a.P = nullptr;
    if (d->d_dchi2_dy) {
        const long long need = rr_grad_workspace_bytes(d);
        if (!d->d_work || d->work_bytes < need)
            return rr::fail(RR_EINVAL, "rr_zscan_grad: the template-flux gradient needs a workspace of %lld "
                                       "bytes, %lld given", need, d->work_bytes);
        a.v = static_cast<double *>(d->d_work);
        a.P = reinterpret_cast<double *>(static_cast<char *>(d->d_work) +
                                         rr::align256((long long)d->nitems * d->npix * 8));
    }
    for (int l = 0; l < d->nlines; ++l) a.lines[l] = d->lines[l];
    hipStream_t st = static_cast<hipStream_t>(stream);
    hipLaunchKernelGGL(grad_kernel, dim3(d->nitems), dim3(256), 0, st, a);
    int rc = rr::check_launch("rr_zscan_grad");
    if (rc || !d->d_dchi2_dy) return rc;
    // the flux rows take at most kMaxSeg camera / exposure segments
    hipLaunchKernelGGL(flux_rows_kernel, dim3(d->nitems), dim3(256), 0, st, a);
    rc = rr::check_launch("rr_zscan_grad(flux rows)");
    if (rc) return rc;
    double *part = reinterpret_cast<double *>(
        static_cast<char *>(d->d_work) + rr::align256((long long)d->nitems * d->npix * 8) +
        rr::align256((long long)d->nitems * d->nx * 8));
    hipLaunchKernelGGL(flux_partial_kernel, dim3((d->nx + 255) / 256, red_chunks(d->nitems)), dim3(256), 0,
                       st, a, part);
    rc = rr::check_launch("rr_zscan_grad(flux partial sums)");
    if (rc) return rc;
    hipLaunchKernelGGL(flux_sum_kernel, dim3((d->nx + 255) / 256, d->nb), dim3(256), 0, st, a,
                       static_cast<const double *>(part));
    return rr::check_launch("rr_zscan_grad(flux sums)");
}
