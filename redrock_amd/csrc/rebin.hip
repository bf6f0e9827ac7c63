// Trapezoidal rebin of a template basis to every trial redshift with the
// Lyman-series IGM factor fused in (rr_rebin).
//
// One thread owns one (redshift, output bin) and loops over the basis, so the
// sample search and the IGM factor are computed once per bin.  Threads are
// laid out so that a wavefront's stores are contiguous: 256 bins of one
// redshift per workgroup when bins are the fast output axis (the template
// window those bins read is then staged in LDS, rebin_lds_kernel), 64
// redshifts x 4 runs of 8 consecutive bins per workgroup when the output is
// z-fastest (the FP64 cache of the redshift-per-lane scan; a thread searches
// for its first bin's samples and steps forward for the next ones).  The bin mean
// follows the serial sweep of py/redrock/rebin.py:126-176 expression by
// expression, and this file is compiled with FP contraction off, so the FP64
// result equals the sweep's double for double: the sample index j is the
// first x[j]*(1+z) above the lower edge (the sweep's `while x[j] <= edges[i]`,
// rebin.py:142), found here by binary search instead of a running pointer.
#include "common.hpp"

#pragma clang fp contract(off)

namespace {

struct RebinArgs {
    const double *x;
    const double *y;
    const double *edges;
    const double *wave;
    const double *z;
    void *out;
    long long y_ld, s_z, s_pix, s_b;
    double scale;
    int nx, nb, nbin, nz, nlines, out_type;
    rr_lyman_line lines[RR_MAX_LYMAN_LINES];
};

// first j in [lo, hi) with x[j]*opz > e (or hi when none)
__device__ __forceinline__ int upper_bound_scaled(const double *x, int lo, int hi, double opz, double e) {
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (x[mid] * opz <= e) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

template <int OUT>
__device__ __forceinline__ void store(const RebinArgs &a, long long off, double v) {
    v *= a.scale;
    if (OUT == RR_F64) {
        static_cast<double *>(a.out)[off] = v;
    } else if (OUT == RR_F32) {
        static_cast<float *>(a.out)[off] = (float)v;
    } else if (OUT == RR_F16) {
        static_cast<uint16_t *>(a.out)[off] = rr::f32_to_f16_bits((float)v);
    } else {
        static_cast<uint16_t *>(a.out)[off] = rr::f32_to_bf16_bits((float)v);
    }
}

// IGM factor (igm.py:346-363): product over lines with lRest < line.
__device__ __forceinline__ double igm_factor(const RebinArgs &a, int iz, int ibin, double opz) {
    double T = 1.0;
    if (a.nlines > 0) {
        const double lobs = a.wave[ibin];
        const double lrest = (1.0 / opz) * lobs;
        for (int l = 0; l < a.nlines; ++l) {
            const double line = a.lines[l].line;
            if (lrest < line) {
                const double zpix = lobs / line - 1.;
                const double tau = a.lines[l].A * pow(1. + zpix, a.lines[l].B);
                T *= exp(-tau);
            }
        }
    }
    return T;
}

// One (redshift, bin), template samples read from global memory.  jraw: on
// entry a lower bound of the first sample beyond the lower edge (0, or that
// index for an earlier bin of the same redshift: the edges increase), on
// exit that index for this bin.  From 0 it is found by binary search, else
// by stepping forward (a bin holds a few samples); both give the same index.
template <int OUT>
__device__ void rebin_bin(const RebinArgs &a, int iz, int ibin, int &jraw) {
    if (ibin >= a.nbin || iz >= a.nz) return;
    const double *x = a.x;
    const double opz = 1. + a.z[iz];
    const double elo = a.edges[ibin];
    const double ehi = a.edges[ibin + 1];
    const int nx = a.nx;

    // rebin.py:142 — first sample beyond the lower edge; clamped so that an
    // out-of-range call stays in bounds (the host checks the range first).
    if (jraw == 0) {
        jraw = upper_bound_scaled(x, 0, nx, opz, elo);
    } else {
        while (jraw < nx && x[jraw] * opz <= elo) ++jraw;
    }
    int j = jraw;
    if (j < 1) j = 1;
    if (j > nx - 1) j = nx - 1;
    const double xj = x[j] * opz;
    const double xjm1 = x[j - 1] * opz;
    // j == nx - 1 cannot be inside after the host's range check; the test
    // keeps x[jend + 1] in bounds for an unchecked call
    const bool inside = xj < ehi && j < nx - 1;
    // rebin.py:154 — last interior sample of this bin
    int jend = j;
    if (inside) {
        while (jend + 1 < nx - 1 && x[jend + 1] * opz < ehi) ++jend;
    }

    const double T = igm_factor(a, iz, ibin, opz);
    const double width = ehi - elo;
    const long long obase = (long long)iz * a.s_z + (long long)ibin * a.s_pix;
    for (int b = 0; b < a.nb; ++b) {
        const double *y = a.y + (long long)b * a.y_ld;
        double r = 0.0;
        if (inside) {
            const double yedge = y[j - 1] + (elo - xjm1) * (y[j] - y[j - 1]) / (xj - xjm1);
            r += 0.5 * (y[j] + yedge) * (xj - elo);
            double xprev = xj;
            for (int jj = j + 1; jj <= jend; ++jj) {
                const double xc = x[jj] * opz;
                r += 0.5 * (y[jj] + y[jj - 1]) * (xc - xprev);
                xprev = xc;
            }
            const double xe1 = x[jend + 1] * opz;
            const double yedge2 = y[jend] + (ehi - xprev) * (y[jend + 1] - y[jend]) / (xe1 - xprev);
            r += 0.5 * (yedge2 + y[jend]) * (ehi - xprev);
        } else {
            const double ylo = y[j] + (elo - xj) * (y[j] - y[j - 1]) / (xj - xjm1);
            const double yhi = y[j] + (ehi - xj) * (y[j] - y[j - 1]) / (xj - xjm1);
            r += 0.5 * (ylo + yhi) * width;
        }
        r /= width;
        if (a.nlines > 0) r *= T;
        store<OUT>(a, obase + (long long)b * a.s_b, r);
    }
}

// z-fastest outputs: kZBins consecutive bins per thread, one binary search
constexpr int kZBins = 8;

template <int OUT, bool ZFAST>
__global__ __launch_bounds__(256) void rebin_kernel(RebinArgs a, int nfast) {
    int jraw = 0;
    if (ZFAST) {   // lane -> redshift, wavefront -> kZBins bins; nfast = redshift blocks of 64
        const int zb = blockIdx.x % nfast, bb = blockIdx.x / nfast;
        const int iz = zb * 64 + (threadIdx.x & 63);
        const int ibin0 = (bb * 4 + (threadIdx.x >> 6)) * kZBins;
        for (int t = 0; t < kZBins; ++t) rebin_bin<OUT>(a, iz, ibin0 + t, jraw);
    } else {       // thread -> bin; nfast = bin blocks of 256
        const int iz = blockIdx.x / nfast;
        const int ibin = (blockIdx.x - iz * nfast) * blockDim.x + threadIdx.x;
        rebin_bin<OUT>(a, iz, ibin, jraw);
    }
}

// Bins-fast layouts with the template window staged in LDS: a workgroup's
// 256 consecutive bins of one redshift read one contiguous run of template
// samples, [wlo, whi] (the first sample before the lower edge of its first bin
// to the first sample beyond the upper edge of its last), so the workgroup
// copies that run's wavelengths and, when they fit beside them, every basis
// vector's values into LDS with coalesced loads (one wait for all of them);
// each thread's binary search and interior sweep then read LDS.  The
// expressions and their order are those of rebin_kernel (same doubles).  Values
// of a window too long for all bases read global memory; a wavelength window
// longer than half the LDS array (a template sampled far more finely than the
// output bins) takes rebin_kernel's global path for that workgroup.
constexpr int kLdsDoubles = 3072;   // 24 KB: 6 workgroups per CU

// Bin mean of one basis vector: ys[k] is the sample of index wlo + k.
template <typename YP>
__device__ __forceinline__ double bin_mean(YP ys, const double *xs, int k, int kend, bool inside,
                                           double elo, double ehi, double xj, double xjm1, double opz) {
    const double width = ehi - elo;
    double r = 0.0;
    if (inside) {
        const double yedge = ys[k - 1] + (elo - xjm1) * (ys[k] - ys[k - 1]) / (xj - xjm1);
        r += 0.5 * (ys[k] + yedge) * (xj - elo);
        double xprev = xj;
        for (int kk = k + 1; kk <= kend; ++kk) {
            const double xc = xs[kk] * opz;
            r += 0.5 * (ys[kk] + ys[kk - 1]) * (xc - xprev);
            xprev = xc;
        }
        const double xe1 = xs[kend + 1] * opz;
        const double yedge2 = ys[kend] + (ehi - xprev) * (ys[kend + 1] - ys[kend]) / (xe1 - xprev);
        r += 0.5 * (yedge2 + ys[kend]) * (ehi - xprev);
    } else {
        const double ylo = ys[k] + (elo - xj) * (ys[k] - ys[k - 1]) / (xj - xjm1);
        const double yhi = ys[k] + (ehi - xj) * (ys[k] - ys[k - 1]) / (xj - xjm1);
        r += 0.5 * (ylo + yhi) * width;
    }
    return r / width;
}

template <int OUT>
__global__ __launch_bounds__(256) void rebin_lds_kernel(RebinArgs a, int nfast) {
    __shared__ double lds[kLdsDoubles];
    const int iz = blockIdx.x / nfast;
    const int b0 = (blockIdx.x - iz * nfast) * 256;
    const int ibin = b0 + threadIdx.x;
    if (iz >= a.nz) return;                                 // uniform over the workgroup
    const int b1 = b0 + 255 < a.nbin - 1 ? b0 + 255 : a.nbin - 1;
    const double *x = a.x;
    const int nx = a.nx;
    const double opz = 1. + a.z[iz];
    // the window: every sample index the bins [b0, b1] read (same clamps as
    // rebin_kernel; whi >= jA keeps an out-of-range call inside the window)
    int jA = upper_bound_scaled(x, 0, nx, opz, a.edges[b0]);
    if (jA < 1) jA = 1;
    if (jA > nx - 1) jA = nx - 1;
    const int wlo = jA - 1;
    int whi = upper_bound_scaled(x, 0, nx, opz, a.edges[b1 + 1]);
    if (whi > nx - 1) whi = nx - 1;
    if (whi < jA) whi = jA;
    const int W = whi - wlo + 1;
    if (2 * W > kLdsDoubles) {                              // uniform: global path
        int jraw = 0;
        rebin_bin<OUT>(a, iz, ibin, jraw);
        return;
    }
    double *xs = lds;
    double *ys = lds + W;
    const bool ystaged = (long long)W * (1 + a.nb) <= kLdsDoubles;
    for (int i = threadIdx.x; i < W; i += 256) xs[i] = x[wlo + i];
    if (ystaged) {
        for (int b = 0; b < a.nb; ++b) {
            const double *y = a.y + (long long)b * a.y_ld + wlo;
            for (int i = threadIdx.x; i < W; i += 256) ys[b * W + i] = y[i];
        }
    }
    __syncthreads();
    if (ibin >= a.nbin) return;

    const double elo = a.edges[ibin];
    const double ehi = a.edges[ibin + 1];
    // rebin.py:142, searched in the window: the answer lies in (wlo, whi + 1]
    int lo = 1, hi = W;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (xs[mid] * opz <= elo) lo = mid + 1;
        else hi = mid;
    }
    int j = wlo + lo;
    if (j < 1) j = 1;
    if (j > nx - 1) j = nx - 1;
    const double xj = xs[j - wlo] * opz;
    const double xjm1 = xs[j - 1 - wlo] * opz;
    const bool inside = xj < ehi && j < nx - 1;
    int jend = j;
    if (inside) {
        while (jend + 1 < nx - 1 && xs[jend + 1 - wlo] * opz < ehi) ++jend;
    }
    const double T = igm_factor(a, iz, ibin, opz);
    const long long obase = (long long)iz * a.s_z + (long long)ibin * a.s_pix;
    const int k = j - wlo, kend = jend - wlo;
    for (int b = 0; b < a.nb; ++b) {
        double r = ystaged ? bin_mean(ys + b * W, xs, k, kend, inside, elo, ehi, xj, xjm1, opz)
                           : bin_mean(a.y + (long long)b * a.y_ld + wlo, xs, k, kend, inside, elo, ehi,
                                      xj, xjm1, opz);
        if (a.nlines > 0) r *= T;
        store<OUT>(a, obase + (long long)b * a.s_b, r);
    }
}

template <int OUT>
void launch(bool zfast, dim3 grid, dim3 block, hipStream_t s, const RebinArgs &a, int nfast) {
    if (zfast) hipLaunchKernelGGL((rebin_kernel<OUT, true>), grid, block, 0, s, a, nfast);
    else hipLaunchKernelGGL((rebin_lds_kernel<OUT>), grid, block, 0, s, a, nfast);
}

}  // namespace

extern "C" int rr_rebin(const double *d_x, int nx, const double *d_y, int nb, long long y_ld,
                        const double *d_edges, int nbin, const double *d_wave,
                        const double *d_z, int nz, const rr_lyman_line *lines, int nlines,
                        void *d_out, int out_type, long long s_z, long long s_pix, long long s_b,
                        double scale, void *stream) {
    if (nx < 2 || nb < 1 || nbin < 1 || nz < 0 || y_ld < nx)
        return rr::fail(RR_EINVAL, "rr_rebin: bad sizes nx=%d nb=%d nbin=%d nz=%d y_ld=%lld",
                        nx, nb, nbin, nz, y_ld);
    if (nlines < 0 || nlines > RR_MAX_LYMAN_LINES || (nlines > 0 && (!lines || !d_wave)))
        return rr::fail(RR_EINVAL, "rr_rebin: bad IGM table (nlines=%d)", nlines);
    if (!d_x || !d_y || !d_edges || !d_z || !d_out)
        return rr::fail(RR_EINVAL, "rr_rebin: null pointer");
    if (nz == 0) return RR_OK;
    RebinArgs a;
    a.x = d_x; a.y = d_y; a.edges = d_edges; a.wave = d_wave; a.z = d_z; a.out = d_out;
    a.y_ld = y_ld; a.s_z = s_z; a.s_pix = s_pix; a.s_b = s_b; a.scale = scale;
    a.nx = nx; a.nb = nb; a.nbin = nbin; a.nz = nz; a.nlines = nlines; a.out_type = out_type;
    for (int l = 0; l < nlines; ++l) a.lines[l] = lines[l];
    const bool zfast = s_z == 1 && s_pix != 1;
    const int nfast = zfast ? (nz + 63) / 64 : (nbin + 255) / 256;
    const long long nblocks = zfast ? (long long)nfast * ((nbin + 4 * kZBins - 1) / (4 * kZBins))
                                    : (long long)nfast * nz;
    if (nblocks > 0x7fffffffLL) return rr::fail(RR_EINVAL, "rr_rebin: grid too large");
    dim3 block(256), grid((unsigned)nblocks);
    hipStream_t s = static_cast<hipStream_t>(stream);
    switch (out_type) {
        case RR_F64: launch<RR_F64>(zfast, grid, block, s, a, nfast); break;
        case RR_F32: launch<RR_F32>(zfast, grid, block, s, a, nfast); break;
        case RR_F16: launch<RR_F16>(zfast, grid, block, s, a, nfast); break;
        case RR_BF16: launch<RR_BF16>(zfast, grid, block, s, a, nfast); break;
        default: return rr::fail(RR_EINVAL, "rr_rebin: unknown out_type %d", out_type);
    }
    return rr::check_launch("rr_rebin");
}
