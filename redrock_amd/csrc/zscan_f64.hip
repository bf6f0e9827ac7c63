// Plain FP64 chi2 scan (rr_zscan with RR_PREC_F64).
//
// Work decomposition: one wavefront = one item (a target scanned over a z
// range) x 64 consecutive redshifts, one lane per redshift (items of at most
// 16 or 4 redshifts: 4 or 16 items per wavefront, see zscan_f64_kernel).  For
// full grids a workgroup's 4 waves are 4 items on the same 64-z block, so
// their template-cache reads hit the same lines in L1/L2.  For short items
// (fine grids, Newton steps, one redshift per candidate) the 4 waves share
// one wavefront's items and take a quarter of the pixels each: a lane's
// dependent chain of loads is a quarter as long and 4x as many waves are in
// flight (these launches hold few items, and one wave per item left most
// SIMDs idle behind a full-length chain).  Per lane and per block of P pixels:
//   A[p][b] = sum_k R[p][k] T[b][p+k-5][z]      (banded R-apply, 11 taps)
//   M[b1][b2] += w_p A[p][b1] A[p][b2],  y[b] += A[p][b] (w f)_p
// With one item per wave R, w and w*f are wave-uniform and come through
// scalar loads; the cache is laid out [b][pixel][z] so every load of a window
// row is one coalesced 512-byte line run (64 / G consecutive redshifts per
// item with G items per wave).  The kernel writes the
// per-redshift normal equations [M | y] to the workspace; the FP64 solve and
// chi2 follow in zscan_finish.hip.  Split workgroups add their quarters
// through LDS in the fixed order m0 + (m1 + (m2 + m3)): deterministic, and
// the same doubles for 4 and 16 items per wave; full-grid rows (and the lane
// kernel, same order as the full-grid form) differ from them by rounding.
//
// Reference path replaced: zscan.py:670-737 (GPU) / :752-780 (CPU).
#include "common.hpp"
#include "zscan_args.hpp"

namespace {

// G items per wavefront, 64 / G redshifts each (G = 1 for full redshift
// grids; 4 or 16 for the short items of fine-grid fits and Newton steps, whose
// redshifts would leave most lanes of a one-item wave idle).  Every lane runs
// the same per-pixel arithmetic in the same order for G = 4 and 16, so their
// results do not depend on G.
// PF: load every basis vector's template window of a pixel block before the
// first is used (launches too small to fill the chip; more registers)
template <int NB, int P, int G, bool PF>
__device__ __forceinline__ void zscan_f64_body(const rr::ZArgs &a, long long ne_off) {
    constexpr int LZ = 64 / G;           // lanes (redshifts) per item
    constexpr bool kSplit = G > 1;       // the 4 waves: one pixel quarter each
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    int item, zl;
    bool live = true;
    if constexpr (G == 1) {
        item = blockIdx.y * 4 + wave;    // wave-uniform: R, w, w*f through scalar loads
        if (item >= a.nitems) return;
        zl = lane;
    } else {
        const int first = blockIdx.y * G;
        if (first >= a.nitems) return;   // uniform over the workgroup
        item = first + lane / LZ;
        zl = lane % LZ;
        if (item >= a.nitems) {          // idle lanes repeat the last item, write nothing
            item = a.nitems - 1;
            live = false;
        }
    }
    const int t = a.item_target[item];
    const int z0 = a.item_z0[item];
    const int nzi = a.item_nz[item];
    const int zbase = blockIdx.x * LZ;
    if constexpr (G == 1) {
        if (zbase >= nzi) return;
    }
    const int iz = zbase + zl;
    live = live && iz < nzi;
    const int zc = z0 + (iz < nzi ? iz : nzi - 1);

    const long long npp = a.npix_pad;
    const double *__restrict__ Rt = a.rband + (long long)t * npp * 11;
    const double *__restrict__ Wt = a.w + (long long)t * npp;
    const double *__restrict__ WFt = a.wf + (long long)t * npp;
    const double *__restrict__ T = static_cast<const double *>(a.tcache);
    const long long ldz = a.cache_ldz;
    const long long bstride = a.cache_rows * ldz;

    constexpr int NM = NB * (NB + 1) / 2;
    double m[NM];
    double y[NB];
#pragma unroll
    for (int i = 0; i < NM; ++i) m[i] = 0.0;
#pragma unroll
    for (int i = 0; i < NB; ++i) y[i] = 0.0;

    // row of pixel p in the cache = p + halo; window of block p0 starts at p0-5
    const double *Tz = T + (long long)(a.halo - 5) * ldz + zc;

    int pbeg = 0, pend = a.npix_pad;
    if constexpr (kSplit) {
        const int nblk = a.npix_pad / P;
        pbeg = (wave * nblk / 4) * P;
        pend = ((wave + 1) * nblk / 4) * P;
    }
    for (int p0 = pbeg; p0 < pend; p0 += P) {
        double r[P][11];
#pragma unroll
        for (int i = 0; i < P; ++i)
#pragma unroll
            for (int k = 0; k < 11; ++k) r[i][k] = Rt[(long long)(p0 + i) * 11 + k];
        double A[P][NB];
        if constexpr (PF) {
            // every basis vector's window loaded before the first is used
            double win[NB][P + 10];
#pragma unroll
            for (int b = 0; b < NB; ++b) {
                const double *col = Tz + b * bstride + (long long)p0 * ldz;
#pragma unroll
                for (int i = 0; i < P + 10; ++i) win[b][i] = col[i * ldz];
            }
#pragma unroll
            for (int b = 0; b < NB; ++b) {
#pragma unroll
                for (int i = 0; i < P; ++i) {
                    double s = r[i][0] * win[b][i];
#pragma unroll
                    for (int k = 1; k < 11; ++k) s = fma(r[i][k], win[b][i + k], s);
                    A[i][b] = s;
                }
            }
        } else {
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
        }
#pragma unroll
        for (int i = 0; i < P; ++i) {
            const double w = Wt[p0 + i];
            const double wf = WFt[p0 + i];
            int idx = 0;
#pragma unroll
            for (int b1 = 0; b1 < NB; ++b1) {
                const double wa = w * A[i][b1];
                y[b1] = fma(A[i][b1], wf, y[b1]);
#pragma unroll
                for (int b2 = b1; b2 < NB; ++b2) {
                    m[idx] = fma(wa, A[i][b2], m[idx]);
                    ++idx;
                }
            }
        }
    }

    if constexpr (kSplit) {
        // quarters summed in the fixed order m0 + (m1 + (m2 + m3))
        __shared__ double red[NM + NB][64];
        for (int s = 3; s >= 1; --s) {
            if (wave == s) {
#pragma unroll
                for (int i = 0; i < NM; ++i) red[i][lane] = s < 3 ? m[i] + red[i][lane] : m[i];
#pragma unroll
                for (int i = 0; i < NB; ++i) red[NM + i][lane] = s < 3 ? y[i] + red[NM + i][lane] : y[i];
            }
            __syncthreads();
        }
        if (wave != 0) return;
#pragma unroll
        for (int i = 0; i < NM; ++i) m[i] += red[i][lane];
#pragma unroll
        for (int i = 0; i < NB; ++i) y[i] += red[NM + i][lane];
    }
    if (!live) return;
    // normal equations of this (item, z) -> workspace, k-major (coalesced)
    const long long row = a.item_out[item] + iz;
    double *ne = reinterpret_cast<double *>(static_cast<char *>(a.work) + ne_off);
    const long long R = a.total_rows;
    int idx = 0;
#pragma unroll
    for (int b1 = 0; b1 < NB; ++b1) {
#pragma unroll
        for (int b2 = b1; b2 < NB; ++b2) {
            ne[rr::ne_index(b1, b2, NB, false) * R + row] = m[idx];
            ++idx;
        }
        ne[rr::ne_index(b1, NB, NB, false) * R + row] = y[b1];
    }
}

template <int NB, int P, int G>
__global__ __launch_bounds__(256) void zscan_f64_kernel(rr::ZArgs a, long long ne_off) {
    zscan_f64_body<NB, P, G, false>(a, ne_off);
}

// The same arithmetic for launches too small to fill the chip (fewer
// workgroups than CUs): one wave per SIMD, every basis vector's window loads
// of a pixel block in flight together instead of one round trip per basis
// vector.  Same operations in the same order: the same doubles.
template <int NB, int P, int G>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(1, 1)))
void zscan_f64_small_kernel(rr::ZArgs a, long long ne_off) {
    zscan_f64_body<NB, P, G, true>(a, ne_off);
}

// launches of fewer workgroups than this take zscan_f64_small_kernel
constexpr long long kSmallLaunch = 256;

template <int NB, int G>
int launch_g(const rr::ZArgs &a, hipStream_t s) {
    constexpr int LZ = 64 / G;
    // full grids: 4 items per workgroup; short items: G items shared by its 4 waves
    const int per_wg = G == 1 ? 4 : G;
    dim3 grid((a.max_item_nz + LZ - 1) / LZ, (a.nitems + per_wg - 1) / per_wg);
    const long long ne_off = rr::align256((long long)a.nitems * 2 * 8);
    if ((long long)grid.x * grid.y < kSmallLaunch)
        hipLaunchKernelGGL((zscan_f64_small_kernel<NB, 2, G>), grid, dim3(256), 0, s, a, ne_off);
    else
        hipLaunchKernelGGL((zscan_f64_kernel<NB, 2, G>), grid, dim3(256), 0, s, a, ne_off);
    return rr::check_launch("rr_zscan(f64)");
}

template <int NB>
int launch_nb(const rr::ZArgs &a, hipStream_t s) {
    if (a.max_item_nz <= 4) return launch_g<NB, 16>(a, s);
    if (a.max_item_nz <= 16) return launch_g<NB, 4>(a, s);
    return launch_g<NB, 1>(a, s);
}

}  // namespace

namespace rr {
int zscan_f64(const ZArgs &a, hipStream_t s) {
    switch (a.nb) {
        case 1: return launch_nb<1>(a, s);
        case 2: return launch_nb<2>(a, s);
        case 3: return launch_nb<3>(a, s);
        case 4: return launch_nb<4>(a, s);
        case 5: return launch_nb<5>(a, s);
        case 6: return launch_nb<6>(a, s);
        case 7: return launch_nb<7>(a, s);
        case 8: return launch_nb<8>(a, s);
        case 9: return launch_nb<9>(a, s);
        case 10: return launch_nb<10>(a, s);
        case 11: return launch_nb<11>(a, s);
        case 12: return launch_nb<12>(a, s);
        default: return fail(RR_EUNSUPPORTED, "rr_zscan(f64): nb=%d unsupported (1..12)", a.nb);
    }
}
}  // namespace rr
