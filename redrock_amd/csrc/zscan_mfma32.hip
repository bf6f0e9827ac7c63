// FP32-input matrix-core chi2 scan (rr_zscan with RR_PREC_F32) and its packing.
//
// Same decomposition as the 16-bit kernel (zscan_mfma.hip) on
// v_mfma_f32_16x16x4_f32, so the operands keep 24-bit significands and the
// chi2 error stays at the fp32 level at any S/N:
//
//   D_h = sum_{kc<8} A_h[kc] B_h[kc]     tile h of the pair: 16 output pixels,
//                                        K = 32 input pixels in 8 slices of 4
//   G  += sum_{h,r} X_h[r] X_h[r]^T      Gram over the 32 pixels, 8 MFMAs
//
// Lane (i|j, g) of slice kc holds input pixel 16q-8+8g+kc (the same pixel
// order as the 16-bit A-operand tiles); the D_h registers r = 0..3 of lane
// (j, g) are pixels 4g+r of tile h, which are exactly the 4-wide K slices the
// Gram MFMAs need.  Column nb of the Gram operand carries sqrt(w) f (fp32, no
// hi/lo split); column nb+1 is zero.  Per 32 pixels, target and redshift:
// 24 MFMAs (16 + 8).  LDS staging and tile shapes follow the 16-bit kernel.
#include "common.hpp"
#include "zscan_args.hpp"
#include "xcd_map.hpp"

namespace {

typedef float float4v __attribute__((ext_vector_type(4)));
typedef float float8v __attribute__((ext_vector_type(8)));

__device__ __forceinline__ float4v mfma4(float a, float b, float4v c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    return v;
}

// rtile (nt, ntile, 64, 8) f32: lane (i, g) slot kc -> sqrt(w_p) rscale
// R[p][8g+kc-i-3], p = 16q+i.  sfcol (nt, npair, 4, 8) f32: lane g of column
// nb, slot v -> sqrt(w) f fscale at pixel 32qp + (v<4 ? 4g+v : 16+4g+v-4).
__global__ __launch_bounds__(64) void pack32_kernel(const double *__restrict__ rband,
                                                    const double *__restrict__ w,
                                                    const double *__restrict__ f, int npix_pad,
                                                    float *__restrict__ rtile,
                                                    float *__restrict__ sfcol,
                                                    double *__restrict__ scales) {
    const int t = blockIdx.x;
    const int lane = threadIdx.x;
    const double *W = w + (long long)t * npix_pad;
    const double *F = f + (long long)t * npix_pad;
    const double *Rb = rband + (long long)t * npix_pad * 11;
    double mw = 0.0, mf = 0.0;
    for (int p = lane; p < npix_pad; p += 64) {
        const double sw = sqrt(W[p]);
        mw = fmax(mw, sw);
        mf = fmax(mf, fabs(sw * F[p]));
    }
    mw = wave_max(mw);
    mf = wave_max(mf);
    const double rscale = (mw > 0.0) ? exp2(-ceil(log2(mw))) : 1.0;
    const double fscale = (mf > 0.0) ? exp2(8.0 - ceil(log2(mf))) : 1.0;
    if (lane == 0) {
        scales[2 * t] = rscale;
        scales[2 * t + 1] = fscale;
    }
    const int ntile = npix_pad / 16;
    float *RT = rtile + (long long)t * ntile * 64 * 8;
    const int i = lane & 15, g = lane >> 4;
    for (int q = 0; q < ntile; ++q) {
        const int p = 16 * q + i;
        const double sw = sqrt(W[p]) * rscale;
        float8v v8;
#pragma unroll
        for (int v = 0; v < 8; ++v) {
            const int k = 8 * g + v - i - 3;
            v8[v] = (k >= 0 && k <= 10) ? (float)(Rb[(long long)p * 11 + k] * sw) : 0.f;
        }
        *reinterpret_cast<float8v *>(RT + ((long long)q * 64 + lane) * 8) = v8;
    }
    const int npair = npix_pad / 32;
    float *SF = sfcol + (long long)t * npair * 32;
    for (int e = lane; e < npair * 4; e += 64) {
        const int qp = e >> 2, gg = e & 3;
        float8v s8;
#pragma unroll
        for (int v = 0; v < 8; ++v) {
            const int p = 32 * qp + (v < 4 ? 4 * gg + v : 16 + 4 * gg + (v - 4));
            s8[v] = (float)(sqrt(W[p]) * F[p] * fscale);
        }
        *reinterpret_cast<float8v *>(SF + (long long)e * 8) = s8;
    }
}

template <int NB, int WA, int WB, int TW, int ZW>
__global__ __launch_bounds__(256) void zscan_mfma32_kernel(rr::ZArgs a, long long ne_off, rr::TileMap map) {
    constexpr int ZP = rr::zpack(NB);    // redshifts side by side in the 16 MFMA columns
    constexpr int TT = WA * TW, ZZ = WB * ZW, ZR = ZZ * ZP;
    constexpr int TCH = ZR * NB * 12;    // 16-byte chunks: 48 rows x 4 B per basis per z
    constexpr int RCH = TT * 2 * 128;    // 2 tiles x 64 lanes x 32 B
    constexpr int SCH = TT * 8;          // 4 g x 32 B
    constexpr int NCH = TCH + RCH + SCH;
    constexpr int PER = (NCH + 255) / 256;
    typedef float float4c __attribute__((ext_vector_type(4)));
    __shared__ float4c lds[2][NCH];

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int wa = wave % WA, wb = wave / WA;
    int ti, tz;
    if (!rr::map_tile(map, blockIdx.x, ti, tz)) return;   // uniform over the workgroup
    const int itemb = ti * TT;
    const int z0 = a.item_z0[itemb];
    const int nzi = a.item_nz[itemb];
    const int zb = tz * ZR;
    if (zb >= nzi) return;

    const int ntile = a.npix_pad / 16;
    const int npair = ntile / 2;
    const long long rows = a.cache_rows;
    const float *__restrict__ T = static_cast<const float *>(a.tcache);
    const float *__restrict__ RT = static_cast<const float *>(a.rtile);
    const float *__restrict__ SF = static_cast<const float *>(a.sfcol);

    const float *src[PER];
    int step[PER];
#pragma unroll
    for (int k = 0; k < PER; ++k) {
        int c = tid + 256 * k;
        if (c >= NCH) c = NCH - 1;
        if (c < TCH) {
            const int s = c / (NB * 12), b = (c / 12) % NB, part = c % 12;
            const int iz = zb + s;
            const int zc = z0 + (iz < nzi ? iz : nzi - 1);
            src[k] = T + ((long long)zc * NB + b) * rows + 4 * part;
            step[k] = 32;
        } else if (c < TCH + RCH) {
            const int r = c - TCH, u = r >> 8, h = (r >> 7) & 1, l = (r >> 1) & 63, half = r & 1;
            const int it = itemb + u < a.nitems ? itemb + u : a.nitems - 1;
            const int t = a.item_target[it];
            src[k] = RT + (((long long)t * ntile + h) * 64 + l) * 8 + 4 * half;
            step[k] = 2 * 64 * 8;
        } else {
            const int r = c - TCH - RCH, u = r >> 3, gs = r & 7;
            const int it = itemb + u < a.nitems ? itemb + u : a.nitems - 1;
            const int t = a.item_target[it];
            src[k] = SF + ((long long)t * npair * 4) * 8 + gs * 4;
            step[k] = 4 * 8;
        }
    }
    float4c pre[PER];
    auto load = [&](int qp) {
#pragma unroll
        for (int k = 0; k < PER; ++k)
            pre[k] = *reinterpret_cast<const float4c *>(src[k] + (long long)qp * step[k]);
    };
    auto store = [&](int buf) {
#pragma unroll
        for (int k = 0; k < PER; ++k) {
            const int c = tid + 256 * k;
            if (c < NCH) lds[buf][c] = pre[k];
        }
    };

    // column j < ZP*NB: basis j%NB of redshift j/NB of the group; column
    // ZP*NB: sqrt(w) f; the others stay zero
    const int j = lane & 15;
    const int g = lane >> 4;
    const bool real = j < ZP * NB;
    const bool fcolumn = (j == ZP * NB);
    const int jz = real ? j / NB : 0, jb = real ? j - jz * NB : 0;

    // fp32 MFMA partial sums of one pair (32 pixels) are flushed into FP64:
    // with fp32 operands the accumulation over ~8000 pixels would otherwise
    // be the dominant error (~sqrt(2000) * 2^-24 relative in M and y).
    float4v G[TW][ZW];
    double Gd[TW][ZW][4];
#pragma unroll
    for (int u = 0; u < TW; ++u)
#pragma unroll
        for (int s = 0; s < ZW; ++s)
#pragma unroll
            for (int r = 0; r < 4; ++r) Gd[u][s][r] = 0.0;
    const float4v zero4 = {0.f, 0.f, 0.f, 0.f};

    load(0);
    store(0);
    __syncthreads();

    for (int qp = 0; qp < npair; ++qp) {
        const int buf = qp & 1;
        const bool more = qp + 1 < npair;
        if (more) load(qp + 1);
        const float *L = reinterpret_cast<const float *>(lds[buf]);
        float ra[TW][2][8], fx[TW][8];
#pragma unroll
        for (int u = 0; u < TW; ++u) {
            const int iu = wa * TW + u;
#pragma unroll
            for (int h = 0; h < 2; ++h)
#pragma unroll
                for (int v = 0; v < 8; ++v) ra[u][h][v] = L[(TCH + iu * 256 + h * 128) * 4 + lane * 8 + v];
#pragma unroll
            for (int v = 0; v < 8; ++v) fx[u][v] = L[(TCH + RCH + iu * 8) * 4 + g * 8 + v];
        }
#pragma unroll
        for (int s = 0; s < ZW; ++s) {
            const int is = wb * ZW + s;
            // template rows 8g+kc (tile 0) and 16+8g+kc (tile 1) of basis j
            float tb[2][8];
#pragma unroll
            for (int v = 0; v < 8; ++v) {
                tb[0][v] = real ? L[((is * ZP + jz) * NB + jb) * 48 + 8 * g + v] : 0.f;
                tb[1][v] = real ? L[((is * ZP + jz) * NB + jb) * 48 + 16 + 8 * g + v] : 0.f;
            }
#pragma unroll
            for (int u = 0; u < TW; ++u) {
                G[u][s] = zero4;
                float4v d[2];
#pragma unroll
                for (int h = 0; h < 2; ++h) {
                    float4v acc = zero4;
#pragma unroll
                    for (int kc = 0; kc < 8; ++kc) acc = mfma4(ra[u][h][kc], tb[h][kc], acc);
                    d[h] = acc;
                }
#pragma unroll
                for (int h = 0; h < 2; ++h)
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        const float xv = fcolumn ? fx[u][4 * h + r] : d[h][r];
                        G[u][s] = mfma4(xv, xv, G[u][s]);
                    }
#pragma unroll
                for (int r = 0; r < 4; ++r) Gd[u][s][r] += (double)G[u][s][r];
            }
        }
        if (more) store(buf ^ 1);
        __syncthreads();
    }

    double *ne = reinterpret_cast<double *>(static_cast<char *>(a.work) + ne_off);
    const long long R = a.total_rows;
#pragma unroll
    for (int u = 0; u < TW; ++u) {
        const int item = itemb + wa * TW + u;
        if (item >= a.nitems) continue;
        const long long row0 = a.item_out[item];
#pragma unroll
        for (int s = 0; s < ZW; ++s) {
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                // per redshift q of the group: its NB x NB block, the f column
                // (y) and the always-zero column after it (y_lo of the 16-bit modes)
                const int i = 4 * g + r;
                const int q = i / NB, ii = i - q * NB;
                const int iz = zb + (wb * ZW + s) * ZP + q;
                const int jj = (j >= i && j < (q + 1) * NB) ? j - q * NB
                               : (j == ZP * NB || j == ZP * NB + 1) ? NB + j - ZP * NB : -1;
                if (i < ZP * NB && jj >= 0 && iz < nzi)
                    ne[rr::ne_index(ii, jj, NB, true) * R + row0 + iz] = Gd[u][s][r];
            }
        }
    }
}

template <int NB, int WA, int WB, int TW, int ZW>
int launch_tile(const rr::ZArgs &a, hipStream_t s) {
    constexpr int TT = WA * TW, ZR = WB * ZW * rr::zpack(NB);
    auto kernel = zscan_mfma32_kernel<NB, WA, WB, TW, ZW>;
    const rr::TileMap map = rr::make_tile_map((a.nitems + TT - 1) / TT, (a.max_item_nz + ZR - 1) / ZR,
                                              rr::resident_per_xcd(kernel, 256),
                                              TT * 2.0 * 128 * 16, ZR * NB * 12.0 * 16);
    const long long ne_off = rr::align256((long long)a.nitems * 2 * 8);
    hipLaunchKernelGGL(kernel, dim3((unsigned)map.blocks), dim3(256), 0, s, a, ne_off, map);
    return rr::check_launch("rr_zscan(f32)");
}

template <int NB>
int launch_nb(const rr::ZArgs &a, bool uniform, hipStream_t s) {
    if (uniform) return launch_tile<NB, 2, 2, 2, 2>(a, s);
    return launch_tile<NB, 1, 4, 1, 2>(a, s);
}

}  // namespace

namespace rr {

int zscan_mfma32(const ZArgs &a, hipStream_t s) {
    const bool u = a.uniform_z != 0;
    switch (a.nb) {
        case 1: return launch_nb<1>(a, u, s);
        case 2: return launch_nb<2>(a, u, s);
        case 3: return launch_nb<3>(a, u, s);
        case 4: return launch_nb<4>(a, u, s);
        case 5: return launch_nb<5>(a, u, s);
        case 6: return launch_nb<6>(a, u, s);
        case 7: return launch_nb<7>(a, u, s);
        case 8: return launch_nb<8>(a, u, s);
        case 9: return launch_nb<9>(a, u, s);
        case 10: return launch_nb<10>(a, u, s);
        case 11: return launch_nb<11>(a, u, s);
        case 12: return launch_nb<12>(a, u, s);
        default: return fail(RR_EUNSUPPORTED, "rr_zscan(f32): nb=%d unsupported", a.nb);
    }
}

int pack_mixed32(const double *d_rband, const double *d_w, const double *d_f, int ntarget,
                 int npix_pad, void *d_rtile, void *d_sfcol, double *d_scales, hipStream_t s) {
    hipLaunchKernelGGL(pack32_kernel, dim3(ntarget), dim3(64), 0, s, d_rband, d_w, d_f, npix_pad,
                       static_cast<float *>(d_rtile), static_cast<float *>(d_sfcol), d_scales);
    return check_launch("rr_pack_mixed(f32)");
}

}  // namespace rr
