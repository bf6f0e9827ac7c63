// Split-fp16 chi2 scan on the matrix cores (rr_zscan with RR_PREC_F16X3).
//
// Every 16-bit operand of the fp16 kernel (zscan_mfma.hip) is carried as an
// unevaluated sum hi + lo of two fp16 numbers (22 significant bits):
//
//   template  T = Th + Tl      (cache planes [z][hi, lo][b][row])
//   R tile    R = Rh + Rl      (packed by rr_pack_mixed, hi and lo tiles)
//
// and the products drop only the lo x lo terms:
//
//   D = Rh Tl + Rl Th + Rh Th          3 x v_mfma_f32_16x16x32_f16, chained
//                                      through the accumulator, small first
//   X = D split into Xh = fp16(D), 2 Xl = 2 fp16(D - Xh)   (v_fma_mix)
//   G += Xh' Xh + Xh' (2 Xl)           2 MFMAs; the symmetric part of G is
//                                      Xh'Xh + Xh'Xl + Xl'Xh
//
// G is accumulated in fp32 over two tile pairs (64 pixels) and then added
// into FP64 registers, so the sum over the ~8000 pixels of a spectrum does
// not collect fp32 rounding.  The epilogue writes the symmetric part of the
// FP64 sums; zscan_finish.hip solves and forms chi2 exactly as for the other
// mixed modes.  sqrt(w) f rides in K slots 0 and 1 of the hi R tiles (outside
// the band, as in the fp16 kernel) and is exact in fp16 hi + lo.
//
// Reference path replaced: zscan.py:670-737 (the GPU path).
#include "common.hpp"
#include "zscan_args.hpp"
#include "xcd_map.hpp"
#include "mfma16.hpp"

namespace {

constexpr int kFlushPairs = 2;   // fp32 partial sums span at most this many tile pairs

// rtile (nt, 2, ntile, 64, 8) fp16: [hi, lo plane][tile q][lane l][slot v];
// lane (i = l%16, g = l/16) slot v -> sqrt(w_p) rscale R[p][8g+v-i-3], p =
// 16q+i, split hi + lo.  K slots 0, 1 of group g = 0 of the hi tile hold
// sqrt(w_p) f_p fscale split hi + lo (zero in the lo tile).  The hi plane of a
// target is, value for value, the fp16 tile set of rr_pack_mixed(F16) (same
// scales, same rounding), so the fp16 normal equations of F16R read it.
// Packing: the per-target power-of-two scales (one wave per target), then
// every tile of every target in parallel (one wave per tile).
__global__ __launch_bounds__(64) void scales3_kernel(const double *__restrict__ w,
                                                     const double *__restrict__ f, int npix_pad,
                                                     double *__restrict__ scales) {
    const int t = blockIdx.x;
    const int lane = threadIdx.x;
    const double *W = w + (long long)t * npix_pad;
    const double *F = f + (long long)t * npix_pad;
    double mw = 0.0, mf = 0.0;
    for (int p = lane; p < npix_pad; p += 64) {
        const double sw = sqrt(W[p]);
        mw = fmax(mw, sw);
        mf = fmax(mf, fabs(sw * F[p]));
    }
    mw = wave_max(mw);
    mf = wave_max(mf);
    if (lane == 0) {
        scales[2 * t] = (mw > 0.0) ? exp2(-ceil(log2(mw))) : 1.0;
        scales[2 * t + 1] = (mf > 0.0) ? exp2(8.0 - ceil(log2(mf))) : 1.0;
    }
}

constexpr int kPackTiles = 4;   // tiles (waves) per packing workgroup

__global__ __launch_bounds__(64 * kPackTiles) void pack3_kernel(const double *__restrict__ rband,
                                                                const double *__restrict__ w,
                                                                const double *__restrict__ f,
                                                                int npix_pad,
                                                                unsigned short *__restrict__ rtile,
                                                                const double *__restrict__ scales) {
    const int t = blockIdx.y;
    const int ntile = npix_pad / 16;
    const int q = blockIdx.x * kPackTiles + (threadIdx.x >> 6);
    if (q >= ntile) return;
    const int lane = threadIdx.x & 63;
    const double rscale = scales[2 * t], fscale = scales[2 * t + 1];
    const int i = lane & 15, g = lane >> 4;
    const int p = 16 * q + i;
    const long long tp = (long long)t * npix_pad + p;
    const double sw = sqrt(w[tp]);
    const double *Rb = rband + tp * 11;
    u16x8 hi8, lo8;
#pragma unroll
    for (int v = 0; v < 8; ++v) {
        const int k = 8 * g + v - i - 3;     // < 0 for g = 0, v <= 1
        const float r = (k >= 0 && k <= 10) ? (float)(Rb[k] * sw * rscale) : 0.f;
        const unsigned short h = to16<RR_PREC_F16>(r);
        hi8[v] = h;
        lo8[v] = to16<RR_PREC_F16>(r - from16<RR_PREC_F16>(h));
    }
    if (g == 0) {
        const float sf = (float)(sw * f[tp] * fscale);
        const unsigned short h = to16<RR_PREC_F16>(sf);
        hi8[0] = h;
        hi8[1] = to16<RR_PREC_F16>(sf - from16<RR_PREC_F16>(h));
    }
    unsigned short *RT = rtile + (long long)t * 2 * ntile * 512;
    *reinterpret_cast<u16x8 *>(RT + ((long long)q * 64 + lane) * 8) = hi8;
    *reinterpret_cast<u16x8 *>(RT + (((long long)ntile + q) * 64 + lane) * 8) = lo8;
}

typedef _Float16 half2v __attribute__((ext_vector_type(2)));

// x (fp32) -> xh = fp16(x) and 2 xl = 2 fp16(x - xh) for the 8 values of a
// Gram operand (two MFMA outputs).  m1 = -1 hidden from constant folding so
// that x - xh maps to v_fma_mix{lo,hi}_f16.
__device__ __forceinline__ void split8(float4v a, float4v b, float m1, u16x8 &xh, u16x8 &xl2) {
    xh = to16x8<RR_PREC_F16>(a, b);
    const half8 h = __builtin_bit_cast(half8, xh);
    half8 l;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        l[r] = (_Float16)__builtin_fmaf((float)h[r], m1, a[r]);
        l[4 + r] = (_Float16)__builtin_fmaf((float)h[4 + r], m1, b[r]);
    }
    l = l + l;   // exact doubling
    xl2 = __builtin_bit_cast(u16x8, l);
}

// Workgroup tile TT items x ZZ redshift groups over four wavefronts (WA x
// WB); a group is ZP = zpack(NB) redshifts side by side in the MFMA columns
// (as in zscan_mfma.hip).  Per tile pair the workgroup stages in LDS the hi
// and lo template rows of its ZZ*ZP redshifts and the hi and lo R tiles of
// its TT items, with direct
// global->LDS loads (lane l of a wave writes 16 bytes at the wave base + 16 l),
// double-buffered.
template <int NB, int WA, int WB, int TW, int ZW>
__global__ __launch_bounds__(256) void zscan_mfma3_kernel(rr::ZArgs a, long long ne_off, rr::TileMap map) {
    constexpr int ZP = rr::zpack(NB);    // redshifts side by side in the 16 MFMA columns
    constexpr int TT = WA * TW, ZZ = WB * ZW, ZR = ZZ * ZP, NB2 = 2 * NB;
    constexpr int TCH = ZR * NB2 * 6;    // 16-byte template chunks per pair: 48 rows x (hi, lo) x NB x ZR
    constexpr int RCH = TT * 256;        // 16-byte R chunks per pair: 2 tiles x (hi, lo) x 64 lanes x TT
    constexpr int NCH = TCH + RCH;
    constexpr int PER = (NCH + 255) / 256;
    __shared__ u16x8 lds[2][PER * 256];
    __shared__ u16x8 zero_chunk[3];

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
    if (zb >= nzi) return;                   // uniform over the workgroup

    const int ntile = a.npix_pad / 16;
    const int npair = ntile / 2;
    const long long rows = a.cache_rows;
    const unsigned short *__restrict__ T = static_cast<const unsigned short *>(a.tcache);
    const unsigned short *__restrict__ RT = static_cast<const unsigned short *>(a.rtile);

    const unsigned short *src[PER];
    int step[PER];
#pragma unroll
    for (int k = 0; k < PER; ++k) {
        int c = tid + 256 * k;
        if (c >= NCH) c = NCH - 1;           // duplicate the last chunk into the padding
        if (c < TCH) {
            const int s = c / (NB2 * 6), b2 = (c / 6) % NB2, part = c % 6;
            const int iz = zb + s;
            const int zc = z0 + (iz < nzi ? iz : nzi - 1);
            src[k] = T + ((long long)zc * NB2 + b2) * rows + 8 * part;
            step[k] = 32;
        } else {
            const int r = c - TCH, u = r >> 8, t = (r >> 7) & 1, h = (r >> 6) & 1, l = r & 63;
            const int it = itemb + u < a.nitems ? itemb + u : a.nitems - 1;
            const int tg = a.item_target[it];
            src[k] = RT + ((((long long)tg * 2 + h) * ntile + t) * 64 + l) * 8;
            step[k] = 2 * 64 * 8;
        }
    }
    auto load = [&](int qp, int buf) {
#pragma unroll
        for (int k = 0; k < PER; ++k)
            __builtin_amdgcn_global_load_lds(
                (const void *)(src[k] + (long long)qp * step[k]),
                (void __attribute__((address_space(3))) *)(&lds[buf][256 * k + 64 * wave]), 16, 0, 0);
    };

    // column j < ZP*NB: basis j%NB of redshift j/NB of the group; columns
    // ZP*NB, ZP*NB+1: sqrt(w) f
    const int j = lane & 15;
    const int g = lane >> 4;
    const bool real = j < ZP * NB;
    const int jz = real ? j / NB : 0, jb = real ? j - jz * NB : 0;
    // first dword of a B operand (K slots 0, 1 of group 0): zero for the basis
    // columns, unit vectors on slots 0 / 1 for the two f columns (hi only)
    constexpr unsigned one16 = 0x3C00u;
    const unsigned keep0 = g == 0 ? 0u : 0xffffffffu;
    const unsigned set0 = (g != 0) ? 0u : (j == ZP * NB) ? one16 : (j == ZP * NB + 1) ? (one16 << 16) : 0u;
    typedef unsigned u32x4 __attribute__((ext_vector_type(4)));
    auto bop = [&](u16x8 v, unsigned set) {
        u32x4 d = __builtin_bit_cast(u32x4, v);
        d[0] = (d[0] & keep0) | set;
        return __builtin_bit_cast(u16x8, d);
    };
    float m1 = -1.0f;
    asm volatile("" : "+s"(m1));

    float4v G[TW][ZW];
    double Gd[TW][ZW][4];
#pragma unroll
    for (int u = 0; u < TW; ++u)
#pragma unroll
        for (int s = 0; s < ZW; ++s) {
            G[u][s] = float4v{0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int r = 0; r < 4; ++r) Gd[u][s][r] = 0.0;
        }
    const float4v zero4 = {0.f, 0.f, 0.f, 0.f};

    if (tid < 3) zero_chunk[tid] = u16x8{0, 0, 0, 0, 0, 0, 0, 0};
    load(0, 0);
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    __syncthreads();

    for (int qp = 0; qp < npair; ++qp) {
        const int buf = qp & 1;
        const u16x8 *L = lds[buf];
        u16x8 ra[TW][2][2], bz[ZW][2][2];    // [item or z][tile][hi, lo]
#pragma unroll
        for (int u = 0; u < TW; ++u) {
            const int iu = wa * TW + u;
#pragma unroll
            for (int t = 0; t < 2; ++t)
#pragma unroll
                for (int h = 0; h < 2; ++h) ra[u][t][h] = L[TCH + iu * 256 + t * 128 + h * 64 + lane];
        }
#pragma unroll
        for (int s = 0; s < ZW; ++s) {
            const int is = wb * ZW + s;
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const u16x8 *tp = real ? L + ((is * ZP + jz) * NB2 + h * NB + jb) * 6 + g : zero_chunk;
                bz[s][0][h] = bop(tp[0], h == 0 ? set0 : 0u);
                bz[s][1][h] = bop(tp[2], h == 0 ? set0 : 0u);
            }
        }
        if (qp + 1 < npair) load(qp + 1, buf ^ 1);
#pragma unroll
        for (int s = 0; s < ZW; ++s) {
#pragma unroll
            for (int u = 0; u < TW; ++u) {
                float4v d[2];
#pragma unroll
                for (int t = 0; t < 2; ++t) {
                    float4v e = mfma<RR_PREC_F16>(ra[u][t][0], bz[s][t][1], zero4);
                    e = mfma<RR_PREC_F16>(ra[u][t][1], bz[s][t][0], e);
                    d[t] = mfma<RR_PREC_F16>(ra[u][t][0], bz[s][t][0], e);
                }
                u16x8 xh, xl2;
                split8(d[0], d[1], m1, xh, xl2);
                G[u][s] = mfma<RR_PREC_F16>(xh, xh, G[u][s]);
                G[u][s] = mfma<RR_PREC_F16>(xh, xl2, G[u][s]);
            }
        }
        if ((qp % kFlushPairs) == kFlushPairs - 1 || qp + 1 == npair) {
#pragma unroll
            for (int u = 0; u < TW; ++u)
#pragma unroll
                for (int s = 0; s < ZW; ++s) {
#pragma unroll
                    for (int r = 0; r < 4; ++r) Gd[u][s][r] += (double)G[u][s][r];
                    G[u][s] = zero4;
                }
        }
        __builtin_amdgcn_sched_barrier(0);
        asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
        __syncthreads();
    }

    // ---- epilogue: symmetric part of the FP64 sums -> workspace ----
    // lane (j, g) register r holds G[4g+r][j]; G[j][4g+r] sits in lane
    // (4g+r) + 16 (j/4), register j%4.  Per redshift q of the group: the upper
    // triangle of its NB x NB diagonal block and its rows of the f columns.
    double *ne = reinterpret_cast<double *>(static_cast<char *>(a.work) + ne_off);
    const long long R = a.total_rows;
#pragma unroll
    for (int u = 0; u < TW; ++u) {
        const int item = itemb + wa * TW + u;
        const long long row0 = item < a.nitems ? a.item_out[item] : 0;
#pragma unroll
        for (int s = 0; s < ZW; ++s) {
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int i = 4 * g + r;
                const int srcl = i + 16 * (j >> 2);
                double tr[4];
#pragma unroll
                for (int k = 0; k < 4; ++k) tr[k] = __shfl(Gd[u][s][k], srcl, 64);
                const double gji = (j & 3) == 0 ? tr[0] : (j & 3) == 1 ? tr[1] : (j & 3) == 2 ? tr[2] : tr[3];
                const int q = i / NB, ii = i - q * NB;
                const int iz = zb + (wb * ZW + s) * ZP + q;
                const int jj = (j >= i && j < (q + 1) * NB) ? j - q * NB
                               : (j == ZP * NB || j == ZP * NB + 1) ? NB + j - ZP * NB : -1;
                if (item < a.nitems && iz < nzi && i < ZP * NB && jj >= 0)
                    ne[rr::ne_index(ii, jj, NB, true) * R + row0 + iz] = 0.5 * (Gd[u][s][r] + gji);
            }
        }
    }
}

struct TileCfg {
    int wa, wb, tw, zw;
};
// entry 0 = default; 1 = the configuration for items with their own z ranges
constexpr TileCfg kTiles3[] = {{2, 2, 2, 2}, {1, 4, 1, 2}, {2, 2, 2, 2}, {1, 4, 2, 2},
                               {4, 1, 2, 2}, {2, 2, 2, 4}, {2, 2, 4, 2}, {1, 4, 4, 1}};
constexpr int kNumTiles3 = sizeof(kTiles3) / sizeof(kTiles3[0]);

template <int NB, int WA, int WB, int TW, int ZW>
int launch_tile(const rr::ZArgs &a, hipStream_t s) {
    constexpr int TT = WA * TW, ZR = WB * ZW * rr::zpack(NB);
    auto kernel = zscan_mfma3_kernel<NB, WA, WB, TW, ZW>;
    const rr::TileMap map = rr::make_tile_map((a.nitems + TT - 1) / TT, (a.max_item_nz + ZR - 1) / ZR,
                                              rr::resident_per_xcd(kernel, 256),
                                              TT * 4.0 * 64 * 16, ZR * NB * 12.0 * 16, true);
    const long long ne_off = rr::align256((long long)a.nitems * 2 * 8);
    hipLaunchKernelGGL(kernel, dim3((unsigned)map.blocks), dim3(256), 0, s, a, ne_off, map);
    return rr::check_launch("rr_zscan(f16x3)");
}

template <int NB>
int launch_nb(const rr::ZArgs &a, int tile, hipStream_t s) {
    switch (tile) {
        case 1: return launch_tile<NB, 1, 4, 1, 2>(a, s);
        case 2: return launch_tile<NB, 2, 2, 2, 2>(a, s);
        case 3: return launch_tile<NB, 1, 4, 2, 2>(a, s);
        case 4: return launch_tile<NB, 4, 1, 2, 2>(a, s);
        case 5: return launch_tile<NB, 2, 2, 2, 4>(a, s);
        case 6: return launch_tile<NB, 2, 2, 4, 2>(a, s);
        case 7: return launch_tile<NB, 1, 4, 4, 1>(a, s);
        default: return rr::fail(RR_EINVAL, "rr_zscan(f16x3): unknown tile %d", tile);
    }
}

}  // namespace

namespace rr {
int zscan_mfma3(const ZArgs &a, hipStream_t s) {
    int tile = a.tile;
    if (tile < 0 || tile >= kNumTiles3)
        return fail(RR_EINVAL, "rr_zscan: tile %d outside 0..%d", tile, kNumTiles3 - 1);
    if (tile == 0) tile = a.uniform_z ? 2 : 1;
    if (!a.uniform_z && kTiles3[tile].wa * kTiles3[tile].tw > 1) tile = 1;
    switch (a.nb) {
        case 1: return launch_nb<1>(a, tile, s);
        case 2: return launch_nb<2>(a, tile, s);
        case 3: return launch_nb<3>(a, tile, s);
        case 4: return launch_nb<4>(a, tile, s);
        case 5: return launch_nb<5>(a, tile, s);
        case 6: return launch_nb<6>(a, tile, s);
        case 7: return launch_nb<7>(a, tile, s);
        case 8: return launch_nb<8>(a, tile, s);
        case 9: return launch_nb<9>(a, tile, s);
        case 10: return launch_nb<10>(a, tile, s);
        case 11: return launch_nb<11>(a, tile, s);
        case 12: return launch_nb<12>(a, tile, s);
        default: return fail(RR_EUNSUPPORTED, "rr_zscan(f16x3): nb=%d unsupported", a.nb);
    }
}

int pack_mixed3(const double *d_rband, const double *d_w, const double *d_f, int ntarget, int npix_pad,
                void *d_rtile, double *d_scales, hipStream_t s) {
    hipLaunchKernelGGL(scales3_kernel, dim3(ntarget), dim3(64), 0, s, d_w, d_f, npix_pad, d_scales);
    const int ntile = npix_pad / 16;
    hipLaunchKernelGGL(pack3_kernel, dim3((ntile + kPackTiles - 1) / kPackTiles, ntarget),
                       dim3(64 * kPackTiles), 0, s, d_rband, d_w, d_f, npix_pad,
                       static_cast<unsigned short *>(d_rtile), d_scales);
    return check_launch("rr_pack_mixed(f16x3)");
}

int zscan_tiles3(int *cfg, int max) {
    const int n = kNumTiles3 < max ? kNumTiles3 : max;
    for (int i = 0; i < n; ++i) {
        cfg[4 * i] = kTiles3[i].wa;
        cfg[4 * i + 1] = kTiles3[i].wb;
        cfg[4 * i + 2] = kTiles3[i].tw;
        cfg[4 * i + 3] = kTiles3[i].zw;
    }
    return kNumTiles3;
}
}  // namespace rr
