// Mixed-precision chi2 scan on the matrix cores (rr_zscan with RR_PREC_F16 /
// RR_PREC_BF16) and the per-target packing it needs (rr_pack_mixed).
//
// For one target, 16 output pixels [16q, 16q+16) and one redshift:
//
//   D = Rq x Tq + Cq                      v_mfma_f32_16x16x32_{f16,bf16}
//
//   Rq (16 x 32)  rows = output pixels, cols = input pixels [16q-8, 16q+24):
//                 the 11-diagonal band scaled by sqrt(w_p) (a dense band
//                 block, packed once per target, reused for every z)
//   Tq (32 x 16)  template cache rows of those input pixels, cols = basis
//                 b < nb; columns nb and nb+1 are unit vectors on K slots 0
//                 and 1, and K slots 0 and 1 of the basis columns are zero
//   Rq K slots 0, 1 (outside the band: input pixels 16q-8, 16q-7 are at
//                 least 8 pixels from every output row) hold sqrt(w) f split
//                 hi + lo
//
// so D[p][b<nb] = sqrt(w_p) (R T)[p][b] and D[p][nb..nb+1] = sqrt(w_p) f_p.
// The f32 outputs of two consecutive tiles (32 pixels) are rounded to 16 bits
// and fed back, unchanged in place, as both operands of a third MFMA:
//
//   G += D^T D                            (16 x 16, K = 32 pixels)
//
// With the 16x16 output layout (lane = column l%16, rows 4(l/16)+r) a lane's
// four accumulators of tile q and of tile q+1 are exactly the 8 K-slots the
// operand layout expects for column l%16, so no data moves between the two
// stages.  G holds M = A'WA (b,b < nb) and y = A'Wf (b, nb / nb+1), scaled;
// the kernel writes them to the workspace and zscan_finish.hip unscales them,
// solves M c = y by FP64 Cholesky (LU fallback) and forms
// chi2 = f'Wf - 2 c'y + c'Mc with f'Wf summed exactly in FP64.
//
// Reference path replaced: zscan.py:670-737 (the GPU path) — the M/y
// contractions run in 16-bit on MFMA with fp32 accumulation, everything after
// them in FP64.
#include "common.hpp"
#include "zscan_args.hpp"
#include "xcd_map.hpp"
#include "mfma16.hpp"

namespace {

// ---------------------------------------------------------------------------
// Packing: one workgroup (one wave) per target.
//   rtile  (ntarget, ntile, 64, 8) 16-bit: lane l of tile q holds row i=l%16,
//          K slots v=0..7 of group g=l/16 -> input pixel 16q-8+8g+v; value
//          sqrt(w_p) * rscale * R[p][k'] with p = 16q+i, k' = 8g+v-i-3.
//          K slots 0 and 1 of group g = 0 (always outside the band) hold
//          sqrt(w_p) f_p * fscale split hi + lo.
//   scales (ntarget, 2) f64: rscale, fscale (powers of two).
// ---------------------------------------------------------------------------
template <int TY>
__global__ __launch_bounds__(64) void pack_mixed_kernel(const double *__restrict__ rband,
                                                        const double *__restrict__ w,
                                                        const double *__restrict__ f, int npix_pad,
                                                        unsigned short *__restrict__ rtile,
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
    // largest power of two with mw*rscale <= 1 and mf*fscale <= 256
    const double rscale = (mw > 0.0) ? exp2(-ceil(log2(mw))) : 1.0;
    const double fscale = (mf > 0.0) ? exp2(8.0 - ceil(log2(mf))) : 1.0;
    if (lane == 0) {
        scales[2 * t] = rscale;
        scales[2 * t + 1] = fscale;
    }
    const int ntile = npix_pad / 16;
    unsigned short *RT = rtile + (long long)t * ntile * 64 * 8;
    const int i = lane & 15, g = lane >> 4;
    for (int q = 0; q < ntile; ++q) {
        const int p = 16 * q + i;
        const double sw = sqrt(W[p]);
        u16x8 v8;
#pragma unroll
        for (int v = 0; v < 8; ++v) {
            const int k = 8 * g + v - i - 3;     // < 0 for g = 0, v <= 1
            const double r = (k >= 0 && k <= 10) ? Rb[(long long)p * 11 + k] * sw * rscale : 0.0;
            v8[v] = to16<TY>((float)r);
        }
        if (g == 0) {
            const float sf = (float)(sw * F[p] * fscale);
            const unsigned short h = to16<TY>(sf);
            v8[0] = h;
            v8[1] = to16<TY>(sf - from16<TY>(h));
        }
        *reinterpret_cast<u16x8 *>(RT + ((long long)q * 64 + lane) * 8) = v8;
    }
}

// ---------------------------------------------------------------------------
// Scan kernel.  Workgroup tile = TT items x ZZ redshift groups, split over
// its four wavefronts as WA (items) x WB (groups); a wavefront owns TW = TT/WA
// items x ZW = ZZ/WB groups.  A group is ZP = zpack(NB) redshifts whose basis
// columns sit side by side in one MFMA tile (ZP > 1 for NB <= 7: one R-apply
// and one Gram MFMA then serve ZP redshifts; the off-diagonal blocks of the
// Gram between them are not used).  For every tile pair (32 pixels) the
// workgroup stages in LDS the template rows of its ZZ*ZP redshifts (shared by
// its TT items) and the R tiles of its TT items (shared by its redshifts), with
// direct global->LDS loads (global_load_lds_dwordx4: lane l of a wave writes
// 16 bytes at the wave's LDS base + 16 l).  LDS is double-buffered: the loads
// of pair p+1 are in flight while the MFMAs consume pair p.  All TT items must
// share one cache z range (uniform_z) unless TT == 1.
// ---------------------------------------------------------------------------
template <int NB, int TY, int WA, int WB, int TW, int ZW>
__global__ __launch_bounds__(256) void zscan_mfma_kernel(rr::ZArgs a, long long ne_off, rr::TileMap map) {
    constexpr int ZP = rr::zpack(NB);    // redshifts side by side in the 16 MFMA columns
    constexpr int TT = WA * TW, ZZ = WB * ZW, ZR = ZZ * ZP;
    constexpr int TCH = ZR * NB * 6;     // 16-byte template chunks per pair (48 rows x NB x ZR)
    constexpr int RCH = TT * 2 * 64;     // 16-byte R-tile chunks per pair (2 tiles x 64 lanes)
    constexpr int NCH = TCH + RCH;
    constexpr int PER = (NCH + 255) / 256;
    // every thread loads PER chunks; the tail of the last round lands in padding
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

    // ---- this thread's share of the staging: source base and per-pair step
    const unsigned short *src[PER];
    int step[PER];
#pragma unroll
    for (int k = 0; k < PER; ++k) {
        int c = tid + 256 * k;
        if (c >= NCH) c = NCH - 1;           // duplicate the last chunk into the padding
        if (c < TCH) {
            const int s = c / (NB * 6), b = (c / 6) % NB, part = c % 6;
            const int iz = zb + s;
            const int zc = z0 + (iz < nzi ? iz : nzi - 1);
            src[k] = T + ((long long)zc * NB + b) * rows + 8 * part;
            step[k] = 32;
        } else {
            const int r = c - TCH, u = r >> 7, h = (r >> 6) & 1, l = r & 63;
            const int it = itemb + u < a.nitems ? itemb + u : a.nitems - 1;
            const int t = a.item_target[it];
            src[k] = RT + (((long long)t * ntile * a.rtile_planes + h) * 64 + l) * 8;
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

    // ---- operand addressing of this lane
    // Column j < ZP*NB of operand B / output is basis j%NB of redshift j/NB
    // of the lane's redshift group; columns ZP*NB, ZP*NB+1 carry sqrt(w) f.
    const int j = lane & 15;
    const int g = lane >> 4;   // K group / output row group
    const bool real = j < ZP * NB;
    // B operand, first dword (K slots 0, 1 of group g = 0): zero for the basis
    // columns (those slots carry sqrt(w) f in the R tile), a unit on slot 0 /
    // slot 1 for the two f columns; other lanes pass their template rows.
    constexpr unsigned one16 = TY == RR_PREC_F16 ? 0x3C00u : 0x3F80u;
    const unsigned keep0 = g == 0 ? 0u : 0xffffffffu;
    const unsigned set0 = (g != 0) ? 0u : (j == ZP * NB) ? one16 : (j == ZP * NB + 1) ? (one16 << 16) : 0u;
    typedef unsigned u32x4 __attribute__((ext_vector_type(4)));
    auto bop = [&](u16x8 v) {
        u32x4 d = __builtin_bit_cast(u32x4, v);
        d[0] = (d[0] & keep0) | set0;
        return __builtin_bit_cast(u16x8, d);
    };

    float4v G[TW][ZW];
#pragma unroll
    for (int u = 0; u < TW; ++u)
#pragma unroll
        for (int s = 0; s < ZW; ++s) G[u][s] = float4v{0.f, 0.f, 0.f, 0.f};
    const float4v zero4 = {0.f, 0.f, 0.f, 0.f};

    if (tid < 3) zero_chunk[tid] = u16x8{0, 0, 0, 0, 0, 0, 0, 0};
    load(0, 0);
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    __syncthreads();

    for (int qp = 0; qp < npair; ++qp) {
        const int buf = qp & 1;
        // all operands of this pair into registers first, then the loads of
        // the next pair: they complete under this pair's MFMAs
        const u16x8 *L = lds[buf];
        u16x8 ra[TW][2], bz[ZW][2];
#pragma unroll
        for (int u = 0; u < TW; ++u) {
            const int iu = wa * TW + u;
            ra[u][0] = L[TCH + iu * 128 + lane];
            ra[u][1] = L[TCH + iu * 128 + 64 + lane];
        }
#pragma unroll
        for (int s = 0; s < ZW; ++s) {
            const int is = wb * ZW + s;
            const u16x8 *tp = real ? L + (is * ZP * NB + j) * 6 + g : zero_chunk;
            bz[s][0] = bop(tp[0]);
            bz[s][1] = bop(tp[2]);
        }
        if (qp + 1 < npair) load(qp + 1, buf ^ 1);
#pragma unroll
        for (int s = 0; s < ZW; ++s) {
#pragma unroll
            for (int u = 0; u < TW; ++u) {
                const float4v d0 = mfma<TY>(ra[u][0], bz[s][0], zero4);
                const float4v d1 = mfma<TY>(ra[u][1], bz[s][1], zero4);
                const u16x8 x = to16x8<TY>(d0, d1);
                G[u][s] = mfma<TY>(x, x, G[u][s]);
            }
        }
        // the wait for the next pair's loads stays after every MFMA of this one
        __builtin_amdgcn_sched_barrier(0);
        asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
        __syncthreads();
    }

    // ---- epilogue: scaled normal equations -> workspace ----
    // lane (j, g) register r holds G[4g+r][j]; per redshift q of the group
    // keep the upper triangle of its NB x NB diagonal block and its rows of
    // the two f columns ([M | y_hi | y_lo]).
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
                const int i = 4 * g + r;
                const int q = i / NB, ii = i - q * NB;
                const int iz = zb + (wb * ZW + s) * ZP + q;
                const int jj = (j >= q * NB + ii && j < (q + 1) * NB) ? j - q * NB
                               : (j == ZP * NB || j == ZP * NB + 1) ? NB + j - ZP * NB : -1;
                if (i < ZP * NB && jj >= 0 && iz < nzi)
                    ne[rr::ne_index(ii, jj, NB, true) * R + row0 + iz] = (double)G[u][s][r];
            }
        }
    }
}

// Tile configurations (WA, WB, TW, ZW) selected by rr_zscan_desc.tile
// (0 = default); configurations with WA*TW > 1 need uniform_z.
struct TileCfg {
    int wa, wb, tw, zw;
};
constexpr TileCfg kTiles[] = {{2, 2, 4, 4}, {1, 4, 1, 4}, {2, 2, 4, 4}, {2, 2, 2, 4},
                              {4, 1, 2, 8}, {1, 4, 4, 2}, {2, 2, 4, 8}, {1, 4, 4, 4},
                              {4, 1, 4, 4}, {2, 2, 8, 2}};
constexpr int kNumTiles = sizeof(kTiles) / sizeof(kTiles[0]);

template <int NB, int TY, int WA, int WB, int TW, int ZW>
int launch_tile(const rr::ZArgs &a, hipStream_t s) {
    constexpr int TT = WA * TW, ZR = WB * ZW * rr::zpack(NB);
    auto kernel = zscan_mfma_kernel<NB, TY, WA, WB, TW, ZW>;
    const rr::TileMap map = rr::make_tile_map((a.nitems + TT - 1) / TT, (a.max_item_nz + ZR - 1) / ZR,
                                              rr::resident_per_xcd(kernel, 256),
                                              TT * 2.0 * 64 * 16, ZR * NB * 6.0 * 16, true);
    const long long ne_off = rr::align256((long long)a.nitems * 2 * 8);
    hipLaunchKernelGGL(kernel, dim3((unsigned)map.blocks), dim3(256), 0, s, a, ne_off, map);
    return rr::check_launch("rr_zscan(mixed)");
}

template <int NB, int TY>
int launch_nb(const rr::ZArgs &a, int tile, hipStream_t s) {
    switch (tile) {
        case 1: return launch_tile<NB, TY, 1, 4, 1, 4>(a, s);
        case 2: return launch_tile<NB, TY, 2, 2, 4, 4>(a, s);
        case 3: return launch_tile<NB, TY, 2, 2, 2, 4>(a, s);
        case 4: return launch_tile<NB, TY, 4, 1, 2, 8>(a, s);
        case 5: return launch_tile<NB, TY, 1, 4, 4, 2>(a, s);
        case 6: return launch_tile<NB, TY, 2, 2, 4, 8>(a, s);
        case 7: return launch_tile<NB, TY, 1, 4, 4, 4>(a, s);
        case 8: return launch_tile<NB, TY, 4, 1, 4, 4>(a, s);
        case 9: return launch_tile<NB, TY, 2, 2, 8, 2>(a, s);
        default: return rr::fail(RR_EINVAL, "rr_zscan(mixed): unknown tile %d", tile);
    }
}

template <int TY>
int launch_ty(const rr::ZArgs &a, int tile, hipStream_t s) {
    switch (a.nb) {
        case 1: return launch_nb<1, TY>(a, tile, s);
        case 2: return launch_nb<2, TY>(a, tile, s);
        case 3: return launch_nb<3, TY>(a, tile, s);
        case 4: return launch_nb<4, TY>(a, tile, s);
        case 5: return launch_nb<5, TY>(a, tile, s);
        case 6: return launch_nb<6, TY>(a, tile, s);
        case 7: return launch_nb<7, TY>(a, tile, s);
        case 8: return launch_nb<8, TY>(a, tile, s);
        case 9: return launch_nb<9, TY>(a, tile, s);
        case 10: return launch_nb<10, TY>(a, tile, s);
        case 11: return launch_nb<11, TY>(a, tile, s);
        case 12: return launch_nb<12, TY>(a, tile, s);
        default: return rr::fail(RR_EUNSUPPORTED, "rr_zscan(mixed): nb=%d unsupported", a.nb);
    }
}

}  // namespace

namespace rr {
int zscan_mfma(const ZArgs &a, hipStream_t s) {
    int tile = a.tile;
    if (tile < 0 || tile >= kNumTiles)
        return fail(RR_EINVAL, "rr_zscan: tile %d outside 0..%d", tile, kNumTiles - 1);
    if (tile == 0) tile = a.uniform_z ? 2 : 1;
    if (!a.uniform_z && kTiles[tile].wa * kTiles[tile].tw > 1) tile = 1;
    if (a.precision == RR_PREC_F16) return launch_ty<RR_PREC_F16>(a, tile, s);
    return launch_ty<RR_PREC_BF16>(a, tile, s);
}
}  // namespace rr

namespace rr {
int zscan_tiles16(int *cfg, int max) {
    const int n = kNumTiles < max ? kNumTiles : max;
    for (int i = 0; i < n; ++i) {
        cfg[4 * i] = kTiles[i].wa;
        cfg[4 * i + 1] = kTiles[i].wb;
        cfg[4 * i + 2] = kTiles[i].tw;
        cfg[4 * i + 3] = kTiles[i].zw;
    }
    return kNumTiles;
}
}  // namespace rr

extern "C" int rr_zscan_tiles(int *cfg, int max) {
    if (!cfg || max < 0) return rr::fail(RR_EINVAL, "rr_zscan_tiles: bad arguments");
    return rr::zscan_tiles16(cfg, max);
}

extern "C" int rr_pack_mixed(const double *d_rband, const double *d_w, const double *d_f,
                             int ntarget, int npix_pad, int precision, void *d_rtile,
                             void *d_sfcol, double *d_scales, void *stream) {
    if (ntarget < 0 || npix_pad < 32 || (npix_pad % 32) != 0)
        return rr::fail(RR_EINVAL, "rr_pack_mixed: bad sizes ntarget=%d npix_pad=%d", ntarget, npix_pad);
    if (!d_rband || !d_w || !d_f || !d_rtile || !d_scales || (precision == RR_PREC_F32 && !d_sfcol))
        return rr::fail(RR_EINVAL, "rr_pack_mixed: null pointer");
    if (ntarget == 0) return RR_OK;
    hipStream_t s = static_cast<hipStream_t>(stream);
    unsigned short *rt = static_cast<unsigned short *>(d_rtile);
    if (precision == RR_PREC_F16)
        hipLaunchKernelGGL(pack_mixed_kernel<RR_PREC_F16>, dim3(ntarget), dim3(64), 0, s, d_rband, d_w,
                           d_f, npix_pad, rt, d_scales);
    else if (precision == RR_PREC_BF16)
        hipLaunchKernelGGL(pack_mixed_kernel<RR_PREC_BF16>, dim3(ntarget), dim3(64), 0, s, d_rband, d_w,
                           d_f, npix_pad, rt, d_scales);
    else if (precision == RR_PREC_F16X3)
        return rr::pack_mixed3(d_rband, d_w, d_f, ntarget, npix_pad, d_rtile, d_scales, s);
    else if (precision == RR_PREC_F32)
        return rr::pack_mixed32(d_rband, d_w, d_f, ntarget, npix_pad, d_rtile, d_sfcol, d_scales, s);
    else
        return rr::fail(RR_EINVAL, "rr_pack_mixed: precision must be RR_PREC_F16, _BF16, _F16X3 or _F32");
    return rr::check_launch("rr_pack_mixed");
}
