// Explicit-residual chi2 of RR_PREC_F16R on the matrix cores.
//
// After the fp16 normal equations and their FP64 solve have given the
// coefficients c of every (target t, redshift z), the chi2 of the mode is the
// reference's residual sum (zscan.py:777-780)
//
//   chi2(t, z) = sum_p ( sqrt(w_p) f_p - sum_k sqrt(w_p) R[p][k] u_{p+k-5} )^2,
//   u_q = sum_b T_z[q][b] c_b(t, z),
//
// which the lane kernel (zscan_lanes.hip) evaluates with nb + 12 FP32 FMAs per
// (pixel, redshift, target).  Here both products run as split-fp16 MFMAs
// (v_mfma_f32_16x16x32_f16, fp32 accumulation), each operand an unevaluated
// sum hi + lo of two fp16 numbers:
//
//   phase A, per redshift z, 16 pixels x 16 targets:
//     U_z[q][t] = sum_K A[q][K] B[K][t]     A = template rows, B = coefficients
//     K slots: Th.Ch | Tl.Ch | Th.Cl [| Tl.Cl]   (nb <= 8 and nb >= 11 keep the
//     fourth product; nb 9, 10 drop Tl.Cl, below 2^-22 of sum |T_b c_b|)
//   transpose through LDS: U_z[.][t] (lane = target) -> U_t[.][z] (lane =
//     redshift), split into Uh = fp16(U) and Ul = fp16(U - Uh) on the way
//   phase C, per target t, 16 output pixels x 16 redshifts:
//     D[p][z] = Rh Ul + Rl Uh + Rh Uh       A = sqrt(w) R tile (rr_pack_mixed
//     F16X3 tiles, the ones F16R scans with; input pixels [16q-8, 16q+24)),
//     B = U_t
//
// sqrt(w) f rides in K slots 0, 1 of the hi R tile (input pixels 16q-8 and
// 16q-7, outside every row's band); the U operand carries -beta there instead
// of U, so D is the scaled residual itself:  D = -rscale s_u r  with
// beta = rscale s_u / fscale, s_u = tscale 2^ec (all powers of two; 2^ec
// scales c so that sum |c_b| 2^ec lies in [32, 64) and every fp16 value stays
// in range).  chi2 = sum D^2 / (rscale s_u)^2: fp32 squares of one 32-pixel
// chunk, then FP64 sums.  The residual is exact to first order in the
// coefficient error (d chi2/dc = 0 at the least-squares c), so chi2 keeps
// the second-order error of the coefficients plus the ~2^-22 relative error
// of the split products.
//
// Workgroup: 16 items (targets) x ZB = 16 NZB consecutive redshifts, NW waves.
// Per 32-pixel chunk: phase A (the wave's ZB / NW redshifts, 2 pixel tiles)
// writes U of pixels [32c + 8, 32c + 40) into a 64-pixel ring in LDS; barrier;
// phase C (the wave's 16 / NW targets x NZB redshift blocks, 2 output tiles)
// reads U of pixels [32c - 8, 32c + 40) into registers; barrier; phase C's
// MFMAs run beside the next chunk's phase A.  Template rows and R tiles
// stream from global memory (L2) into registers one chunk ahead.
//
// Reference path replaced: zscan.py:752-780 (chi2 of each (target, z)).
#include "common.hpp"
#include "zscan_args.hpp"
#include "xcd_map.hpp"
#include "mfma16.hpp"

namespace {

constexpr int kRing = 64;     // pixels of U kept per (target, redshift)
constexpr int kRowH = 72;     // halfs per (target, redshift) ring row (64 + 8 pad: 9 x 16 B)

typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));

// K slot s of the phase-A products: 0 unused, 1 Th.Ch, 2 Tl.Ch, 3 Th.Cl, 4 Tl.Cl
// (basis index s % nb).
__host__ __device__ constexpr int uslot_kind(int nb, int nprod, int s) {
    return s < nb ? 1 : s < 2 * nb ? 2 : s < 3 * nb ? 3 : (nprod == 4 && s < 4 * nb) ? 4 : 0;
}

template <int NB>
struct ULay {
    static constexpr int nprod = rr::resid_u_nprod(NB);
    static constexpr int nmf = (nprod * NB + 31) / 32;   // phase-A MFMAs per tile
    static constexpr int ks = 32 * nmf;                  // halfs per (z, row) in the cache
};

// Coefficient scale 2^ec and beta = 2^eb of (item, z) from its coefficients;
// er, et, ef: exponents of rscale, tscale, fscale.  Evaluated identically by
// the phase-A lane (scales c) and the phase-C lane (unscales chi2).
template <int NB>
__device__ __forceinline__ void coef_scale(const double (&c)[NB], int er, int et, int ef, int &ec, int &eb) {
    double s = 0.0;
#pragma unroll
    for (int b = 0; b < NB; ++b) s += fabs(c[b]);
    ec = (s > 0.0 && s < 1e300) ? 5 - ilogb(s) : 0;
    eb = er + ec + et - ef;
    if (eb > 15) {            // fp16 max exponent: scale c down instead (u shrinks, stays exact)
        ec -= eb - 15;
        eb = 15;
    }
    if (eb < -24) eb = -24;   // then |f term| < 2^-24 of the model term: negligible
}

template <int NB>
__device__ __forceinline__ void load_coef(const rr::ZArgs &a, long long row, bool live, double (&c)[NB]) {
    const bool ok = live && (!a.status || a.status[row] == RR_STATUS_OK);
#pragma unroll
    for (int b = 0; b < NB; ++b) c[b] = ok ? a.zcoeff[row * NB + b] : 0.0;
}

template <int NB, int NZB, int NW>
__global__ __launch_bounds__(64 * NW) void zscan_resid_mfma_kernel(rr::ZArgs a, rr::TileMap map,
                                                                    const unsigned short *__restrict__ TU,
                                                                    const unsigned short *__restrict__ R3,
                                                                    const double *__restrict__ SC3) {
    using L = ULay<NB>;
    constexpr int NMF = L::nmf, KS = L::ks;
    constexpr int ZB = 16 * NZB;          // redshifts per workgroup
    constexpr int ZPW = ZB / NW;          // phase A: redshifts per wave
    constexpr int TPW = 16 / NW;          // phase C: targets per wave (each over NZB blocks)
    constexpr int TSH = ZB * kRowH + 8;   // halfs per target: a stride of 4 mod 64 dwords
    static_assert(ZPW >= 1 && TPW >= 1 && ZPW * NW == ZB && TPW * NW == 16, "bad configuration");
    __shared__ __attribute__((aligned(16))) unsigned short lds[2 * 16 * TSH];
    unsigned short *const Uh = lds;
    unsigned short *const Ul = lds + 16 * TSH;

    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int i16 = lane & 15, g = lane >> 4;
    int ti, tz;
    if (!rr::map_tile(map, blockIdx.x, ti, tz)) return;   // uniform over the workgroup
    const int nzi = a.item_nz[0];
    const int z0 = a.item_z0[0];
    const int zb0 = tz * ZB;              // < nzi by the map's z-tile count
    const int et = ilogb(a.tscale);
    const long long rows = a.cache_rows;
    const int ntile = a.npix_pad / 16;
    const int nchunk = ntile / 2;

    // ---- phase-A coefficient operands: lane = (target i16, K group g) ----
    const int itemA = ti * 16 + i16;
    const bool liveA = itemA < a.nitems;
    const int itA = liveA ? itemA : a.nitems - 1;
    const int tgtA = a.item_target[itA];
    const int erA = ilogb(SC3[2 * tgtA]), efA = ilogb(SC3[2 * tgtA + 1]);
    u16x8 cop[ZPW][NMF];
#pragma unroll
    for (int j = 0; j < ZPW; ++j) {
        const int iz = zb0 + wave * ZPW + j;
        const bool live = liveA && iz < nzi;
        double c[NB];
        load_coef<NB>(a, a.item_out[itA] + (live ? iz : 0), live, c);
        int ec, eb;
        coef_scale<NB>(c, erA, et, efA, ec, eb);
        unsigned short ch[NB], cl[NB];
#pragma unroll
        for (int b = 0; b < NB; ++b) {
            const double v = ldexp(c[b], ec);
            const _Float16 h = (_Float16)(float)v;
            ch[b] = __builtin_bit_cast(unsigned short, h);
            cl[b] = __builtin_bit_cast(unsigned short, (_Float16)(float)(v - (double)(float)h));
        }
#pragma unroll
        for (int m = 0; m < NMF; ++m)
#pragma unroll
            for (int v = 0; v < 8; ++v) {
                unsigned short val = 0;
#pragma unroll
                for (int gg = 0; gg < 4; ++gg) {
                    const int s = 32 * m + 8 * gg + v;
                    const int k = uslot_kind(NB, L::nprod, s);
                    const unsigned short cand = k == 0 ? (unsigned short)0 : k <= 2 ? ch[s % NB] : cl[s % NB];
                    val = g == gg ? cand : val;
                }
                cop[j][m][v] = val;
            }
    }

    // ---- phase-C per (target j, redshift block k): lane = (redshift i16, pixel group g) ----
    u32x2 negb[TPW][NZB];        // -beta as two fp16 (K slots 0, 1 of the U operand)
    int escale[TPW][NZB];        // chi2 = sum D^2 * 2^escale
    bool wlive[TPW][NZB];        // this lane writes chi2 of (item, z)
    int tgtC[TPW];
    bool itC[TPW];               // wave-uniform: target slot holds an item
#pragma unroll
    for (int j = 0; j < TPW; ++j) {
        const int item = ti * 16 + wave * TPW + j;
        itC[j] = item < a.nitems;
        const int it = itC[j] ? item : a.nitems - 1;
        tgtC[j] = a.item_target[it];
        const int er = ilogb(SC3[2 * tgtC[j]]), ef = ilogb(SC3[2 * tgtC[j] + 1]);
        const double *st = static_cast<const double *>(a.work);
        const bool weighted = st[2 * it + 1] != 0.0;     // zero weight: HUGE_CHI2 stays
#pragma unroll
        for (int k = 0; k < NZB; ++k) {
            const int iz = zb0 + 16 * k + i16;
            const bool live = itC[j] && iz < nzi;
            const long long row = a.item_out[it] + (live ? iz : 0);
            double c[NB];
            load_coef<NB>(a, row, live, c);
            int ec, eb;
            coef_scale<NB>(c, er, et, ef, ec, eb);
            const unsigned short nb16 = __builtin_bit_cast(unsigned short, (_Float16)(-ldexpf(1.f, eb)));
            negb[j][k] = u32x2{(unsigned)nb16 | ((unsigned)nb16 << 16), 0u};
            escale[j][k] = -2 * (er + ec + et);
            wlive[j][k] = live && weighted && g == 0 && (!a.status || a.status[row] == RR_STATUS_OK);
        }
    }

    // template rows of phase A: the row of pixel px0 + i16 of redshift slot j of
    // this wave.  Bases are wave-uniform (scalar), lane offsets 32-bit.
    const char *tzp[ZPW];
#pragma unroll
    for (int j = 0; j < ZPW; ++j) {
        int iz = zb0 + wave * ZPW + j;
        iz = iz < nzi ? iz : nzi - 1;
        tzp[j] = reinterpret_cast<const char *>(TU + ((long long)(z0 + iz) * rows + a.halo) * KS);
    }
    const unsigned t_lane = (unsigned)(i16 * KS + 8 * g) * 2u;    // bytes
    auto load_t = [&](int px0, u16x8 (&dst)[ZPW][NMF]) {
#pragma unroll
        for (int j = 0; j < ZPW; ++j) {
            const char *src = tzp[j] + __builtin_amdgcn_readfirstlane(px0 * (KS * 2)) + t_lane;
#pragma unroll
            for (int m = 0; m < NMF; ++m) dst[j][m] = *reinterpret_cast<const u16x8 *>(src + 64 * m);
        }
    };
    float m1 = -1.0f;                     // hidden from constant folding: U - Uh is one v_fma_mix
    asm volatile("" : "+s"(m1));
    // phase A on the 16-pixel tile starting at px0 (px0 + 8 a multiple of 16):
    // the MFMAs of every redshift of the wave first, then the splits and stores
    auto phase_a = [&](int px0, const u16x8 (&top)[ZPW][NMF]) {
        const int ro = (px0 + 8) & (kRing - 1);
        float4v d[ZPW];
#pragma unroll
        for (int j = 0; j < ZPW; ++j) {
            d[j] = float4v{0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int m = 0; m < NMF; ++m) d[j] = mfma<RR_PREC_F16>(top[j][m], cop[j][m], d[j]);
        }
#pragma unroll
        for (int j = 0; j < ZPW; ++j) {
            // lane holds U[pixel px0 + 4g + r][target i16] of redshift wave*ZPW + j
            const half4 h = __builtin_convertvector(d[j], half4);
            half4 l;
#pragma unroll
            for (int r = 0; r < 4; ++r) l[r] = (_Float16)__builtin_fmaf((float)h[r], m1, d[j][r]);
            const int off = i16 * TSH + (wave * ZPW + j) * kRowH + ro + 4 * g;
            *reinterpret_cast<half4 *>(Uh + off) = h;
            *reinterpret_cast<half4 *>(Ul + off) = l;
        }
    };

    double acc[TPW][NZB];
#pragma unroll
    for (int j = 0; j < TPW; ++j)
#pragma unroll
        for (int k = 0; k < NZB; ++k) acc[j][k] = 0.0;

    // phase C of chunk c: output tiles 2c, 2c + 1.  read_c takes every U
    // operand of the wave (its targets, both tiles, every redshift block) into
    // registers; once every wave has done so (a barrier) the ring may be
    // rewritten, and mfma_c runs the chains beside the next chunk's phase A.
    typedef u32x4 UOps[TPW][NZB][2];
    auto read_c = [&](int c, UOps &uh, UOps &ul) {
#pragma unroll
        for (int j = 0; j < TPW; ++j) {
            const int tl = wave * TPW + j;
#pragma unroll
            for (int k = 0; k < NZB; ++k)
#pragma unroll
                for (int qq = 0; qq < 2; ++qq) {
                    const int q = 2 * c + qq;
                    const int off = tl * TSH + (16 * k + i16) * kRowH + ((16 * q + 8 * g) & (kRing - 1));
                    uh[j][k][qq] = *reinterpret_cast<const u32x4 *>(Uh + off);
                    ul[j][k][qq] = *reinterpret_cast<const u32x4 *>(Ul + off);
                }
        }
    };
    // slots past the last item compute on a repeated target (no branch: every
    // path waits for the same loads) and write nothing
    auto mfma_c = [&](const u16x8 (&rt)[TPW][2][2], const UOps &uh, const UOps &ul) {
#pragma unroll
        for (int j = 0; j < TPW; ++j) {
#pragma unroll
            for (int k = 0; k < NZB; ++k) {
                float part = 0.f;
#pragma unroll
                for (int qq = 0; qq < 2; ++qq) {
                    u32x4 h = uh[j][k][qq], l = ul[j][k][qq];
                    h.x = g == 0 ? negb[j][k].x : h.x;
                    l.x = g == 0 ? 0u : l.x;
                    const u16x8 bh = __builtin_bit_cast(u16x8, h), bl = __builtin_bit_cast(u16x8, l);
                    float4v d = {0.f, 0.f, 0.f, 0.f};
                    d = mfma<RR_PREC_F16>(rt[j][qq][0], bl, d);
                    d = mfma<RR_PREC_F16>(rt[j][qq][1], bh, d);
                    d = mfma<RR_PREC_F16>(rt[j][qq][0], bh, d);
#pragma unroll
                    for (int r = 0; r < 4; ++r) part = __builtin_fmaf(d[r], d[r], part);
                }
                acc[j][k] += (double)part;
            }
        }
    };
    // R tiles: wave-uniform base of target j's tiles, 16 bytes per lane
    const char *rz[TPW];
#pragma unroll
    for (int j = 0; j < TPW; ++j)
        rz[j] = reinterpret_cast<const char *>(R3 + (long long)tgtC[j] * 2 * ntile * 512);
    const unsigned r_lane = (unsigned)lane * 16u;
    auto load_r = [&](int c, u16x8 (&rt)[TPW][2][2]) {
#pragma unroll
        for (int j = 0; j < TPW; ++j)
#pragma unroll
            for (int qq = 0; qq < 2; ++qq)
#pragma unroll
                for (int h = 0; h < 2; ++h)
                    rt[j][qq][h] = *reinterpret_cast<const u16x8 *>(
                        rz[j] + __builtin_amdgcn_readfirstlane((h * ntile + 2 * c + qq) * 1024) + r_lane);
    };

    // operands of two chunks in flight: the loads for chunk c + 1 are issued
    // at the start of step c, before anything of step c waits on memory
    u16x8 ta[2][ZPW][NMF], tb[2][ZPW][NMF];   // template rows (tiles 0, 1 of a chunk)
    u16x8 ra[TPW][2][2], rb[TPW][2][2];       // R tiles (hi, lo of the chunk's 2 output tiles)
    {   // prologue: U of pixels [-8, 8)
        u16x8 tp[ZPW][NMF];
        load_t(-8, tp);
        phase_a(-8, tp);
    }
    load_t(8, ta[0]);
    load_t(24, ta[1]);
    load_r(0, ra);
    // step c: loads of chunk c + 1, phase A of chunk c (pixels [32c + 8,
    // 32c + 40)), barrier, the U operands of chunk c into registers, barrier,
    // the phase C chains of chunk c (which the next step's phase A follows
    // without a barrier).  The loads are issued on every step (the last one
    // re-reads its own chunk) so that every path reaches phase A with the same
    // loads outstanding and the wait for the current operands leaves the next
    // chunk's loads in flight.
    auto step = [&](int c, u16x8 (&tcur)[2][ZPW][NMF], u16x8 (&tnxt)[2][ZPW][NMF],
                    u16x8 (&rcur)[TPW][2][2], u16x8 (&rnxt)[TPW][2][2]) {
        const int cn = c + 1 < nchunk ? c + 1 : c;
        load_t(32 * cn + 8, tnxt[0]);
        load_t(32 * cn + 24, tnxt[1]);
        load_r(cn, rnxt);
        phase_a(32 * c + 8, tcur[0]);
        phase_a(32 * c + 24, tcur[1]);
        __syncthreads();
        UOps uh, ul;
        read_c(c, uh, ul);
        __syncthreads();
        mfma_c(rcur, uh, ul);
    };
    // pairs of steps without a condition inside (a load whose only use sat in
    // a conditional block would be sunk into it, next to its use), then the
    // last chunk of an odd count
    int c = 0;
    for (; c + 1 < nchunk; c += 2) {
        step(c, ta, tb, ra, rb);
        step(c + 1, tb, ta, rb, ra);
    }
    if (c < nchunk) step(c, ta, tb, ra, rb);

    // sum over the four pixel groups (lanes i16, i16 + 16, + 32, + 48), fixed order
#pragma unroll
    for (int j = 0; j < TPW; ++j) {
        if (!itC[j]) continue;
        const int item = ti * 16 + wave * TPW + j;
#pragma unroll
        for (int k = 0; k < NZB; ++k) {
            double v = acc[j][k];
            v += __shfl_xor(v, 16, 64);
            v += __shfl_xor(v, 32, 64);
            if (wlive[j][k]) a.zchi2[a.item_out[item] + zb0 + 16 * k + i16] = ldexp(v, escale[j][k]);
        }
    }
}

// Residual-kernel configurations on the matrix cores: (redshift blocks of 16
// per workgroup, waves).  LDS: 147 KB at NZB 2 (one workgroup per CU), 74 KB
// at NZB 1.  Galaxy petal: {2, 16} 4.7 ms (4 waves per SIMD), {2, 8} 5.3 ms,
// {1, 4} 6.8 ms.
struct MfmaTile {
    int nzb, nw;
};
constexpr MfmaTile kMfmaTiles[] = {{2, 16}, {2, 8}, {1, 4}};
constexpr int kNumMfmaTiles = sizeof(kMfmaTiles) / sizeof(kMfmaTiles[0]);

template <int NB, int NZB, int NW>
int launch_cfg(const rr::ZArgs &a, hipStream_t s) {
    auto kernel = zscan_resid_mfma_kernel<NB, NZB, NW>;
    constexpr int ZB = 16 * NZB;
    // bytes per 32-pixel chunk: an item tile reads 16 targets' hi + lo R tiles,
    // a z tile ZB template rows of 32 pixels
    // per 32-pixel chunk an item tile reads 16 targets' hi + lo R tiles (64 KB),
    // a z tile ZB template rows (ZB x 2 KB); super-tiles that fit the XCD
    const rr::TileMap map = rr::make_tile_map((a.nitems + 15) / 16, (a.max_item_nz + ZB - 1) / ZB,
                                              rr::resident_per_xcd(kernel, 64 * NW), 65536.0,
                                              ZB * 2048.0, true);
    hipLaunchKernelGGL(kernel, dim3((unsigned)map.blocks), dim3(64 * NW), 0, s, a, map,
                       static_cast<const unsigned short *>(a.tcache_u),
                       static_cast<const unsigned short *>(a.rtile), a.scales);
    return rr::check_launch("rr_zscan(residual, MFMA)");
}

template <int NB>
int launch_nb(const rr::ZArgs &a, int cfg, hipStream_t s) {
    switch (cfg) {
        case 0: return launch_cfg<NB, 2, 16>(a, s);
        case 1: return launch_cfg<NB, 2, 8>(a, s);
        case 2: return launch_cfg<NB, 1, 4>(a, s);
        default: return rr::fail(RR_EINVAL, "rr_zscan(residual, MFMA): configuration %d", cfg);
    }
}

}  // namespace

namespace rr {

int zscan_resid_mfma_tiles(int *cfg, int max) {
    const int n = kNumMfmaTiles < max ? kNumMfmaTiles : max;
    for (int i = 0; i < n; ++i) {
        cfg[4 * i] = 2;                          // kind: matrix-core residual
        cfg[4 * i + 1] = kMfmaTiles[i].nw;       // waves per workgroup
        cfg[4 * i + 2] = 16;                     // items per workgroup
        cfg[4 * i + 3] = 16 * kMfmaTiles[i].nzb; // redshifts per workgroup
    }
    return kNumMfmaTiles;
}

int zscan_resid_mfma(const ZArgs &a, int cfg, hipStream_t s) {
    if (cfg < 0 || cfg >= kNumMfmaTiles) return fail(RR_EINVAL, "rr_zscan(residual, MFMA): configuration %d", cfg);
    switch (a.nb) {
        case 1: return launch_nb<1>(a, cfg, s);
        case 2: return launch_nb<2>(a, cfg, s);
        case 3: return launch_nb<3>(a, cfg, s);
        case 4: return launch_nb<4>(a, cfg, s);
        case 5: return launch_nb<5>(a, cfg, s);
        case 6: return launch_nb<6>(a, cfg, s);
        case 7: return launch_nb<7>(a, cfg, s);
        case 8: return launch_nb<8>(a, cfg, s);
        case 9: return launch_nb<9>(a, cfg, s);
        case 10: return launch_nb<10>(a, cfg, s);
        case 11: return launch_nb<11>(a, cfg, s);
        case 12: return launch_nb<12>(a, cfg, s);
        default: return fail(RR_EUNSUPPORTED, "rr_zscan(residual, MFMA): nb=%d unsupported", a.nb);
    }
}

}  // namespace rr
