// Scan kernels with one item (target) per lane: the plain FP64 scan of a
// uniform redshift grid, and the explicit-residual chi2 of the fp16 scan
// (RR_PREC_F16R), plus the per-target packing both read (rr_pack_lanes).
//
// A wavefront holds 64 items at one redshift (or ZW redshifts), so every
// template value it needs is wave-uniform: the template cache is laid out
// [z][row][b] (basis fastest) and read with scalar loads straight into SGPRs,
// which the vector FMAs take as their scalar operand (the residual kernel can
// instead hold a block's rows in VGPRs and broadcast them with DPP).  Per-target data (R
// band, weights, flux) are per lane; rr_pack_lanes stores them lane-major,
// [target / 64][pixel][slot][target % 64], so each slot of a pixel is one
// coalesced 512-byte (FP64) or 256-byte (FP32) load across the wavefront.
//
// zscan_f64r_kernel (RR_PREC_F64, uniform items): per lane and pixel p
//   A[b] = sum_k R[p][k] T[p+k-5][b]                 (the 11-tap R-apply)
//   M[b1][b2] += w_p A[b1] A[b2],  y[b] += A[b] (w f)_p
// in exactly the order of zscan_f64_kernel (so the normal equations are the
// same doubles); the finish stage solves them (zscan_finish.hip).
// Reference path replaced: zscan.py:752-780 (CPU) / :670-737 (GPU).
//
// zscan_resid_kernel (RR_PREC_F16R, after the fp16 scan and its solve): the
// reference's explicit residual chi2 = sum_p w_p (f_p - (R T c)_p)^2
// (zscan.py:777-780) at the coefficients c of the fp16 normal equations, in
// FP32 with FP64 partial sums:
//   u_q = sum_b T[q][b] c_b                           (template model, pixel q)
//   r_p = sqrt(w_p) f_p - sum_k sqrt(w_p) R[p][k] u_{p+k-5}
//   chi2 = sum_p r_p^2
// chi2 evaluated at an approximate c exceeds the minimum by exactly
// (c - c*)' M (c - c*): second order in the coefficient error, so the fp16
// normal equations' first-order error (which the expansion f'Wf - 2c'y + c'Mc
// carries, amplified by its cancellation) does not reach chi2.
#include <utility>
#include "common.hpp"
#include "zscan_args.hpp"
#include "xcd_map.hpp"

namespace {

constexpr int kSlots64 = 13;   // FP64 lane pack: R[0..10], w, w*f
constexpr int kSlots32 = 12;   // FP32 lane pack: sqrt(w) R[0..10], sqrt(w) f

// Slot i of the pack holds target order[i] (order NULL: target i).
__global__ __launch_bounds__(64) void pack_lanes64_kernel(const double *__restrict__ rband,
                                                          const double *__restrict__ w,
                                                          const double *__restrict__ f,
                                                          const int *__restrict__ order, int nslots,
                                                          int npix_pad, long long npl,
                                                          double *__restrict__ out) {
    const long long p = blockIdx.x;
    const int g = blockIdx.y;
    const int l = threadIdx.x;
    const int slot = 64 * g + l;
    const bool on = slot < nslots && p < npix_pad;
    const int t = !on ? 0 : order ? order[slot] : slot;
    double *o = out + ((long long)g * npl + p) * kSlots64 * 64 + l;
    const long long tp = (long long)t * npix_pad + p;
#pragma unroll
    for (int k = 0; k < 11; ++k) o[k * 64] = on ? rband[tp * 11 + k] : 0.0;
    const double wv = on ? w[tp] : 0.0;
    o[11 * 64] = wv;
    o[12 * 64] = on ? wv * f[tp] : 0.0;
}

__global__ __launch_bounds__(64) void pack_lanes32_kernel(const double *__restrict__ rband,
                                                          const double *__restrict__ w,
                                                          const double *__restrict__ f,
                                                          const int *__restrict__ order, int nslots,
                                                          int npix_pad, long long npl,
                                                          float *__restrict__ out) {
    const long long p = blockIdx.x;
    const int g = blockIdx.y;
    const int l = threadIdx.x;
    const int slot = 64 * g + l;
    const bool on = slot < nslots && p < npix_pad;
    const int t = !on ? 0 : order ? order[slot] : slot;
    float *o = out + ((long long)g * npl + p) * kSlots32 * 64 + l;
    const long long tp = (long long)t * npix_pad + p;
    const double sw = on ? sqrt(w[tp]) : 0.0;
#pragma unroll
    for (int k = 0; k < 11; ++k) o[k * 64] = on ? (float)(sw * rband[tp * 11 + k]) : 0.f;
    o[11 * 64] = on ? (float)(sw * f[tp]) : 0.f;
}

// ---------------------------------------------------------------------------
// Lane data staged in LDS.  A workgroup's NW waves all take the same 64 items
// (at different redshifts), so the lane data of a block of PB pixels is
// copied once per workgroup, global -> LDS (global_load_lds_dwordx4: lane l
// of a wave writes 16 bytes at the wave's LDS base + 16 l), double-buffered:
// the copy of block j+1 is in flight while the waves compute block j.  The
// block is one contiguous run of PB*S*64 elements in the pack.
// ---------------------------------------------------------------------------
template <typename E, int S, int PB, int NW>
struct LaneStage {
    static constexpr int kBytes = PB * S * 64 * (int)sizeof(E);          // one block
    static constexpr int kInsts = (kBytes + 1023) / 1024;                // 1 KB wave copies per block
    // LDS elements per buffer: whole wave copies (the last one runs up to 1 KB
    // past the block, into the pack's tail padding, rr_pack_lanes_bytes)
    static constexpr int kElems = kInsts * 1024 / (int)sizeof(E);
    static constexpr int kBlockElems = PB * S * 64;                      // one block in the pack

    // issue the copy of block j (pixels [j PB, j PB + PB)) into buffer buf
    __device__ __forceinline__ static void load(const E *group, int j, E *lds_buf, int wave, int lane) {
        const char *src = reinterpret_cast<const char *>(group + (long long)j * kBlockElems);
        for (int i = wave; i < kInsts; i += NW)
            __builtin_amdgcn_global_load_lds(
                (const void *)(src + ((long long)i * 64 + lane) * 16),
                (void __attribute__((address_space(3))) *)(reinterpret_cast<char *>(lds_buf) + i * 1024), 16, 0, 0);
    }
};

// ---------------------------------------------------------------------------
// Plain FP64 scan: a workgroup = NW waves = NW consecutive redshifts of the
// same 64 items (one item per lane); workgroups follow the XCD tile map.
// ---------------------------------------------------------------------------
constexpr int kF64Block = 11;   // pixels per staged block (the pack is a multiple of 22 pixels)

// NW waves per workgroup: 12 = 3 waves per SIMD (<= 168 VGPRs), 8 = 2 (<= 256)
// The cache and the pack come in as separate restrict parameters: the template
// loads then cannot alias the staging copies, so they stay scalar loads.
template <int NB, int NW>
__global__ __launch_bounds__(64 * NW) void zscan_f64r_kernel(rr::ZArgs a, long long ne_off,
                                                              rr::TileMap map,
                                                              const double *__restrict__ tcache,
                                                              const double *__restrict__ rlane) {
    using Stage = LaneStage<double, kSlots64, kF64Block, NW>;
    __shared__ double lds[2][Stage::kElems];
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    int ti, tz;
    if (!rr::map_tile(map, blockIdx.x, ti, tz)) return;   // uniform over the workgroup
    const int nzi = a.item_nz[0];
    const int iz0 = tz * NW + wave;
    const bool zlive = iz0 < nzi;                          // idle waves still copy and sync
    const int iz = zlive ? iz0 : nzi - 1;
    int item = ti * 64 + lane;
    const bool live = item < a.nitems && zlive;
    if (item >= a.nitems) item = a.nitems - 1;             // idle lanes repeat the last item
    const long long npl = rr::lane_pixels(a.npix_pad);
    // the pack is in item order: the workgroup's 64 items are pack group ti
    const double *group = rlane + (long long)ti * npl * (kSlots64 * 64);
    // row of pixel p - 5 of this redshift: wave-uniform (scalar loads)
    const double *Tz = tcache + ((long long)(a.item_z0[0] + iz) * a.cache_rows + (a.halo - 5)) * NB;

    constexpr int NM = NB * (NB + 1) / 2;
    double m[NM];
    double y[NB];
#pragma unroll
    for (int i = 0; i < NM; ++i) m[i] = 0.0;
#pragma unroll
    for (int i = 0; i < NB; ++i) y[i] = 0.0;

    const int nblk = (int)(npl / kF64Block);
    Stage::load(group, 0, lds[0], wave, lane);
    // one loop over pixels (the block turnover inside it): a nested block loop
    // made the register allocator rotate the Gram sums through copies
    const double *L = lds[0] + lane;
    for (int p = 0, s = 0, j = 0; p < (int)npl; ++p, ++s) {
        if (s == kF64Block) s = 0;
        if (s == 0) {
            asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
            __syncthreads();
            if (j + 1 < nblk) Stage::load(group, j + 1, lds[(j + 1) & 1], wave, lane);
            L = lds[j & 1] + lane;
            ++j;
        }
        {
            const double *Lp = L + s * (kSlots64 * 64);
            double r[11];
#pragma unroll
            for (int k = 0; k < 11; ++k) r[k] = Lp[k * 64];
            const double w = Lp[11 * 64], wf = Lp[12 * 64];
            const double *row = Tz + (long long)p * NB;
            double A[NB];
#pragma unroll
            for (int b = 0; b < NB; ++b) A[b] = r[0] * row[b];
#pragma unroll
            for (int k = 1; k < 11; ++k)
#pragma unroll
                for (int b = 0; b < NB; ++b) A[b] = fma(r[k], row[k * NB + b], A[b]);
            int idx = 0;
#pragma unroll
            for (int b1 = 0; b1 < NB; ++b1) {
                const double wa = w * A[b1];
                y[b1] = fma(A[b1], wf, y[b1]);
#pragma unroll
                for (int b2 = b1; b2 < NB; ++b2) {
                    m[idx] = fma(wa, A[b2], m[idx]);
                    ++idx;
                }
            }
        }
    }

    if (!live) return;
    const long long orow = a.item_out[item] + iz;
    double *ne = reinterpret_cast<double *>(static_cast<char *>(a.work) + ne_off);
    const long long R = a.total_rows;
    int idx = 0;
#pragma unroll
    for (int b1 = 0; b1 < NB; ++b1) {
#pragma unroll
        for (int b2 = b1; b2 < NB; ++b2) {
            ne[rr::ne_index(b1, b2, NB, false) * R + orow] = m[idx];
            ++idx;
        }
        ne[rr::ne_index(b1, NB, NB, false) * R + orow] = y[b1];
    }
}

// ---------------------------------------------------------------------------
// Explicit residual chi2 at the coefficients already in zcoeff: a workgroup
// = NW waves x ZW redshifts each, of the same 64 items.  The model u of each
// redshift is kept in an 11-pixel window (register slot (q + 5) % 11 holds
// u_q), so every template row is read and every u computed once.  Rows whose
// solve failed (status != OK) or whose target has no weight keep the finish
// stage's HUGE_CHI2.
// ---------------------------------------------------------------------------

// Template-row modes (TM):
//   kRowsScalar: scalar loads into SGPRs, the FMAs' scalar operand;
//   kRowsLds:    staged in LDS with the lane data (global_load_lds_dword, 4
//                bytes per lane), read back as wave-uniform LDS loads;
//   kRowsDpp:    a block's rows of each redshift in ceil(PB NB / 16) VGPRs,
//                float f of the block in lane f % 16 of every 16-lane row of
//                VGPR f / 16 (vector loads, one block ahead), and fed to the
//                model FMAs through the DPP row broadcast (row_newbcast:
//                every lane reads lane f % 16 of its own row), so the pixel
//                loop issues no scalar loads at all.  Scalar and LDS loads
//                share one counter that scalar loads decrement out of order,
//                so a wave waiting for its LDS data also waits for every
//                scalar load in flight; vector loads count separately.
constexpr int kRowsScalar = 0, kRowsLds = 1, kRowsDpp = 2;

// u = t[lane F of this row] * c  (the compiler folds the move into v_mul_f32_dpp)
template <int F>
__device__ __forceinline__ float bcast_mul(float t, float c) {
    return __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, t), 0x150 + F,
                                                              0xF, 0xF, true)) * c;
}

// u += t[lane F of this row] * c as one v_fmac_f32_dpp (the compiler does not
// fold a DPP move into a tied-accumulator FMA); its hazard recognizer treats
// the asm conservatively and pads VGPR write -> DPP read distances itself
template <int F>
__device__ __forceinline__ void bcast_fmac(float &u, float t, float c) {
    asm("v_fmac_f32_dpp %0, %1, %2 row_newbcast:%3 row_mask:0xf bank_mask:0xf"
        : "+v"(u)
        : "v"(t), "v"(c), "n"(F));
}

// model u = sum_b T[row][b] c_b of block row ST, its floats ST NB .. ST NB + NB - 1
template <int NB, int TV, int ST, int... B>
__device__ __forceinline__ float model_dpp(const float (&t)[TV], const float (&c)[NB],
                                           std::integer_sequence<int, B...>) {
    float u = bcast_mul<(ST * NB) % 16>(t[(ST * NB) / 16], c[0]);
    (bcast_fmac<(ST * NB + B + 1) % 16>(u, t[(ST * NB + B + 1) / 16], c[B + 1]), ...);
    return u;
}

// one pixel (block row ST) of the residual pass for ZW redshifts
template <int NB, int ZW, int TV, int ST>
__device__ __forceinline__ void resid_pixel_dpp(const float *L, const float (&t)[ZW][TV],
                                                const float (&c)[ZW][NB], float (&U)[ZW][11],
                                                float (&part)[ZW]) {
    const float *Lp = L + ST * (kSlots32 * 64);
    float r[11];
#pragma unroll
    for (int k = 0; k < 11; ++k) r[k] = Lp[k * 64];
    const float sf = Lp[11 * 64];
#pragma unroll
    for (int s = 0; s < ZW; ++s) {
        U[s][(ST + 10) % 11] = model_dpp<NB, TV, ST>(t[s], c[s], std::make_integer_sequence<int, NB - 1>{});
        float res = sf;
#pragma unroll
        for (int k = 0; k < 11; ++k) res = __builtin_fmaf(-r[k], U[s][(ST + k) % 11], res);
        part[s] = __builtin_fmaf(res, res, part[s]);
    }
}

template <int NB, int ZW, int TV, int... ST>
__device__ __forceinline__ void resid_block_dpp(const float *L, const float (&t)[ZW][TV],
                                                const float (&c)[ZW][NB], float (&U)[ZW][11],
                                                float (&part)[ZW], std::integer_sequence<int, ST...>) {
    (resid_pixel_dpp<NB, ZW, TV, ST>(L, t, c, U, part), ...);
}

template <int NB, int ZW, int NW, int WPE, int PB, int TM>
__global__ __launch_bounds__(64 * NW, WPE) void zscan_resid_kernel(rr::ZArgs a, rr::TileMap map,
                                                               const float *__restrict__ T32,
                                                               const float *__restrict__ rlane) {
    static_assert(PB % 11 == 0, "the model window turns over every 11 pixels");
    constexpr bool TL = TM == kRowsLds;
    using Stage = LaneStage<float, kSlots32, PB, NW>;
    __shared__ float lds[2][Stage::kElems];
    constexpr int TROW = PB * NB;                          // floats of one redshift's block rows
    constexpr int TINST = TL ? (NW * ZW * TROW + 63) / 64 : 0;
    constexpr int TV = TM == kRowsDpp ? (TROW + 15) / 16 : 1;
    __shared__ float tl[TL ? 2 : 1][TL ? TINST * 64 : 1];
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    int ti, tz;
    if (!rr::map_tile(map, blockIdx.x, ti, tz)) return;
    const int nzi = a.item_nz[0];
    const int zb = (tz * NW + wave) * ZW;
    int item = ti * 64 + lane;
    const bool live = item < a.nitems;
    if (!live) item = a.nitems - 1;
    const long long npl = rr::lane_pixels(a.npix_pad);
    const float *group = rlane + (long long)ti * npl * (kSlots32 * 64);
    const long long rows = a.cache_rows;
    const long long orow0 = a.item_out[item];

    float c[ZW][NB];
    const float *Tz[ZW];
#pragma unroll
    for (int s = 0; s < ZW; ++s) {
        const int iz = zb + s < nzi ? zb + s : nzi - 1;
#pragma unroll
        for (int b = 0; b < NB; ++b) c[s][b] = (float)a.zcoeff[(orow0 + iz) * NB + b];
        Tz[s] = T32 + ((long long)(a.item_z0[0] + iz) * rows + (a.halo - 5)) * NB;
    }
    auto model = [&](int s, long long q5) {   // u_q with q5 = q + 5 (row offset)
        const float *row = Tz[s] + q5 * NB;
        float u = row[0] * c[s][0];
#pragma unroll
        for (int b = 1; b < NB; ++b) u = __builtin_fmaf(row[b], c[s][b], u);
        return u;
    };
    float U[ZW][11];
#pragma unroll
    for (int s = 0; s < ZW; ++s)
#pragma unroll
        for (int j = 0; j < 10; ++j) U[s][j] = model(s, j);      // u_{-5} .. u_4
    double acc[ZW];
#pragma unroll
    for (int s = 0; s < ZW; ++s) acc[s] = 0.0;

    // the template-row copy: lane slot e of the block -> redshift e / TROW of
    // the workgroup, float e % TROW of its rows [j PB + 10, j PB + PB + 10)
    // (row offsets from pixel -5); slots past the last redshift repeat it
    auto load_rows = [&](int j, float *buf) {
        const int zw0 = tz * NW * ZW;
        for (int i = wave; i < TINST; i += NW) {
            int e = i * 64 + lane;
            if (e >= NW * ZW * TROW) e = NW * ZW * TROW - 1;
            const int zi = e / TROW, off = e - zi * TROW;
            const int iz = zw0 + zi < nzi ? zw0 + zi : nzi - 1;
            const float *src = T32 + ((long long)(a.item_z0[0] + iz) * rows + (a.halo - 5)) * NB +
                               (long long)(j * PB + 10) * NB + off;
            __builtin_amdgcn_global_load_lds((const void *)src,
                                             (void __attribute__((address_space(3))) *)(buf + i * 64), 4, 0, 0);
        }
    };

    // kRowsDpp: block j's rows of redshift s, float f at lane f % 16 of
    // tv[s][f / 16] (the 4 rows of 16 lanes load the same 64 bytes); floats
    // past the block repeat its last one
    const int l16 = lane & 15;
    auto load_tv = [&](int j, float (&dst)[ZW][TV]) {
#pragma unroll
        for (int s = 0; s < ZW; ++s) {
            const float *base = Tz[s] + (long long)(j * PB + 10) * NB;
#pragma unroll
            for (int v = 0; v < TV; ++v) {
                const int f = v * 16 + l16;
                dst[s][v] = base[f < TROW ? f : TROW - 1];
            }
        }
    };
    float tcur[ZW][TV], tnext[ZW][TV];

    const int nblk = (int)(npl / PB);               // the pack holds a multiple of 22 pixels
    Stage::load(group, 0, lds[0], wave, lane);
    if constexpr (TL) load_rows(0, tl[0]);
    if constexpr (TM == kRowsDpp) load_tv(0, tcur);
    for (int j = 0; j < nblk; ++j) {
        asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
        __syncthreads();
        if (j + 1 < nblk) {
            Stage::load(group, j + 1, lds[(j + 1) & 1], wave, lane);
            if constexpr (TL) load_rows(j + 1, tl[(j + 1) & 1]);
            if constexpr (TM == kRowsDpp) load_tv(j + 1, tnext);
        }
        const float *L = lds[j & 1] + lane;
        const float *TLb = tl[TL ? (j & 1) : 0] + (wave * ZW) * TROW;
        const long long p0 = (long long)j * PB;
        float part[ZW];
#pragma unroll
        for (int s = 0; s < ZW; ++s) part[s] = 0.f;
        if constexpr (TM == kRowsDpp) {
            resid_block_dpp<NB, ZW, TV>(L, tcur, c, U, part, std::make_integer_sequence<int, PB>{});
#pragma unroll
            for (int s = 0; s < ZW; ++s) {
                acc[s] += (double)part[s];
#pragma unroll
                for (int v = 0; v < TV; ++v) tcur[s][v] = tnext[s][v];
            }
            continue;
        }
#pragma unroll
        for (int st = 0; st < PB; ++st) {
            const long long p = p0 + st;
            const float *Lp = L + st * (kSlots32 * 64);
            float r[11];
#pragma unroll
            for (int k = 0; k < 11; ++k) r[k] = Lp[k * 64];
            const float sf = Lp[11 * 64];
#pragma unroll
            for (int s = 0; s < ZW; ++s) {
                if constexpr (TL) {
                    const float *row = TLb + s * TROW + st * NB;      // row of u_{p+5}
                    float u = row[0] * c[s][0];
#pragma unroll
                    for (int b = 1; b < NB; ++b) u = __builtin_fmaf(row[b], c[s][b], u);
                    U[s][(st + 10) % 11] = u;
                } else {
                    U[s][(st + 10) % 11] = model(s, p + 10);       // u_{p+5}
                }
                float res = sf;
#pragma unroll
                for (int k = 0; k < 11; ++k) res = __builtin_fmaf(-r[k], U[s][(st + k) % 11], res);
                part[s] = __builtin_fmaf(res, res, part[s]);
            }
        }
#pragma unroll
        for (int s = 0; s < ZW; ++s) acc[s] += (double)part[s];
    }

    if (!live) return;
    const double *st = static_cast<const double *>(a.work);
    if (st[2 * item + 1] == 0.0) return;                  // no weight: HUGE_CHI2 stays
#pragma unroll
    for (int s = 0; s < ZW; ++s) {
        const int iz = zb + s;
        if (iz >= nzi) continue;
        const long long orow = orow0 + iz;
        if (a.status && a.status[orow] != RR_STATUS_OK) continue;
        a.zchi2[orow] = acc[s];
    }
}

// Lane-kernel configurations (rr_zscan_tile_table(RR_PREC_F64)): waves per
// workgroup; entry 0 = default (by basis size).
constexpr int kF64Tiles[] = {0, 12, 8};
constexpr int kNumF64Tiles = sizeof(kF64Tiles) / sizeof(kF64Tiles[0]);

template <int NB, int NW>
int launch_f64r_nw(const rr::ZArgs &a, hipStream_t s) {
    auto kernel = zscan_f64r_kernel<NB, NW>;
    // bytes per pixel: one item tile reads 64 lanes x 13 doubles, one z tile NW rows of NB doubles
    const rr::TileMap map = rr::make_tile_map((a.nitems + 63) / 64, (a.max_item_nz + NW - 1) / NW,
                                              rr::resident_per_xcd(kernel, 64 * NW),
                                              64.0 * kSlots64 * 8, NW * NB * 8.0);
    const long long ne_off = rr::align256((long long)a.nitems * 2 * 8);
    hipLaunchKernelGGL(kernel, dim3((unsigned)map.blocks), dim3(64 * NW), 0, s, a, ne_off, map,
                       static_cast<const double *>(a.tcache_r), static_cast<const double *>(a.rlane));
    return rr::check_launch("rr_zscan(f64, lanes)");
}

template <int NB>
int launch_f64r(const rr::ZArgs &a, hipStream_t s) {
    // default: 3 waves per SIMD (a few spilled registers at nb 7..10 cost less
    // than dropping to 2 waves); 2 waves from nb 11 on (the spills grow)
    int nw = (a.tile > 0 && a.tile < kNumF64Tiles) ? kF64Tiles[a.tile] : (NB <= 10 ? 12 : 8);
    if (nw == 12) return launch_f64r_nw<NB, 12>(a, s);
    if (nw == 8) return launch_f64r_nw<NB, 8>(a, s);
    return launch_f64r_nw<NB, 8>(a, s);
}

// Residual-kernel configurations (rr_zscan_resid_tiles): redshifts per wave,
// waves per workgroup, waves per SIMD the registers must allow, pixels per
// LDS block; entries 11-13 stage the template rows in LDS, 16-20 broadcast
// them from VGPRs, the others read them with scalar loads; entry 0 = default
// (by basis size).
struct ResTile {
    int zw, nw, wpe, pb;
};
constexpr ResTile kResTiles[] = {{2, 16, 4, 0},  {2, 16, 4, 11}, {3, 16, 4, 11}, {2, 16, 4, 22},
                                 {3, 16, 4, 22}, {4, 16, 4, 22}, {6, 8, 2, 22},  {1, 16, 8, 11},
                                 {1, 16, 8, 22}, {1, 8, 8, 22},  {2, 8, 4, 22},  {1, 16, 8, 11},
                                 {2, 16, 4, 11}, {3, 16, 4, 11}, {2, 16, 8, 11}, {2, 12, 6, 11},
                                 {1, 16, 8, 11}, {2, 16, 4, 11}, {2, 16, 4, 22}, {3, 16, 4, 11},
                                 {2, 12, 6, 11}};
constexpr int kNumResTiles = sizeof(kResTiles) / sizeof(kResTiles[0]);

template <int NB, int ZW, int NW, int WPE, int PB, int TM = kRowsScalar>
int launch_resid_cfg(const rr::ZArgs &a, hipStream_t s) {
    auto kernel = zscan_resid_kernel<NB, ZW, NW, WPE, PB, TM>;
    const rr::TileMap map = rr::make_tile_map((a.nitems + 63) / 64,
                                              (a.max_item_nz + NW * ZW - 1) / (NW * ZW),
                                              rr::resident_per_xcd(kernel, 64 * NW), 64.0 * kSlots32 * 4,
                                              NW * ZW * NB * 4.0);
    hipLaunchKernelGGL(kernel, dim3((unsigned)map.blocks), dim3(64 * NW), 0, s, a, map,
                       static_cast<const float *>(a.tcache_r), static_cast<const float *>(a.rlane));
    return rr::check_launch("rr_zscan(residual)");
}

template <int NB>
int launch_resid(const rr::ZArgs &a, hipStream_t s) {
    switch (a.resid_tile) {
        case 0:   // default: 2 redshifts per wave; small bases in 22-pixel blocks with
                  // scalar template rows, larger ones in 11-pixel blocks with the rows
                  // broadcast from VGPRs (the scalar loads of nb >= 7 rows stall the
                  // LDS waits behind them)
            if constexpr (NB <= 6) return launch_resid_cfg<NB, 2, 16, 4, 22>(a, s);
            else return launch_resid_cfg<NB, 2, 16, 4, 11, kRowsDpp>(a, s);
        case 1: return launch_resid_cfg<NB, 2, 16, 4, 11>(a, s);
        case 2: return launch_resid_cfg<NB, 3, 16, 4, 11>(a, s);
        case 3: return launch_resid_cfg<NB, 2, 16, 4, 22>(a, s);
        case 4: return launch_resid_cfg<NB, 3, 16, 4, 22>(a, s);
        case 5: return launch_resid_cfg<NB, 4, 16, 4, 22>(a, s);
        case 6: return launch_resid_cfg<NB, 6, 8, 2, 22>(a, s);
        case 7: return launch_resid_cfg<NB, 1, 16, 8, 11>(a, s);
        case 8: return launch_resid_cfg<NB, 1, 16, 8, 22>(a, s);
        case 9: return launch_resid_cfg<NB, 1, 8, 8, 22>(a, s);
        case 10: return launch_resid_cfg<NB, 2, 8, 4, 22>(a, s);
        case 11: return launch_resid_cfg<NB, 1, 16, 8, 11, kRowsLds>(a, s);
        case 12: return launch_resid_cfg<NB, 2, 16, 4, 11, kRowsLds>(a, s);
        case 13: return launch_resid_cfg<NB, 3, 16, 4, 11, kRowsLds>(a, s);
        case 14: return launch_resid_cfg<NB, 2, 16, 8, 11>(a, s);
        case 15: return launch_resid_cfg<NB, 2, 12, 6, 11>(a, s);
        case 16: return launch_resid_cfg<NB, 1, 16, 8, 11, kRowsDpp>(a, s);
        case 17: return launch_resid_cfg<NB, 2, 16, 4, 11, kRowsDpp>(a, s);
        case 18: return launch_resid_cfg<NB, 2, 16, 4, 22, kRowsDpp>(a, s);
        case 19: return launch_resid_cfg<NB, 3, 16, 4, 11, kRowsDpp>(a, s);
        case 20: return launch_resid_cfg<NB, 2, 12, 6, 11, kRowsDpp>(a, s);
        default: return rr::fail(RR_EINVAL, "rr_zscan(residual): resid_tile %d outside 0..%d", a.resid_tile,
                                 kNumResTiles - 1);
    }
}

}  // namespace

namespace rr {

int zscan_f64r_tiles(int *cfg, int max) {
    const int n = kNumF64Tiles < max ? kNumF64Tiles : max;
    for (int i = 0; i < n; ++i) {
        cfg[4 * i] = 1;                                  // one 64-item tile
        cfg[4 * i + 1] = kF64Tiles[i] ? kF64Tiles[i] : 8;   // waves (redshifts) per workgroup
        cfg[4 * i + 2] = 64;                             // items per wave
        cfg[4 * i + 3] = 1;                              // redshifts per wave
    }
    return kNumF64Tiles;
}

int zscan_f64r(const ZArgs &a, hipStream_t s) {
    if (a.tile < 0 || a.tile >= kNumF64Tiles)
        return fail(RR_EINVAL, "rr_zscan(f64, lanes): tile %d outside 0..%d", a.tile, kNumF64Tiles - 1);
    switch (a.nb) {
        case 1: return launch_f64r<1>(a, s);
        case 2: return launch_f64r<2>(a, s);
        case 3: return launch_f64r<3>(a, s);
        case 4: return launch_f64r<4>(a, s);
        case 5: return launch_f64r<5>(a, s);
        case 6: return launch_f64r<6>(a, s);
        case 7: return launch_f64r<7>(a, s);
        case 8: return launch_f64r<8>(a, s);
        case 9: return launch_f64r<9>(a, s);
        case 10: return launch_f64r<10>(a, s);
        case 11: return launch_f64r<11>(a, s);
        case 12: return launch_f64r<12>(a, s);
        default: return fail(RR_EUNSUPPORTED, "rr_zscan(f64, lanes): nb=%d unsupported", a.nb);
    }
}

int zscan_resid_tiles(int *cfg, int max) {
    const int n = kNumResTiles < max ? kNumResTiles : max;
    for (int i = 0; i < n; ++i) {
        const ResTile &t = kResTiles[i];
        cfg[4 * i] = 1;             // one 64-item tile
        cfg[4 * i + 1] = t.nw;      // waves per workgroup
        cfg[4 * i + 2] = 64;        // items per wave
        cfg[4 * i + 3] = t.zw;      // redshifts per wave
    }
    return kNumResTiles;
}

int zscan_resid(const ZArgs &a, hipStream_t s) {
    switch (a.nb) {
        case 1: return launch_resid<1>(a, s);
        case 2: return launch_resid<2>(a, s);
        case 3: return launch_resid<3>(a, s);
        case 4: return launch_resid<4>(a, s);
        case 5: return launch_resid<5>(a, s);
        case 6: return launch_resid<6>(a, s);
        case 7: return launch_resid<7>(a, s);
        case 8: return launch_resid<8>(a, s);
        case 9: return launch_resid<9>(a, s);
        case 10: return launch_resid<10>(a, s);
        case 11: return launch_resid<11>(a, s);
        case 12: return launch_resid<12>(a, s);
        default: return fail(RR_EUNSUPPORTED, "rr_zscan(residual): nb=%d unsupported", a.nb);
    }
}

}  // namespace rr

extern "C" long long rr_pack_lanes_bytes(int nslots, int npix_pad, int precision) {
    if (nslots < 0 || npix_pad < 32 || (npix_pad % 32) != 0)
        return rr::fail(RR_EINVAL, "rr_pack_lanes_bytes: bad sizes nslots=%d npix_pad=%d", nslots, npix_pad);
    const long long groups = (nslots + 63) / 64;
    const long long npl = rr::lane_pixels(npix_pad);
    // + 1 KB: the staging copy of a group's last pixel block may run past it
    if (precision == RR_PREC_F64) return groups * npl * kSlots64 * 64 * 8 + 1024;
    if (precision == RR_PREC_F16R) return groups * npl * kSlots32 * 64 * 4 + 1024;
    return rr::fail(RR_EINVAL, "rr_pack_lanes_bytes: precision must be RR_PREC_F64 or RR_PREC_F16R");
}

extern "C" int rr_pack_lanes(const double *d_rband, const double *d_w, const double *d_f,
                             const int *d_order, int nslots, int npix_pad, int precision, void *d_out,
                             void *stream) {
    const long long need = rr_pack_lanes_bytes(nslots, npix_pad, precision);
    if (need < 0) return (int)need;
    if (nslots == 0) return RR_OK;
    if (!d_rband || !d_w || !d_f || !d_out) return rr::fail(RR_EINVAL, "rr_pack_lanes: null pointer");
    const long long npl = rr::lane_pixels(npix_pad);
    dim3 grid((unsigned)npl, (unsigned)((nslots + 63) / 64));
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (precision == RR_PREC_F64)
        hipLaunchKernelGGL(pack_lanes64_kernel, grid, dim3(64), 0, s, d_rband, d_w, d_f, d_order, nslots,
                           npix_pad, npl, static_cast<double *>(d_out));
    else
        hipLaunchKernelGGL(pack_lanes32_kernel, grid, dim3(64), 0, s, d_rband, d_w, d_f, d_order, nslots,
                           npix_pad, npl, static_cast<float *>(d_out));
    return rr::check_launch("rr_pack_lanes");
}
