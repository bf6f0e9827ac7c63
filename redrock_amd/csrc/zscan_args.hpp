// Kernel-argument form of rr_zscan_desc (passed by value to the kernels).
#pragma once
#include "common.hpp"

namespace rr {

struct ZArgs {
    int nitems, npix, npix_pad, nb, precision, halo;
    const void *tcache;
    long long cache_rows, cache_ldz;
    int cache_nz;
    double tscale;
    const double *rband, *w, *wf, *f;
    const void *rtile;
    const void *sfcol;
    const double *scales;
    const void *tcache_r;  // [z][row][b] cache of the lane-per-item kernels
    const void *rlane;     // rr_pack_lanes output
    int uniform_z;
    int tile;
    int resid_tile;        // RR_PREC_F16R: configuration of the residual kernel
    const int *item_target, *item_z0, *item_nz;
    const long long *item_out;
    const double *prior, *oiisum;
    double *zchi2, *zcoeff, *penalty;
    unsigned char *status;
    int max_item_nz;
    void *work;            // workspace: item stats then normal equations
    long long total_rows;  // rows of the output arrays (>= max item_out + nz)
    int solve;             // RR_SOLVE_PCA, RR_SOLVE_NNLS or RR_SOLVE_BVLS
    int nonneg_mask;       // RR_SOLVE_NNLS: bit j set = coefficient j >= 0
    const double *bounds;  // RR_SOLVE_BVLS: (2, nb) lower then upper bounds
    const void *tcache_u;  // RR_PREC_F16R, matrix-core residual: template operand cache
    int rtile_planes;      // R tile planes per target in rtile: 2 (hi, lo) for F16X3 / F16R
};

// Hi/lo products of the matrix-core residual's u = T c (Th.Ch, Tl.Ch, Th.Cl
// and, when it still fits the same number of 32-slot MFMAs or needs two
// anyway, Tl.Cl) and the slots per (z, row) of its template operand cache.
__host__ __device__ constexpr int resid_u_nprod(int nb) { return 4 * nb <= 32 ? 4 : (3 * nb <= 32 ? 3 : 4); }
__host__ __device__ constexpr int resid_u_slots(int nb) { return 32 * ((resid_u_nprod(nb) * nb + 31) / 32); }

// Kernel-argument form of rr_polish_desc.
struct PolishArgs {
    const double *tcache64;
    long long ldz;
    double tol;
    int max_windows;
    int nw_all;            // windows of the longest item, ceil(max_item_nz / 16)
    int *windows, *nwindows;
    int *queue;            // [0] = entries, then item * RR_POLISH_MAX_WINDOWS + slot
};

// Workspace layout (all offsets 256-byte aligned):
//   stats  double[nitems][2]            f'Wf, sum(w) of the item's target
//   ne     double[ne_count][total_rows]   (k-major, coalesced); the mixed
//          modes store their scaled MFMA sums (unscaled by zscan_finish)
// ne entry k of row r is the (i, j) element, i <= j, i < nb, of the upper
// triangle of [M | y] (F64: j <= nb; mixed: j <= nb+1, y split hi/lo in the
// two last columns), k = ne_index(i, j).
__host__ __device__ constexpr int ne_cols(int nb, bool mixed) { return mixed ? nb + 2 : nb + 1; }
__host__ __device__ constexpr int ne_count(int nb, bool mixed) {
    return nb * ne_cols(nb, mixed) - nb * (nb - 1) / 2;
}
__host__ __device__ constexpr int ne_index(int i, int j, int nb, bool mixed) {
    return i * ne_cols(nb, mixed) - i * (i - 1) / 2 + (j - i);
}
long long workspace_bytes(int nitems, int nb, bool mixed, long long total_rows);
inline long long align256(long long x) { return (x + 255) & ~255LL; }

// Redshifts packed side by side into the 16 columns of one 16-bit MFMA tile:
// ZP*nb basis columns + 2 sqrt(w) f columns <= 16 (at most 4).
__host__ __device__ constexpr int zpack(int nb) {
    return nb <= 3 ? 4 : (14 / nb >= 2 ? 14 / nb : 1);
}

int zscan_stats(const ZArgs &a, hipStream_t s);
int zscan_finish(const ZArgs &a, hipStream_t s);

int zscan_f64(const ZArgs &a, hipStream_t s);
int zscan_f64r(const ZArgs &a, hipStream_t s);
int zscan_f64r_tiles(int *cfg, int max);
int zscan_resid(const ZArgs &a, hipStream_t s);
int zscan_resid_tiles(int *cfg, int max);
int zscan_resid_mfma(const ZArgs &a, int cfg, hipStream_t s);
int zscan_resid_mfma_tiles(int *cfg, int max);
// pixels of the lane-major per-target pack (a multiple of 22, >= npix_pad)
__host__ __device__ constexpr long long lane_pixels(long long npix_pad) { return rr_lane_pixels(npix_pad); }
int zscan_mfma(const ZArgs &a, hipStream_t s);
int zscan_mfma32(const ZArgs &a, hipStream_t s);
int zscan_mfma3(const ZArgs &a, hipStream_t s);
int zscan_polish(const ZArgs &a, const PolishArgs &p, hipStream_t s);
int pack_mixed3(const double *d_rband, const double *d_w, const double *d_f, int ntarget, int npix_pad,
                void *d_rtile, double *d_scales, hipStream_t s);
int zscan_tiles3(int *cfg, int max);
int zscan_tiles16(int *cfg, int max);
int pack_mixed32(const double *d_rband, const double *d_w, const double *d_f, int ntarget,
                 int npix_pad, void *d_rtile, void *d_sfcol, double *d_scales, hipStream_t s);

}  // namespace rr
