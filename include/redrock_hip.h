/*
 * redrock_hip.h — C ABI of libredrock_hip.so, the gfx950 implementation of
 * redrock's redshift chi2 scan.
 *
 * Conventions
 *  - Every pointer argument named d_* is a device pointer (HBM) owned by the
 *    caller; the library never frees caller memory and allocates nothing.
 *  - Every entry point is stream-ordered on `stream` (a hipStream_t passed as
 *    void*; NULL = the null stream) and does not synchronise the device.
 *  - Every entry point returns 0 on success and a negative code on an
 *    argument or launch error; rr_last_error() then returns a thread-local
 *    message.  Numerical failures (a singular normal matrix) are not errors:
 *    they are reported per (target, redshift) in the status output.
 *  - Calls that share output buffers on one stream must not overlap.
 */
#ifndef REDROCK_HIP_H
#define REDROCK_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

#define RR_ABI_VERSION 14

/* Error codes. */
#define RR_OK 0
#define RR_EINVAL (-1)
#define RR_ELAUNCH (-2)
#define RR_EUNSUPPORTED (-3)

/* Output element types for rr_rebin. */
#define RR_F64 0
#define RR_F32 1
#define RR_F16 2
#define RR_BF16 3

/* Precision modes of rr_zscan. */
#define RR_PREC_F64 0   /* plain FP64: vector FP64 FMA, FP64 LU solve           */
#define RR_PREC_F16 1   /* MFMA fp16 inputs, fp32 accumulate, FP64 Cholesky     */
#define RR_PREC_BF16 2  /* MFMA bf16 inputs, fp32 accumulate, FP64 Cholesky     */
#define RR_PREC_F32 3   /* MFMA fp32 inputs, fp32 accumulate, FP64 Cholesky     */
#define RR_PREC_F16X3 4 /* MFMA fp16 hi+lo split inputs (three products per
                           term), fp32 partial sums added in FP64, FP64
                           Cholesky                                           */
#define RR_PREC_F16R 5  /* fp16 MFMA normal equations and FP64 Cholesky for the
                           coefficients, then chi2 as the explicit residual
                           sum(w (f - R T c)^2) at those coefficients in FP32
                           with FP64 partial sums (uniform items only)        */

/* Resolution band: half width 5 (11 diagonals), the DESI layout. */
#define RR_BAND_HALF 5
#define RR_BAND_NDIAG 11

/* Largest template basis supported by the scan kernels. */
#define RR_MAX_NBASIS 12

/* Per-(target, z) status codes written by rr_zscan. */
#define RR_STATUS_OK 0
#define RR_STATUS_SINGULAR 1   /* LU found an exactly zero pivot -> HUGE_CHI2 */

int rr_abi_version(void);
const char *rr_last_error(void);

/*
 * Lyman-series IGM line table, one entry per line: transmission factor
 * exp(-A (lambda_obs/line)^B) wherever lambda_obs/(1+z) < line.
 * Replaces the constants.Lyman_series tables consumed by
 * transmission_Lyman_CaluraKamble (py/redrock/igm.py:321-364).
 */
typedef struct {
    double line;
    double A;
    double B;
} rr_lyman_line;

#define RR_MAX_LYMAN_LINES 8

/*
 * rr_rebin — trapezoidal rebin of a template basis to every trial redshift,
 * with the Lyman-series transmission multiply fused in.
 *
 * Replaces rebin.trapz_rebin / _trapz_rebin_batch / batch_trapz_rebin
 * (py/redrock/rebin.py:20-96,126-209,212-451) followed by the IGM multiply
 * of DistTemplate.__init__ (py/redrock/templates.py:586-597) and of fitz
 * (py/redrock/fitz.py:233-243).
 *
 *   d_x      (nx)        rest-frame template wavelengths, increasing
 *   d_y      (nb, y_ld)  template fluxes, basis-major
 *   d_edges  (nbin+1)    observed-frame bin edges
 *   d_wave   (nbin)      bin centres for the IGM factor (may be NULL when
 *                        nlines == 0)
 *   d_z      (nz)        redshifts
 *   lines    host array of nlines rr_lyman_line (nlines <= 8; NULL if 0)
 *   d_out    element (iz, ibin, ib) is written at
 *            d_out + iz*s_z + ibin*s_pix + ib*s_b (in elements of out_type),
 *            value multiplied by `scale` before conversion.
 *
 * The caller must have checked the edge range (rebin.py:332-341); the kernel
 * clamps its sample search so that an unchecked call cannot read out of
 * bounds, but the values are then meaningless.
 */
int rr_rebin(const double *d_x, int nx, const double *d_y, int nb, long long y_ld,
             const double *d_edges, int nbin, const double *d_wave,
             const double *d_z, int nz, const rr_lyman_line *lines, int nlines,
             void *d_out, int out_type, long long s_z, long long s_pix, long long s_b,
             double scale, void *stream);

/*
 * Scan problem descriptor for rr_zscan.  One "item" is one target scanned
 * over a contiguous range of the template cache's redshift axis.
 *
 * Pixel axis: the target's cameras concatenated in spectrum order
 * (zscan.py:383-389, spectral_data zscan.py:192-209), padded to npix_pad
 * (a multiple of 32) with zero weight.  The template cache rows are the same
 * pixel axis shifted by `halo` rows of zeros.
 */
typedef struct {
    int nitems;        /* number of (target, z-range) items                      */
    int npix;          /* real pixels per target                                 */
    int npix_pad;      /* padded pixel count (multiple of 32)                    */
    int nb;            /* template basis size (<= RR_MAX_NBASIS)                 */
    int precision;     /* RR_PREC_*                                              */
    int halo;          /* zero rows before pixel 0 in the cache                  */

    /* Template cache.
     *  F64  : d_tcache is double, element (b, row, zc) at
     *         (b*cache_rows + row)*cache_ldz + zc           ([b][row][z])
     *  F16/BF16: d_tcache is 16-bit, element (zc, b, row) at
     *         (zc*nb + b)*cache_rows + row                  ([z][b][row]),
     *         followed by at least 64 zero elements
     *  F32  : as F16 with fp32 elements
     *  F16X3: fp16, element (zc, h, b, row) at
     *         ((zc*2 + h)*nb + b)*cache_rows + row  ([z][hi,lo][b][row]),
     *         h = 0: fp16(v), h = 1: fp16(v - hi), followed by at least 64
     *         zero elements
     * row = pixel + halo.  cache_nz is the extent of the z axis.            */
    const void *d_tcache;
    long long cache_rows;
    long long cache_ldz;
    int cache_nz;
    double tscale;     /* cache values = template * tscale (power of two)        */

    /* Per-target data, indexed by target slot. */
    const double *d_rband;   /* (ntarget, npix_pad, 11) R[p, p+k-5]             */
    const double *d_w;       /* (ntarget, npix_pad) ivar                        */
    const double *d_wf;      /* (ntarget, npix_pad) ivar*flux                   */
    const double *d_f;       /* (ntarget, npix_pad) flux                        */
    /* Mixed precision modes only (NULL otherwise), written by rr_pack_mixed: */
    const void *d_rtile;     /* (ntarget, npix_pad/16, 64, 8) MFMA A-operand
                                tiles of sqrt(w) R (16-bit: K slots 0, 1 of
                                each row hold sqrt(w) f split hi + lo; F16X3
                                and F16R: the F16X3 tiles, (ntarget, 2,
                                npix_pad/16, 64, 8), hi then lo plane; F16R's
                                fp16 normal equations read the hi plane)        */
    const void *d_sfcol;     /* RR_PREC_F32 only (NULL allowed otherwise):
                                (ntarget, npix_pad/32, 4, 8) fp32 sqrt(w) f in
                                MFMA operand order                              */
    const double *d_scales;  /* (ntarget, 2) power-of-two scales of R and f     */
    /* Lane-per-item kernels (NULL = not used):
     *  d_tcache_r: template cache [z][row][b] (basis fastest), element
     *         (zc, row, b) at (zc*cache_rows + row)*nb + b; double for F64,
     *         float for F16R (values not scaled by tscale).  Needs
     *         cache_rows >= halo + rr_lane_pixels(npix_pad) + 5.
     *  d_rlane: rr_pack_lanes output in item order (slot i = the target
     *         of item i), same precision (F64 or F16R).
     * F64 with both set and uniform_z: the scan runs one item per lane
     * (template values through scalar loads); otherwise zscan_f64 reads
     * d_tcache.  F16R needs both (and the F16 cache and tiles above).      */
    const void *d_tcache_r;
    const void *d_rlane;
    int uniform_z;           /* 1 if every item has the same z0 and nz: lets a
                                wavefront share template tiles across items     */
    int tile;                /* mixed: tile configuration, 0 = default (see
                                rr_zscan_tiles); F64 with the lane inputs: the
                                lane kernel's (rr_zscan_tile_table)             */
    int resid_tile;          /* F16R: configuration of the explicit-residual
                                kernel, 0 = default (rr_zscan_resid_tiles)      */

    /* Work list, one entry per item. */
    const int *d_item_target;  /* target slot                                   */
    const int *d_item_z0;      /* first cache z index                           */
    const int *d_item_nz;      /* number of redshifts                           */
    const long long *d_item_out; /* offset of the item's first row in outputs   */

    /* Optional inputs (NULL = absent). */
    const double *d_prior;     /* (nb, nb) added to M (zscan.py:711-712,761-762) */
    const double *d_oiisum;    /* (nb) OIItemplate summed over wavelength        */

    /* Outputs, row r = item_out + iz. */
    double *d_zchi2;           /* (rows)                                         */
    double *d_zcoeff;          /* (rows, nb)                                     */
    double *d_penalty;         /* (rows) or NULL                                 */
    unsigned char *d_status;   /* (rows) or NULL                                 */
    int max_item_nz;           /* largest item nz (sizes the grid)               */
    long long total_rows;      /* number of output rows (> every item_out + nz)  */

    /* Caller-provided device workspace of at least rr_zscan_workspace_bytes()
     * bytes (per-item statistics and the per-row normal equations).         */
    void *d_work;
    long long work_bytes;

    /* Which of the three launches to enqueue: RR_STAGE_* bits, 0 = all.
     * The stages communicate through the workspace, so a partial call
     * relies on an earlier call having run the other stages on it.        */
    int stages;
    /* Coefficient solve per (item, redshift), the template's method
     * (solve_matrices, zscan.py:1107-1205):
     *   RR_SOLVE_PCA : M c = y (LU in FP64, Cholesky in the mixed modes)
     *   RR_SOLVE_NNLS: min |M c - y|^2 over c >= 0, the reference's NMF / NNLS
     *                  templates (scipy.optimize.nnls(M, y), zscan.py:1153-1172),
     *                  Lawson-Hanson active sets; no solution within 3 nb
     *                  iterations -> HUGE_CHI2, status RR_STATUS_SINGULAR
     *                  (the reference's LinAlgError, zscan.py:769-771)
     *   RR_SOLVE_BVLS: min |M c - y|^2 over d_bounds[0][j] <= c_j <=
     *                  d_bounds[1][j] (scipy lsq_linear(M, y, bounds, 'bvls'),
     *                  the reference's BVLS branch, zscan.py:1175-1201, for
     *                  any bounds), bounded-variable active sets; the same
     *                  failure convention                                      */
    int solve;
    /* RR_SOLVE_NNLS: bit j set = coefficient j >= 0, the others free.  Every
     * bit for NMF / NNLS templates; the archetype bounds of the reference's
     * BVLS calls (zscan.py:289-294: archetype coefficients >= 0, Legendre
     * terms free) otherwise.  Ignored by RR_SOLVE_PCA.                       */
    int nonneg_mask;
    /* RR_SOLVE_BVLS: (2, nb) lower bounds then upper bounds, +-inf allowed,
     * lower < upper for every coefficient; NULL otherwise.                  */
    const double *d_bounds;

    /* F16R explicit residual on the matrix cores (the configurations of
     * kind 2 in rr_zscan_resid_tiles, and the default when set; NULL = not
     * used).  Needs uniform_z and halo == 8; reads d_rtile (F16X3 tiles).
     *  d_tcache_u: fp16 template operand, element (zc, row, s) at
     *         (zc*cache_rows + row)*KS + s, KS = rr_resid_u_slots(nb, &np):
     *         s < nb: hi(T[s]); nb <= s < 2nb: lo(T[s-nb]); 2nb <= s < 3nb:
     *         hi(T[s-2nb]); 3nb <= s < 4nb when np == 4: lo(T[s-3nb]); other
     *         slots 0.  T = template value * tscale, hi = fp16(T),
     *         lo = fp16(T - hi), round to nearest even.                     */
    const void *d_tcache_u;
} rr_zscan_desc;
#define RR_SOLVE_PCA 0
#define RR_SOLVE_NNLS 1
#define RR_SOLVE_BVLS 2

#define RR_STAGE_STATS 1    /* per-item f'Wf and sum(w)                     */
#define RR_STAGE_MAIN 2     /* normal equations (the scan kernel)           */
#define RR_STAGE_FINISH 4   /* FP64 solve, chi2, penalty, status            */
#define RR_STAGE_RESID 8    /* F16R: explicit residual chi2 (after FINISH)   */

/* Workspace size rr_zscan needs for this descriptor (nitems, nb, precision,
 * total_rows are read).  Returns a negative code on a bad descriptor. */
long long rr_zscan_workspace_bytes(const rr_zscan_desc *desc);

/*
 * rr_zscan — chi2 of every (item, redshift): R-apply, weighted normal
 * equations, solve, chi2 and the [OII] penalty, fused in one launch.
 *
 * Replaces zscan.calc_zchi2_batch (py/redrock/zscan.py:626-781) including
 * batch_dot_product_sparse (:45-74,:392-444), batch_dot_product_3d3d
 * (:76-133,:551-617), solve_matrices PCA (:1107-1152), batch_dot_product_3d2d
 * (:135-162,:502-541) and calc_negOII_penalty (:783-800); with many items it
 * also replaces the per-target loop of zscan.calc_zchi2 (:853-887).
 *
 * chi2 = f'Wf - 2 c'y + c'Mc with the unregularised M (compensated FP64
 * sum): the reference's explicit residual sum(w (f - A c)^2) (:777-780)
 * expanded, exact with a prior too.  Singular M (exactly zero LU
 * pivot, np.linalg.solve's LinAlgError) gives chi2 = HUGE_CHI2, zero coeffs
 * and status RR_STATUS_SINGULAR (zscan.py:769-771).  Zero total weight gives
 * HUGE_CHI2 and zero coeffs with status OK (zscan.py:657-660).
 *
 * RR_PREC_F16R: the fp16 normal equations (as RR_PREC_F16) give the
 * coefficients; stage RR_STAGE_RESID then overwrites chi2 with the explicit
 * residual sum(w (f - R T c)^2) (:777-780 itself, no expansion) at those
 * coefficients, whose error is second order in the coefficients' error.
 *
 * Mixed modes need halo == 8 and the rr_pack_mixed outputs; F64 needs
 * cache_ldz >= cache_nz, or the lane-per-item inputs (d_tcache_r, d_rlane)
 * with uniform_z.
 */
int rr_zscan(const rr_zscan_desc *desc, void *stream);

/*
 * rr_zscan_noalloc — rr_zscan with no caller workspace: the statistics and
 * normal equations live in a static device array of RR_NOALLOC_SCRATCH_BYTES
 * held by every GPU the library runs on until the process ends.  No
 * counterpart in the reference; it is for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops).
 * desc->d_work and desc->work_bytes are ignored; everything else is as for
 * rr_zscan, which it runs.  A scan whose rr_zscan_workspace_bytes() exceeds
 * the array returns RR_EINVAL.  Calls that overlap on one GPU (other
 * streams, other host threads) share the array and race.  The array is not
 * rr_zscan_polish's workspace: polish scans made by rr_zscan.
 */
#define RR_NOALLOC_SCRATCH_BYTES (64LL << 20)
int rr_zscan_noalloc(const rr_zscan_desc *desc, void *stream);

/*
 * rr_pack_lanes — per-item data of the lane-per-item kernels, lane-major:
 * value v of pixel p of slot i at ((i/64)*L + p)*S*64 + v*64 + i%64 with
 * L = rr_lane_pixels(npix_pad) (pixels from npix_pad on, and slots from
 * nslots on, are zero).  Slot i holds target d_order[i] (d_order NULL:
 * target i); rr_zscan reads slot i for item i, so pack the plan's
 * item_target as d_order (or NULL when item i is target i).
 *   RR_PREC_F64 : double, S = 13: R[p][0..10] (R[p, p+k-5]), w, w*f
 *   RR_PREC_F16R: float,  S = 12: sqrt(w) R[p][0..10], sqrt(w) f
 * d_out holds rr_pack_lanes_bytes(nslots, ...) bytes (the packed slots plus
 * 1 KB the kernels' staging copies may read past the end).  Inputs as in
 * rr_zscan_desc.  Replaces the per-call CSR upload of zscan.py:428-430
 * (done once per target batch here).
 */
#define rr_lane_pixels(npix_pad) ((((npix_pad) + 21) / 22) * 22)
long long rr_pack_lanes_bytes(int nslots, int npix_pad, int precision);
int rr_pack_lanes(const double *d_rband, const double *d_w, const double *d_f, const int *d_order,
                  int nslots, int npix_pad, int precision, void *d_out, void *stream);

/*
 * rr_pack_mixed — per-target preparation for the mixed precision modes:
 * packs sqrt(w) R into MFMA A-operand tiles (tile q covers output pixels
 * [16q, 16q+16) and input pixels [16q-8, 16q+24)) and chooses power-of-two
 * scales that keep R and sqrt(w) f in range.  Inputs are the (ntarget,
 * npix_pad[, 11]) arrays of rr_zscan_desc; precision is RR_PREC_F16 or
 * RR_PREC_BF16 (16-bit tiles; sqrt(w) f split hi + lo into the two K slots of
 * each tile row that lie outside the band; d_sfcol is not used and may be
 * NULL), RR_PREC_F16X3 (hi and lo fp16 tiles, twice the 16-bit size, in
 * planes (ntarget, 2, npix_pad/16, 64, 8); the hi plane equals the
 * RR_PREC_F16 tiles value for value; also the tiles of RR_PREC_F16R;
 * d_sfcol not used) or RR_PREC_F32 (fp32 tiles of twice the 16-bit size,
 * rtile (ntarget, npix_pad/16, 64, 8) floats, and sqrt(w) f in d_sfcol
 * (ntarget, npix_pad/32, 4, 8) floats).
 * Replaces the per-call CSR upload
 * of zscan.py:428-430 (done once per target here).
 */
/*
 * rr_zscan_tiles — the tile configurations of the mixed-precision scan
 * kernel.  Entry i is cfg[4i..4i+3] = (WA, WB, TW, ZW): a workgroup of four
 * wavefronts arranged WA x WB covers WA*TW items x WB*ZW redshifts, each
 * wavefront TW items x ZW redshifts.  Entry 0 is the default.  Host array of
 * 4*max ints; returns the number of configurations.  Used by the autotuner
 * to pick rr_zscan_desc.tile per (nbasis, pixel count, precision).
 */
int rr_zscan_tiles(int *cfg, int max);

/*
 * rr_zscan_tile_table — as rr_zscan_tiles for the kernel of `precision`
 * (RR_PREC_F16 / _BF16 / _F16R: the table of rr_zscan_tiles; RR_PREC_F16X3:
 * its own table; RR_PREC_F64: the lane-per-item kernel's, entry i =
 * (1, waves per workgroup, 64, 1), each wave one redshift of the same 64
 * items; other modes have a single configuration and return 0).
 */
int rr_zscan_tile_table(int precision, int *cfg, int max);

/*
 * rr_zscan_resid_tiles — configurations of the explicit-residual stage of
 * RR_PREC_F16R (rr_zscan_desc.resid_tile): first the FP32 lane kernel's,
 * entry i = (1, waves per workgroup, 64 items per wave, redshifts per wave),
 * then the matrix-core kernel's, entry i = (2, waves per workgroup, 16 items
 * per workgroup, redshifts per workgroup).  Entry 0 is the default: the
 * matrix-core kernel when d_tcache_u is set, else
 * the lane kernel.  Returns the number of entries.
 */
int rr_zscan_resid_tiles(int *cfg, int max);

/*
 * rr_resid_u_slots — row length KS (fp16 elements per (z, row)) of the
 * d_tcache_u layout for a basis of nb vectors, and in *nprod the number of
 * hi/lo products it holds (3 or 4).  Returns a negative code for nb outside
 * 1..RR_MAX_NBASIS.  No reference counterpart: the layout of this build's
 * matrix-core residual kernel.
 */
int rr_resid_u_slots(int nb, int *nprod);

int rr_pack_mixed(const double *d_rband, const double *d_w, const double *d_f, int ntarget,
                  int npix_pad, int precision, void *d_rtile, void *d_sfcol,
                  double *d_scales, void *stream);

/*
 * rr_zscan_polish — FP64 re-evaluation of the rows of a mixed-precision scan
 * near each item's minimum, so that the item's best redshift is the FP64
 * scan's.  No counterpart in the reference (its scan is FP64 throughout);
 * it stands between the mixed scan and the consumers of zchi2 (find_minima
 * in fitz.py:28-54, the best-redshift choice of zfind.py:458-500).
 *
 * For every item of `scan` (a descriptor that rr_zscan has run in full, in a
 * mixed precision): the windows of RR_POLISH_WINDOW consecutive redshifts
 * (item z index 4k .. 4k+3) holding a chi2, or chi2 + penalty, within
 * tol * max(f'Wf, |minimum|) of the item's minimum of that quantity (f'Wf of
 * the item's target, from the scan's statistics stage) are re-evaluated in
 * FP64 (R-apply, normal equations, LU solve, compensated chi2, penalty,
 * status) from the FP64 template cache, and their rows of d_zchi2, d_zcoeff,
 * d_penalty and d_status are overwritten.  If more than max_windows windows
 * qualify, the tolerance is halved until they fit (the windows nearest the
 * minimum are kept).  An item with a single row under the full thresholds
 * gets no window (that row is the FP64 argmin already; its chi2 keeps the
 * mixed value).  Rows with a status other than RR_STATUS_OK take no part,
 * and an item whose target has no weight (HUGE_CHI2 everywhere) is left
 * alone; a real chi2 above HUGE_CHI2 (a bright target) is polished like any
 * other.  A mixed scan's chi2 error scales with f'Wf (the sum chi2 cancels
 * against): with |chi2_mixed - chi2_f64| <= e f'Wf and tol >= 2e the argmin
 * of every polished item equals the FP64 scan's (ties within FP64 rounding
 * aside).
 *
 * scan->precision RR_PREC_F64: nothing to do, returns RR_OK.  scan->solve
 * must be RR_SOLVE_PCA (the re-evaluation solves M c = y): RR_EUNSUPPORTED
 * otherwise.
 * Reads the scan's workspace statistics (stage RR_STAGE_STATS); the FP64
 * cache has the scan's rows, halo and pixel layout.  max_item_nz <= 16384.
 */
#define RR_POLISH_WINDOW 4
#define RR_POLISH_MAX_WINDOWS 64
typedef struct {
    const double *d_tcache64;  /* FP64 cache [b][row][z] of the scanned template  */
    long long cache_ldz;       /* >= scan->cache_nz                                 */
    double tol;                /* relative threshold, >= 0                          */
    int max_windows;           /* 1 .. RR_POLISH_MAX_WINDOWS per item               */
    int *d_windows;            /* (nitems, max_windows) out: polished window k's    */
    int *d_nwindows;           /* (nitems) out: windows polished per item           */
    int *d_queue;              /* (1 + nitems * max_windows) workspace: count and
                                  the (item, window) work list                      */
} rr_polish_desc;
int rr_zscan_polish(const rr_zscan_desc *scan, const rr_polish_desc *polish, void *stream);

/*
 * Adjoint problem descriptor for rr_zscan_grad.  One item = one target at one
 * redshift with its best-fit coefficients (from rr_zscan).  The pixel axis is
 * the one of rr_zscan_desc; each real pixel carries its observed-frame bin
 * edges and centre.
 */
typedef struct {
    int nitems;
    int npix, npix_pad, nb;
    const double *d_x;          /* (nx) rest-frame template wavelengths            */
    int nx;
    const double *d_y;          /* (nb, nx) template fluxes                         */
    const double *d_pe_lo;      /* (npix) lower bin edge of every pixel             */
    const double *d_pe_hi;      /* (npix) upper bin edge                            */
    const double *d_pwave;      /* (npix) bin centre (IGM); NULL if nlines == 0     */
    const rr_lyman_line *lines; /* host array, IGM lines of the template            */
    int nlines;
    const double *d_rband, *d_w, *d_f;   /* per target, as in rr_zscan_desc         */
    const int *d_item_target;   /* (nitems)                                         */
    const double *d_item_z;     /* (nitems) redshift                                */
    const double *d_item_coeff; /* (nitems, nb) coefficients at that redshift       */
    double *d_dchi2_dz;         /* (nitems) out                                     */
    double *d_chi2;             /* (nitems) out, explicit FP64 residual, or NULL    */
    double *d_dchi2_dy;         /* (nb, nx) accumulated (+=; items summed in item
                                   order, no atomics), or NULL                       */
    const double *d_item_scale; /* (nitems) weight of each item's d_dchi2_dy term,
                                   or NULL for 1 (a loss sum_i s_i chi2_i)          */
    void *d_work;               /* workspace of rr_grad_workspace_bytes() bytes
                                   (only with d_dchi2_dy)                            */
    long long work_bytes;
} rr_grad_desc;

/* Workspace rr_zscan_grad needs for this descriptor: 0 without d_dchi2_dy,
 * else per item one double per pixel and one per template sample, and
 * min(nitems, 64) x nbasis doubles per template sample (the partial sums of
 * the fixed-order reduction over items). */
long long rr_grad_workspace_bytes(const rr_grad_desc *desc);

/*
 * rr_zscan_grad — gradient of chi2(z) = min_c |W^1/2 (f - R T(z) c)|^2 with
 * respect to z and to the template fluxes, at the given coefficients (the
 * envelope theorem makes the coefficient derivative vanish).  The rebin is
 * recomputed from the template, not read from a cache.  No reference
 * counterpart (redrock has no gradients); it serves the autograd Function and
 * the Newton refinement that stand beside fitz's parabola fit
 * (py/redrock/fitz.py:259-261).  The chi2 output is the explicit residual
 * sum(w (f - R T c)^2) (zscan.py:777-780), a check on the forward pass.
 * The template-flux gradient is summed in a fixed order (per item, per
 * template sample, items in order): the same bits on every run.
 */
int rr_zscan_grad(const rr_grad_desc *desc, void *stream);

/*
 * Archetype fit (redrock's archetype mode, the final fit of every minimum).
 *
 * Replaces Archetype.get_best_archetype and nearest_neighbour_model
 * (py/redrock/archetypes.py:241-400) over per_camera_coeff_with_least_square_batch
 * (py/redrock/zscan.py:245-348) and calc_zchi2_batch with solve_matrices
 * "BVLS" (py/redrock/zscan.py:626-781,1107-1152): for one target at one
 * redshift the design matrix is [R a_1 .. R a_na | R L_1 .. R L_nleg] (a_j
 * rebinned archetypes times the IGM factor, L_l Legendre columns on the
 * pixel axis), M = A'WA + prior, y = A'Wf, the coefficients solve
 * min |M c - y|^2 with c_j >= 0 for j < nbound (scipy lsq_linear 'bvls' on
 * (M, y), zscan.py:1175-1201), and chi2 = sum w (f - A c)^2.
 */
#define RR_ARCH_MAX_NA 4     /* archetype columns per problem (n_nearest)      */
#define RR_ARCH_MAX_NLEG 12  /* Legendre columns (cameras x degree)            */

/*
 * rr_arch_prep — the z-independent Legendre terms of every target:
 *   d_wrl (ntarget, nleg, npix_pad)  w * (R L_l)
 *   d_gll (ntarget, nleg, nleg)      (R L)' W (R L)
 *   d_yl  (ntarget, nleg)            (R L)' W f
 *   d_fwf (ntarget)                  f' W f
 * d_leg (nleg, npix_pad) holds the Legendre columns on the pixel axis
 * (zero outside the camera a column belongs to); d_rband, d_w, d_f as in
 * rr_zscan_desc.  nleg may be 0 (then only d_fwf is written).
 */
int rr_arch_prep(const double *d_rband, const double *d_w, const double *d_f, int ntarget,
                 int npix_pad, const double *d_leg, int nleg, double *d_wrl, double *d_gll,
                 double *d_yl, double *d_fwf, void *stream);

typedef struct {
    int npair;                  /* (target, redshift) pairs                         */
    int nprob;                  /* problems per pair                                */
    int na;                     /* archetype columns per problem, 1..RR_ARCH_MAX_NA */
    int nleg;                   /* Legendre columns, 0..RR_ARCH_MAX_NLEG            */
    int nbound;                 /* coefficients 0..nbound-1 are bounded below by 0  */
    int npix_pad, halo;         /* pixel axis of the targets; halo >= 5 zero rows   */
    const int *d_pair_target;   /* (npair) target of each pair                      */
    const double *d_acache;     /* (npair, npix_pad + 2 halo, ld) rebinned
                                   archetypes: row halo + p is pixel p              */
    long long ld;               /* leading dimension of d_acache, >= narch          */
    int narch;                  /* archetypes in d_acache                           */
    const int *d_columns;       /* (npair, nprob, na) archetype of every column, or
                                   NULL: problem q uses archetype q (na == 1)       */
    const double *d_rband, *d_w, *d_wf;             /* per target, rr_zscan_desc    */
    const double *d_wrl, *d_gll, *d_yl, *d_fwf;     /* rr_arch_prep                 */
    const double *d_prior;      /* (na+nleg)^2 added to M, or NULL                  */
    const double *d_wave;       /* (npix_pad) pixel wavelengths; NULL if nlines 0  */
    const double *d_zigm;       /* (npair) redshift of the IGM factor; NaN = none   */
    const rr_lyman_line *lines; /* host array                                       */
    int nlines;
    double *d_coeff;            /* (npair, nprob, na + nleg) out                     */
    double *d_chi2;             /* (npair, nprob) out                                */
    void *d_work;               /* rr_arch_workspace_bytes bytes                     */
    long long work_bytes;
    /* Fused rebin, used when d_acache is NULL: every window row of the fit is
       the bin mean of the archetype (rebin.py:126-176 expression by
       expression, as rr_rebin) computed in the kernel, so no cache exists.   */
    const double *d_x;          /* (nx) rest-frame archetype wavelengths, increasing */
    int nx;
    const double *d_yt;         /* (nx, ldy) archetype fluxes, archetype fastest     */
    long long ldy;              /* >= narch                                          */
    const double *d_pe_lo;      /* (npix) lower observed-frame edge of every pixel   */
    const double *d_pe_hi;      /* (npix) upper edge                                 */
    int npix;                   /* real pixels; rows npix..npix_pad-1 are zero       */
    const double *d_zrebin;     /* (npair) redshift of the archetype rebin           */
} rr_arch_desc;

long long rr_arch_workspace_bytes(const rr_arch_desc *desc);

/*
 * rr_arch_fit — normal equations, bounded solve and chi2 of every problem.
 * A basis column with zero weighted norm (and no prior) gets coefficient 0,
 * the minimum-norm choice of the least-squares solve.  The caller checks the
 * rebin range (each pixel's edges inside the redshifted archetype range,
 * rebin.py:332-341) for the fused path; searches are clamped in bounds.
 */
int rr_arch_fit(const rr_arch_desc *desc, void *stream);

/*
 * rr_minima_head — host function (no device work): the first k entries of
 * find_minima(x[t]) (fitz.py:46-67: local minima, edges included, sorted by
 * value) for every row t of the nt x nz array x (row stride ldx doubles),
 * the candidates the fine-fit loop walks first (fitz.py:198-210).
 * head: nt x k, -1 past a row's candidates; ncand[t]: the row's number of
 * candidates; flag[t] = 1 where ties or non-finite values could leave the
 * order to the sort's tie handling (the caller takes find_minima there).
 */
int rr_minima_head(const double *x, long long nt, long long nz, long long ldx, int k,
                   long long *head, long long *ncand, unsigned char *flag);

/*
 * rr_gather_rows — host function (no device work): dst[i * row_bytes ..] =
 * the row_bytes bytes at src[i], for i < n, on nthreads threads.  The host
 * half of a batch upload (spectral_data, zscan.py:192-209, for every target
 * at once): per-target arrays into one pinned staging buffer before the DMA.
 */
int rr_gather_rows(const unsigned long long *src, long long n, long long row_bytes, void *dst,
                   int nthreads);

#ifdef __cplusplus
}
#endif
#endif /* REDROCK_HIP_H */
