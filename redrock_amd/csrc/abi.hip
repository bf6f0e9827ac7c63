// Exported C ABI: version, error text, descriptor validation and dispatch.
#include "common.hpp"
#include "zscan_args.hpp"

#include <cstring>
#include <mutex>

namespace rr {
static thread_local char g_err[1024] = "";

void set_error(const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}
}  // namespace rr

extern "C" int rr_abi_version(void) { return RR_ABI_VERSION; }

extern "C" const char *rr_last_error(void) { return rr::g_err; }

static bool is_mixed(int precision) {
    return precision == RR_PREC_F16 || precision == RR_PREC_BF16 || precision == RR_PREC_F32 ||
           precision == RR_PREC_F16X3 || precision == RR_PREC_F16R;
}

// The lane-per-item kernels read rows up to halo + lane_pixels + 4 of their cache.
static bool lane_inputs(const rr_zscan_desc *d) {
    return d->d_tcache_r && d->d_rlane && d->uniform_z &&
           d->cache_rows >= d->halo + rr::lane_pixels(d->npix_pad) + 5;
}

// Inputs of the matrix-core explicit residual of F16R.
static bool resid_mfma_inputs(const rr_zscan_desc *d) {
    return d->d_tcache_u && d->uniform_z && d->halo == 8;
}

// Default configuration of the matrix-core residual (index into its table).
constexpr int kResidMfmaDefault = 0;

extern "C" long long rr_zscan_workspace_bytes(const rr_zscan_desc *d) {
    if (!d || d->nitems < 0 || d->nb < 1 || d->nb > RR_MAX_NBASIS || d->total_rows < 0)
        return rr::fail(RR_EINVAL, "rr_zscan_workspace_bytes: bad descriptor");
    return rr::workspace_bytes(d->nitems, d->nb, is_mixed(d->precision), d->total_rows);
}

// Validate a scan descriptor and fill the kernel arguments.  Returns RR_OK,
// an error code, or 1 for a valid call with nothing to do.
static int scan_args(const rr_zscan_desc *d, rr::ZArgs &a) {
    if (!d) return rr::fail(RR_EINVAL, "rr_zscan: null descriptor");
    if (d->nitems < 0 || d->nitems > 65535)
        return rr::fail(RR_EINVAL, "rr_zscan: nitems=%d outside 0..65535", d->nitems);
    if (d->npix < 1 || d->npix_pad < d->npix || (d->npix_pad % 32) != 0)
        return rr::fail(RR_EINVAL, "rr_zscan: bad pixel sizes npix=%d npix_pad=%d", d->npix,
                        d->npix_pad);
    if (d->nb < 1 || d->nb > RR_MAX_NBASIS)
        return rr::fail(RR_EUNSUPPORTED, "rr_zscan: nb=%d outside 1..%d", d->nb, RR_MAX_NBASIS);
    if (d->halo < 8 || d->cache_rows < (long long)d->npix_pad + 2 * d->halo)
        return rr::fail(RR_EINVAL, "rr_zscan: cache rows %lld too small for npix_pad %d halo %d",
                        d->cache_rows, d->npix_pad, d->halo);
    // no items (or only empty ones): a valid call with nothing to do; zero-size
    // device arrays may then have null addresses
    if (d->nitems == 0 || d->max_item_nz <= 0) return 1;
    if (d->precision == RR_PREC_F64 && !d->d_tcache && !lane_inputs(d))
        return rr::fail(RR_EINVAL, "rr_zscan: F64 needs d_tcache, or d_tcache_r and d_rlane with "
                                   "uniform items and cache_rows >= halo + rr_lane_pixels(npix_pad) + 5");
    if ((d->precision != RR_PREC_F64 && !d->d_tcache) || !d->d_rband || !d->d_w || !d->d_wf || !d->d_f ||
        !d->d_item_target ||
        !d->d_item_z0 || !d->d_item_nz || !d->d_item_out || !d->d_zchi2 || !d->d_zcoeff)
        return rr::fail(RR_EINVAL, "rr_zscan: null required pointer");
    const bool mixed = is_mixed(d->precision);
    const long long need = rr::workspace_bytes(d->nitems, d->nb, mixed, d->total_rows);
    if (!d->d_work || d->work_bytes < need)
        return rr::fail(RR_EINVAL, "rr_zscan: workspace of %lld bytes needed, %lld given", need,
                        d->work_bytes);
    if (d->precision == RR_PREC_F64) {
        if (!lane_inputs(d) && d->cache_ldz < d->cache_nz)
            return rr::fail(RR_EINVAL, "rr_zscan: cache_ldz < cache_nz");
    } else if (mixed) {
        if (d->precision == RR_PREC_F16R && !lane_inputs(d) && !resid_mfma_inputs(d))
            return rr::fail(RR_EINVAL, "rr_zscan: F16R needs uniform items and either d_tcache_u "
                                       "(matrix-core residual) or d_tcache_r and d_rlane with "
                                       "cache_rows >= halo + rr_lane_pixels(npix_pad) + 5");
        if (!d->d_rtile || !d->d_scales || (d->precision == RR_PREC_F32 && !d->d_sfcol))
            return rr::fail(RR_EINVAL, "rr_zscan: mixed precision needs rr_pack_mixed outputs");
        if (d->halo != 8) return rr::fail(RR_EINVAL, "rr_zscan: mixed precision needs halo == 8");
        if (d->nb + 2 > 16) return rr::fail(RR_EUNSUPPORTED, "rr_zscan: mixed needs nb <= 14");
    } else {
        return rr::fail(RR_EINVAL, "rr_zscan: unknown precision %d", d->precision);
    }
    if (d->solve != RR_SOLVE_PCA && d->solve != RR_SOLVE_NNLS && d->solve != RR_SOLVE_BVLS)
        return rr::fail(RR_EINVAL, "rr_zscan: unknown solve %d", d->solve);
    if (d->solve == RR_SOLVE_BVLS && !d->d_bounds)
        return rr::fail(RR_EINVAL, "rr_zscan: RR_SOLVE_BVLS needs d_bounds");

    a.nitems = d->nitems; a.npix = d->npix; a.npix_pad = d->npix_pad; a.nb = d->nb;
    a.precision = d->precision; a.halo = d->halo;
    a.tcache = d->d_tcache; a.cache_rows = d->cache_rows; a.cache_ldz = d->cache_ldz;
    a.cache_nz = d->cache_nz; a.tscale = d->tscale;
    a.rband = d->d_rband; a.w = d->d_w; a.wf = d->d_wf; a.f = d->d_f;
    a.rtile = d->d_rtile; a.sfcol = d->d_sfcol; a.scales = d->d_scales;
    a.tcache_r = d->d_tcache_r; a.rlane = d->d_rlane;
    a.uniform_z = d->uniform_z;
    a.tile = d->tile;
    a.resid_tile = d->resid_tile;
    a.item_target = d->d_item_target; a.item_z0 = d->d_item_z0; a.item_nz = d->d_item_nz;
    a.item_out = d->d_item_out; a.prior = d->d_prior; a.oiisum = d->d_oiisum;
    a.zchi2 = d->d_zchi2; a.zcoeff = d->d_zcoeff; a.penalty = d->d_penalty; a.status = d->d_status;
    a.max_item_nz = d->max_item_nz;
    a.work = d->d_work; a.total_rows = d->total_rows;
    a.solve = d->solve;
    a.nonneg_mask = d->nonneg_mask;
    a.bounds = d->d_bounds;
    a.tcache_u = d->d_tcache_u;
    a.rtile_planes = (d->precision == RR_PREC_F16X3 || d->precision == RR_PREC_F16R) ? 2 : 1;
    return RR_OK;
}

extern "C" int rr_zscan(const rr_zscan_desc *d, void *stream) {
    rr::ZArgs a;
    const int v = scan_args(d, a);
    if (v) return v < 0 ? v : RR_OK;
    const bool mixed = is_mixed(d->precision);
    hipStream_t s = static_cast<hipStream_t>(stream);

    const int st = d->stages ? d->stages
                             : (RR_STAGE_STATS | RR_STAGE_MAIN | RR_STAGE_FINISH | RR_STAGE_RESID);
    int rc = RR_OK;
    if (st & RR_STAGE_STATS) rc = rr::zscan_stats(a, s);
    if (rc) return rc;
    if (st & RR_STAGE_MAIN) {
        if (d->precision == RR_PREC_F16R) {
            rr::ZArgs a16 = a;            // the fp16 normal equations of RR_PREC_F16
            a16.precision = RR_PREC_F16;
            rc = rr::zscan_mfma(a16, s);
        } else {
            rc = d->precision == RR_PREC_F32     ? rr::zscan_mfma32(a, s)
                 : d->precision == RR_PREC_F16X3 ? rr::zscan_mfma3(a, s)
                 : mixed                         ? rr::zscan_mfma(a, s)
                 : lane_inputs(d)                ? rr::zscan_f64r(a, s)
                                                 : rr::zscan_f64(a, s);
        }
    }
    if (rc) return rc;
    if (st & RR_STAGE_FINISH) rc = rr::zscan_finish(a, s);
    if (rc) return rc;
    if ((st & RR_STAGE_RESID) && d->precision == RR_PREC_F16R) {
        // resid_tile: the lane kernels' table, then the matrix-core kernel's
        const int nlane = rr::zscan_resid_tiles(nullptr, 0);
        const int nmfma = rr::zscan_resid_mfma_tiles(nullptr, 0);
        int t = d->resid_tile;
        if (t < 0 || t >= nlane + nmfma)
            return rr::fail(RR_EINVAL, "rr_zscan(residual): resid_tile %d outside 0..%d", t, nlane + nmfma - 1);
        // default: 16 waves (4 per SIMD, 128 VGPRs) up to nb 10, where it
        // fits without spills; 8 waves beyond
        if (t == 0 && resid_mfma_inputs(d)) t = nlane + (d->nb <= 10 ? kResidMfmaDefault : 1);
        if (t >= nlane) {
            if (!resid_mfma_inputs(d))
                return rr::fail(RR_EINVAL, "rr_zscan(residual): resid_tile %d needs d_tcache_u, uniform "
                                           "items and halo 8", t);
            rc = rr::zscan_resid_mfma(a, t - nlane, s);
        } else {
            if (!lane_inputs(d))
                return rr::fail(RR_EINVAL, "rr_zscan(residual): resid_tile %d needs d_tcache_r and d_rlane", t);
            rc = rr::zscan_resid(a, s);
        }
    }
    return rc;
}

// ---------------------------------------------------------------------------
// rr_zscan_noalloc: rr_zscan with its workspace in a static device array.
// `scratch` is a host-side stand-in for the array; on every device the
// library runs on, its device address comes from hipGetSymbolAddress (once
// per device, with that device current) and only that address is passed to
// the kernels.  Every call on a device shares the device's array: calls that
// overlap on one device race.  Nothing in the workspace must start at zero
// (the statistics and normal-equation stages write every entry the later
// stages read).
// ---------------------------------------------------------------------------
static __device__ __align__(256) unsigned char scratch[RR_NOALLOC_SCRATCH_BYTES];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
void *g_scratch_ptr[kMaxDevices] = {};

int scratch_address(void **out) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return rr::fail(RR_ELAUNCH, "rr_zscan_noalloc: hipGetDevice: %s", hipGetErrorString(e));
    if (dev < 0 || dev >= kMaxDevices)
        return rr::fail(RR_EUNSUPPORTED, "rr_zscan_noalloc: device %d beyond %d", dev, kMaxDevices);
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return rr::fail(RR_ELAUNCH, "rr_zscan_noalloc: hipGetSymbolAddress(scratch) on device %d: %s", dev,
                            hipGetErrorString(e));
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return RR_OK;
}
}  // namespace

extern "C" int rr_zscan_noalloc(const rr_zscan_desc *d, void *stream) {
    if (!d) return rr::fail(RR_EINVAL, "rr_zscan_noalloc: null descriptor");
    const long long need = rr_zscan_workspace_bytes(d);
    if (need < 0) return (int)need;
    if (need > (long long)RR_NOALLOC_SCRATCH_BYTES)
        return rr::fail(RR_EINVAL,
                        "rr_zscan_noalloc: this scan needs a %lld-byte workspace, the static scratch holds "
                        "%lld bytes (fewer items or redshifts per call, or rr_zscan with a caller workspace)",
                        need, (long long)RR_NOALLOC_SCRATCH_BYTES);
    void *ptr = nullptr;
    const int rc = scratch_address(&ptr);
    if (rc) return rc;
    rr_zscan_desc dd = *d;
    dd.d_work = ptr;
    dd.work_bytes = RR_NOALLOC_SCRATCH_BYTES;
    return rr_zscan(&dd, stream);
}

extern "C" int rr_zscan_polish(const rr_zscan_desc *d, const rr_polish_desc *p, void *stream) {
    rr::ZArgs a;
    const int v = scan_args(d, a);
    if (v) return v < 0 ? v : RR_OK;
    if (d->precision == RR_PREC_F64) return RR_OK;
    if (d->solve != RR_SOLVE_PCA)
        return rr::fail(RR_EUNSUPPORTED,
                        "rr_zscan_polish: re-solves M c = y; bounded (NNLS, BVLS) scans are not polished");
    if (!p) return rr::fail(RR_EINVAL, "rr_zscan_polish: null polish descriptor");
    if (!p->d_tcache64 || !p->d_windows || !p->d_nwindows || !p->d_queue)
        return rr::fail(RR_EINVAL, "rr_zscan_polish: null required pointer");
    if (p->cache_ldz < d->cache_nz)
        return rr::fail(RR_EINVAL, "rr_zscan_polish: cache_ldz %lld < cache_nz %d", p->cache_ldz,
                        d->cache_nz);
    if (!(p->tol >= 0.0) || p->tol > 1e300)
        return rr::fail(RR_EINVAL, "rr_zscan_polish: tol must be finite and >= 0");
    if (p->max_windows < 1 || p->max_windows > RR_POLISH_MAX_WINDOWS)
        return rr::fail(RR_EINVAL, "rr_zscan_polish: max_windows %d outside 1..%d", p->max_windows,
                        RR_POLISH_MAX_WINDOWS);
    if (d->max_item_nz > 16384)   // window minima in 64 KB of LDS
        return rr::fail(RR_EUNSUPPORTED, "rr_zscan_polish: items of more than 16384 redshifts");
    rr::PolishArgs pa;
    pa.tcache64 = p->d_tcache64;
    pa.ldz = p->cache_ldz;
    pa.tol = p->tol;
    pa.max_windows = p->max_windows;
    pa.nw_all = (d->max_item_nz + RR_POLISH_WINDOW - 1) / RR_POLISH_WINDOW;
    pa.windows = p->d_windows;
    pa.nwindows = p->d_nwindows;
    pa.queue = p->d_queue;
    return rr::zscan_polish(a, pa, static_cast<hipStream_t>(stream));
}

extern "C" int rr_zscan_tile_table(int precision, int *cfg, int max) {
    if (!cfg || max < 0) return rr::fail(RR_EINVAL, "rr_zscan_tile_table: bad arguments");
    if (precision == RR_PREC_F16 || precision == RR_PREC_BF16 || precision == RR_PREC_F16R)
        return rr::zscan_tiles16(cfg, max);
    if (precision == RR_PREC_F16X3) return rr::zscan_tiles3(cfg, max);
    if (precision == RR_PREC_F64) return rr::zscan_f64r_tiles(cfg, max);
    return 0;
}

// Entries: the lane kernel's configurations (kind 1), then the matrix-core
// kernel's (kind 2).
extern "C" int rr_zscan_resid_tiles(int *cfg, int max) {
    if (!cfg || max < 0) return rr::fail(RR_EINVAL, "rr_zscan_resid_tiles: bad arguments");
    const int nlane = rr::zscan_resid_tiles(cfg, max);
    const int room = max > nlane ? max - nlane : 0;
    const int nmfma = rr::zscan_resid_mfma_tiles(cfg + 4 * (max < nlane ? max : nlane), room);
    (void)nmfma;
    return nlane + rr::zscan_resid_mfma_tiles(nullptr, 0);
}

extern "C" int rr_resid_u_slots(int nb, int *nprod) {
    if (nb < 1 || nb > RR_MAX_NBASIS) return rr::fail(RR_EINVAL, "rr_resid_u_slots: nb=%d outside 1..%d", nb, RR_MAX_NBASIS);
    if (nprod) *nprod = rr::resid_u_nprod(nb);
    return rr::resid_u_slots(nb);
}
