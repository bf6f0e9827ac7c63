"""Redshift chi2 scan on the GPU (py/redrock/zscan.py).

Drop-in functions with the reference signatures:

* calc_zchi2(target_ids, target_data, dtemplate, progress=None, use_gpu=False)
  (zscan.py:807-889) — all targets of a template in one rr_zscan launch.
* calc_zchi2_batch(spectra, tdata, weights, flux, wflux, nz, nbasis, ...)
  (zscan.py:626-781) — one target over the redshifts of `tdata`.
* calc_zchi2_targets(targets, templates, mp_procs=1, use_gpu=False)
  (zscan.py:910-1105), calc_negOII_penalty (:783-800), spectral_data
  (:192-209), HUGE_CHI2 (:18).

Every call runs the HIP kernels of libredrock_hip.so; there is no CPU path.
The ``use_gpu`` and ``fullprecision`` arguments are accepted for signature
compatibility (the scan is always on the GPU and always deterministic).
``precision`` selects 'f64' (default, the reference's arithmetic) or a
mixed-precision MFMA form: 'f16r' (fp16 normal equations for the
coefficients, chi2 as the explicit residual at them, evaluated with
split-fp16 MFMA products: within ~1e-6 of FP64), 'f16x3' (split fp16, chi2 within ~1e-5 of FP64 at per-pixel S/N 1),
'f16', 'bf16' or 'f32'; the default can be changed with the environment
variable RR_PRECISION.

Device-side objects:

* PixelLayout   — the concatenated pixel axis of a target (spectra order).
* TemplateCache — a template rebinned (IGM included) to every redshift on
  every wavelength grid, resident in HBM in the layout of each kernel.
* TargetBatch   — per-target R bands, ivar, flux, ivar*flux in HBM.
* ScanPlan      — a prepared rr_zscan launch (descriptor, outputs, workspace)
  that can be replayed.
* ScanGraph     — every template class over one TargetBatch captured in one
  HIP graph.
"""
from __future__ import annotations

import ctypes
import math
import os
import threading
import warnings

import numpy as np

from . import _hip
from .igm import igm_applies, lyman_table
from .rebin import centers2edges, rebin_into

HUGE_CHI2 = 9999999.0
HALO = 8


def _roundup(n, m):
    return (n + m - 1) // m * m


def default_precision():
    return os.environ.get('RR_PRECISION', 'f64')


# Tolerance of the FP64 polish per mixed precision, relative to the target's
# f'Wf (a mixed chi2's error scales with f'Wf, the sum chi2 cancels against,
# not with chi2): at least four times the largest |chi2_mixed - chi2_f64| /
# f'Wf measured for that mode over GALAXY / QSO / STAR targets at per-pixel
# S/N 0.3 to 30 and the masked GALAXY/QSO/STAR batch.
POLISH_TOL = {'f16r': 1e-6, 'f16x3': 1e-5, 'f32': 1e-5, 'f16': 1e-3, 'bf16': 5e-3}
POLISH_WINDOWS = 16
STAGE_POLISH = 16     # ScanPlan.run stage bit (rr_zscan_polish), beside RR_STAGE_*


SOLVES = ('PCA', 'NMF', 'NNLS', 'BVLS')


def solve_spec(method, nb, bounds=None):
    """(RR_SOLVE_*, non-negativity mask, bounds) of a solve_matrices algorithm
    (zscan.py:1107-1205): PCA -> M c = y; NMF / NNLS -> nnls(M, y), every
    coefficient >= 0; BVLS -> lsq_linear(M, y, bounds, 'bvls') for bounds
    (2, nb) or (2,) (None: no bound at all).  Bounds whose lower bounds are 0
    or -inf and upper bounds +inf (the archetype bounds of zscan.py:289-294)
    run as the masked NNLS solve; any others as RR_SOLVE_BVLS with the (2, nb)
    float64 bounds array returned third (None for the other solves).  As
    lsq_linear, every lower bound must be strictly below its upper bound."""
    key = 'PCA' if method is None else str(method).upper()
    if key == 'PCA':
        return _hip.RR_SOLVE_PCA, 0, None
    if key in ('NMF', 'NNLS'):
        return _hip.RR_SOLVE_NNLS, (1 << nb) - 1, None
    if key == 'BVLS':
        if bounds is None:
            return _hip.RR_SOLVE_NNLS, 0, None
        b = np.asarray(bounds, dtype=np.float64)
        if b.shape == (2,):                      # scalar bounds, broadcast as lsq_linear does
            b = np.repeat(b[:, None], nb, axis=1)
        if b.shape != (2, nb):
            raise ValueError(f'BVLS bounds must have shape (2, {nb}) or (2,), got {b.shape}')
        lo, hi = b
        if np.any(np.isnan(b)) or not np.all(lo < hi):
            raise ValueError('Each lower bound must be strictly less than each upper bound.')
        if np.all(np.isposinf(hi)) and np.all((lo == 0.0) | np.isneginf(lo)):
            return _hip.RR_SOLVE_NNLS, int(sum(1 << j for j in range(nb) if lo[j] == 0.0)), None
        return _hip.RR_SOLVE_BVLS, 0, np.ascontiguousarray(b)
    raise NotImplementedError(f'solve_matrices algorithm {method!r} is not available in the scan')


def solve_code(method):
    """RR_SOLVE_* of a template method (PCA or NMF / NNLS)."""
    return solve_spec(method, 1)[0]


def polish_tol(precision, polish):
    """Tolerance of the FP64 polish for `precision`, or None when it does not run.

    polish: None or True = the mode's POLISH_TOL, False = off, a number = that
    tolerance.  FP64 scans are never polished.
    """
    if precision == 'f64' or polish is False:
        return None
    if polish is None or polish is True:
        return POLISH_TOL[precision]
    tol = float(polish)
    if not (tol >= 0.0 and math.isfinite(tol)):
        raise ValueError(f'polish tolerance must be finite and >= 0, got {polish!r}')
    return tol


def spectral_data(spectra):
    """Concatenated (weights, flux, wflux) of a list of spectra (zscan.py:192-209)."""
    weights = np.concatenate([s.ivar for s in spectra])
    flux = np.concatenate([s.flux for s in spectra])
    return weights, flux, weights * flux


def calc_negOII_penalty(OIItemplate, coeff):
    """Penalty for negative [OII] flux (zscan.py:783-800)."""
    if hasattr(coeff, 'detach'):
        import torch
        oii = torch.as_tensor(OIItemplate, device=coeff.device, dtype=coeff.dtype)
        OIIflux = (coeff @ oii).sum(dim=1)
        return -OIIflux * (OIIflux < 0)
    OIIflux = np.sum(coeff @ OIItemplate, axis=1)
    return -OIIflux * (OIIflux < 0)


class PixelLayout(object):
    """Concatenated pixel axis of a target: segments (wavehash, length) in order."""

    def __init__(self, keys, sizes):
        self.keys = tuple(keys)
        self.sizes = tuple(int(n) for n in sizes)
        self.offsets = tuple(int(o) for o in np.concatenate([[0], np.cumsum(self.sizes)])[:-1])
        self.npix = int(sum(self.sizes))
        self.npix_pad = _roundup(self.npix, 32)
        self.halo = HALO
        self.cache_rows = self.npix_pad + 48   # halo + pixels + the lane kernels' window overrun

    @property
    def signature(self):
        return (self.keys, self.sizes)

    @staticmethod
    def of(spectra):
        return PixelLayout([s.wavehash for s in spectra], [s.nwave for s in spectra])

    @staticmethod
    def signature_of(spectra):
        """PixelLayout.of(spectra).signature without building the layout."""
        return (tuple(s.wavehash for s in spectra), tuple(int(s.nwave) for s in spectra))


class TemplateCache(object):
    """A template rebinned to a redshift grid on every wavelength grid, in HBM.

    Equivalent of DistTemplate's ``local.data`` (templates.py:584-597): the
    rebin (rebin.py:453) and the IGM multiply (igm.py:369) are one rr_rebin
    launch per wavelength grid.  Raises ValueError like the reference when a
    wavelength grid leaves the redshifted template range.
    """

    def __init__(self, template, dwave, redshifts=None, check_range=True, flux=None):
        """flux: optional (nbasis, nwave) float64 device tensor replacing template.flux."""
        self.template = template
        self._flux = flux
        self.dwave = dict(dwave)
        self.redshifts = np.ascontiguousarray(
            template.redshifts if redshifts is None else redshifts, dtype=np.float64)
        self.nz = len(self.redshifts)
        self.nb = template.nbasis
        self.edges = {k: centers2edges(w) for k, w in self.dwave.items()}
        if check_range and self.nz > 0:
            xmin = template.minwave * (1 + self.redshifts.max())
            xmax = template.maxwave * (1 + self.redshifts.min())
            for e in self.edges.values():
                if e[0] < xmin or e[-1] > xmax:
                    raise ValueError('edges must be within input x range')
        fmax = (float(flux.abs().max()) if flux is not None
                else float(np.max(np.abs(template.flux))))
        self.tscale = 2.0 ** (8 - math.ceil(math.log2(fmax))) if fmax > 0 else 1.0
        self._dev = None
        self._bufs = dict()

    def lyman(self, key):
        """Line table for grid `key`, or None when transmission_Lyman returns None."""
        model = self.template.igm_model
        if not igm_applies(self.redshifts, self.dwave[key], model):
            return None
        return lyman_table(model)

    def device(self):
        import torch
        if self._dev is None:
            dev = torch.device('cuda')
            def up(a):
                return torch.as_tensor(np.ascontiguousarray(a, dtype=np.float64),
                                       device=dev).contiguous()
            self._dev = dict(
                x=up(self.template.wave),
                y=(self._flux.detach().contiguous() if self._flux is not None
                   else up(self.template.flux)),
                z=up(self.redshifts),
                edges={k: up(e) for k, e in self.edges.items()},
                wave={k: up(w) for k, w in self.dwave.items()})
        return self._dev

    def layout(self, keys):
        return PixelLayout(keys, [len(self.dwave[k]) for k in keys])

    def get(self, layout, precision):
        """Device cache for `layout` in the layout of `precision`'s kernel.

        f64 : float64 [nb][cache_rows][ldz]   (z fastest), ldz = nz rounded to 64
        f64r/f32r: float64 / float32 [nz][cache_rows][nb] (basis fastest),
              unscaled: the caches of the lane-per-item kernels
        f16/bf16/f32: [nz][nb][cache_rows] (pixel fastest), values * tscale
        f16x3: fp16 [nz][hi, lo][nb][cache_rows], hi + lo = values * tscale
        f16u: fp16 [nz][cache_rows][KS], the template operand of the 'f16r'
              matrix-core residual: slots hi | lo | hi [| lo] of values *
              tscale per basis vector (_hip.resid_u_slots)
        Returns (tensor, ldz).
        """
        import torch
        key = (layout.signature, precision)
        if key in self._bufs:
            return self._bufs[key]
        d = self.device()
        nb, nz, rows = self.nb, self.nz, layout.cache_rows
        if precision in ('f64r', 'f32r'):
            ldz = 0
            dt, ot = ((torch.float64, _hip.RR_F64) if precision == 'f64r'
                      else (torch.float32, _hip.RR_F32))
            buf = torch.zeros(nz * rows * nb, dtype=dt, device='cuda')
            for k, off in zip(layout.keys, layout.offsets):
                rebin_into(buf[(layout.halo + off) * nb:], d['x'], d['y'], d['edges'][k], d['z'],
                           rows * nb, nb, 1, ot, 1.0, self.lyman(k), d['wave'][k])
        elif precision == 'f64':
            ldz = _roundup(max(nz, 1), 64)
            # zeros only where the rebin writes nothing (halo rows, rows after
            # the pixels, redshift padding): a fine-grid cache is written once
            buf = torch.empty(nb * rows * ldz, dtype=torch.float64, device='cuda')
            v = buf.view(nb, rows, ldz)
            h, e = layout.halo, layout.halo + layout.npix
            v[:, :h].zero_()
            v[:, e:].zero_()
            v[:, h:e, nz:].zero_()
            for k, off in zip(layout.keys, layout.offsets):
                rebin_into(buf[(layout.halo + off) * ldz:], d['x'], d['y'], d['edges'][k], d['z'],
                           1, ldz, rows * ldz, _hip.RR_F64, 1.0, self.lyman(k), d['wave'][k])
        elif precision == 'f16u':
            ldz = 0
            full = torch.zeros(nz * rows * nb, dtype=torch.float32, device='cuda')
            for k, off in zip(layout.keys, layout.offsets):
                rebin_into(full[(layout.halo + off) * nb:], d['x'], d['y'], d['edges'][k], d['z'],
                           rows * nb, nb, 1, _hip.RR_F32, self.tscale, self.lyman(k), d['wave'][k])
            buf = _u_operand(full.reshape(nz, rows, nb))
        elif precision == 'f16x3':
            ldz = 0
            full = torch.zeros(nz * nb * rows, dtype=torch.float32, device='cuda')
            for k, off in zip(layout.keys, layout.offsets):
                rebin_into(full[layout.halo + off:], d['x'], d['y'], d['edges'][k], d['z'],
                           nb * rows, 1, rows, _hip.RR_F32, self.tscale, self.lyman(k), d['wave'][k])
            buf = _split_planes(full.reshape(nz, nb, rows))
        else:
            ldz = 0
            # + 64 zero elements read by the MFMA lanes of unused basis columns
            dt = torch.float32 if precision == 'f32' else torch.int16
            buf = torch.zeros(nz * nb * rows + 64, dtype=dt, device='cuda')
            ot = {'f16': _hip.RR_F16, 'bf16': _hip.RR_BF16, 'f32': _hip.RR_F32}[precision]
            for k, off in zip(layout.keys, layout.offsets):
                rebin_into(buf[layout.halo + off:], d['x'], d['y'], d['edges'][k], d['z'],
                           nb * rows, 1, rows, ot, self.tscale, self.lyman(k), d['wave'][k])
        self._bufs[key] = (buf, ldz)
        return self._bufs[key]

    def release(self):
        self._bufs.clear()

    def reference_layout(self):
        """{wavehash: (nz, nwave, nbasis) float64 device tensor} (DistTemplate.local.data)."""
        import torch
        d = self.device()
        out = dict()
        for k, w in self.dwave.items():
            n = len(w)
            t = torch.empty((self.nz, n, self.nb), dtype=torch.float64, device='cuda')
            rebin_into(t, d['x'], d['y'], d['edges'][k], d['z'], n * self.nb, self.nb, 1,
                       _hip.RR_F64, 1.0, self.lyman(k), d['wave'][k])
            out[k] = t
        return out


class TargetBatch(object):
    """Per-target scan inputs in HBM for targets that share one PixelLayout.

    rband (nt, npix_pad, 11), w, f, wf (nt, npix_pad) float64; pixels past
    npix have zero weight.  w and f are the concatenated ivar and flux
    (spectral_data, zscan.py:192-209); wf = w*f.  good_pixels: host int
    array, each target's number of pixels with ivar > 0.
    """

    def __init__(self, targets=None, layout=None, arrays=None, grouped=False):
        """grouped: the targets are known to share `layout` (target_batches
        grouped them by its signature), so they are not checked again."""
        import torch
        dev = torch.device('cuda')
        if arrays is not None:
            layout, rband, w, f, wf = arrays
            self.rband = torch.as_tensor(np.ascontiguousarray(rband), dtype=torch.float64, device=dev)
            self.w = torch.as_tensor(np.ascontiguousarray(w), dtype=torch.float64, device=dev)
            self.f = torch.as_tensor(np.ascontiguousarray(f), dtype=torch.float64, device=dev)
            self.wf = torch.as_tensor(np.ascontiguousarray(wf), dtype=torch.float64, device=dev)
            self._good = np.count_nonzero(np.asarray(w) > 0, axis=1)
        else:
            layout = layout or PixelLayout.of(targets[0].spectra)
            nt, P = len(targets), layout.npix_pad
            for tg in ([] if grouped else targets):
                if PixelLayout.signature_of(tg.spectra) != layout.signature:
                    raise ValueError('targets of one batch must share their wavelength grids')
            self.rband = torch.zeros((nt, P, 11), dtype=torch.float64, device=dev)
            self.w = torch.zeros((nt, P), dtype=torch.float64, device=dev)
            self.f = torch.zeros((nt, P), dtype=torch.float64, device=dev)
            for k, off in enumerate(layout.offsets):
                spectra = [tg.spectra[k] for tg in targets]
                n = spectra[0].nwave
                _upload_bands(self.rband[:, off:off + n], spectra)
                seg = torch.empty((2, nt, n), dtype=torch.float64, device=dev)
                _upload_rows(seg[0], [sp.ivar for sp in spectra])
                _upload_rows(seg[1], [sp.flux for sp in spectra])
                self.w[:, off:off + n] = seg[0]
                self.f[:, off:off + n] = seg[1]
            self._good = None
            self.wf = self.w * self.f
        self.layout = layout
        self.nt = self.rband.shape[0]
        self._mixed = dict()

    @property
    def good_pixels(self):
        """Host int array: each target's number of pixels with ivar > 0."""
        if self._good is None:
            self._good = (self.w > 0).sum(dim=1).cpu().numpy()
        return self._good

    def lanes(self, precision, order=None):
        """Lane-major pack of the lane-per-item kernels (rr_pack_lanes), 'f64' or
        'f16r': slot i holds target order[i] (order None: target i, cached)."""
        import torch
        key = 'lanes-' + precision
        if order is None and key in self._mixed:
            return self._mixed[key]
        P = self.layout.npix_pad
        code = _hip.PRECISIONS[precision]
        nslots = self.nt if order is None else len(order)
        nbytes = _hip.lib().rr_pack_lanes_bytes(nslots, P, code)
        if nbytes < 0:
            _hip.check(int(nbytes), 'rr_pack_lanes_bytes')
        buf = torch.empty(max(int(nbytes), 256), dtype=torch.uint8, device='cuda')
        t_order = (None if order is None else
                   torch.as_tensor(np.ascontiguousarray(order, dtype=np.int32), device='cuda'))
        rc = _hip.lib().rr_pack_lanes(_hip.ptr(self.rband), _hip.ptr(self.w), _hip.ptr(self.f),
                                      _hip.ptr(t_order), nslots, P, code, _hip.ptr(buf),
                                      _hip.stream_handle())
        _hip.check(rc, 'rr_pack_lanes')
        if order is None:
            self._mixed[key] = buf
        return buf

    def mixed(self, precision):
        """(rtile, sfcol, scales) for a mixed precision (rr_pack_mixed), cached."""
        import torch
        if precision == 'f16r':
            precision = 'f16x3'              # hi planes: the fp16 normal equations; both: the residual
        if precision not in self._mixed:
            P = self.layout.npix_pad
            dt = torch.float32 if precision == 'f32' else torch.int16
            planes = 2 if precision == 'f16x3' else 1      # hi and lo tiles
            rtile = torch.empty(self.nt * (P // 16) * planes * 64 * 8, dtype=dt, device='cuda')
            # fp32 only: sqrt(w) f in operand order (16-bit tiles carry it themselves)
            sfcol = (torch.empty(self.nt * (P // 32) * 32, dtype=dt, device='cuda')
                     if precision == 'f32' else None)
            scales = torch.empty((self.nt, 2), dtype=torch.float64, device='cuda')
            rc = _hip.lib().rr_pack_mixed(_hip.ptr(self.rband), _hip.ptr(self.w), _hip.ptr(self.f),
                                          self.nt, P, _hip.PRECISIONS[precision], _hip.ptr(rtile),
                                          _hip.ptr(sfcol), _hip.ptr(scales), _hip.stream_handle())
            _hip.check(rc, 'rr_pack_mixed')
            self._mixed[precision] = (rtile, sfcol, scales)
        return self._mixed[precision]


# Host-to-device uploads of per-target rows go through two pinned staging
# buffers kept for the process: a small thread pool fills one (numpy copies
# release the GIL) while the asynchronous copy of the other runs.
STAGE_BYTES = 64 << 20
UPLOAD_THREADS = 2        # filling staging from many small host arrays: more contend (measured)
GATHER_THREADS = 8        # rr_gather_rows: float64 rows copied by the library, without the GIL
DOWNLOAD_THREADS = 8      # copying staging out into new arrays: 3.9 -> 3.4 ms for 95 MB from 4 to 8
_STAGING = []
_POOL = []
_STAGE_LOCK = threading.Lock()        # one upload at a time uses the staging buffers


def _staging(k):
    import torch
    while len(_STAGING) <= k:
        _STAGING.append((torch.empty(STAGE_BYTES, dtype=torch.uint8, pin_memory=True),
                         torch.cuda.Event()))
    return _STAGING[k]


def _copy_pool(nthreads=None):
    """(thread pool, its number of threads) for the staging copies: one pool
    per thread count (UPLOAD_THREADS by default)."""
    n = max(1, min(UPLOAD_THREADS if nthreads is None else nthreads, len(os.sched_getaffinity(0))))
    for pool, m in _POOL:
        if m == n:
            return pool, m
    import concurrent.futures
    _POOL.append((concurrent.futures.ThreadPoolExecutor(max_workers=n), n))
    return _POOL[-1]


def _upload_rows(dst, rows):
    """dst[i] = rows[i] for a contiguous float64 device tensor dst (n, ...) and
    n host arrays of shape dst.shape[1:] (or an (n, ...) host array), in
    chunks of STAGE_BYTES through the pinned staging buffers; the copies are
    ordered on the current stream."""
    import torch
    n = dst.shape[0]
    if n == 0:
        return
    shape = tuple(dst.shape[1:])
    row = int(np.prod(shape, dtype=np.int64)) * 8
    if row > STAGE_BYTES:
        for i in range(n):
            dst[i].copy_(torch.from_numpy(np.ascontiguousarray(rows[i], dtype=np.float64)))
        return
    per = max(1, STAGE_BYTES // row)
    pool, nthr = _copy_pool()
    stream = torch.cuda.current_stream()
    with _STAGE_LOCK:
        _upload_chunks(dst, rows, n, shape, row, per, pool, nthr, stream)


def _row_pointers(rows, n, shape):
    """Data addresses of n host rows when every one is a C-contiguous float64
    array of `shape` (what rr_gather_rows copies), else None."""
    if isinstance(rows, np.ndarray):
        return None
    ptrs = np.empty(n, dtype=np.uint64)
    for i in range(n):
        r = rows[i]
        if not isinstance(r, np.ndarray):
            return None
        ai = r.__array_interface__
        if ai['typestr'] != '<f8' or ai['shape'] != shape or ai['strides'] is not None:
            return None
        ptrs[i] = ai['data'][0]
    return ptrs


def _upload_chunks(dst, rows, n, shape, row, per, pool, nthr, stream):
    import torch
    ptrs = _row_pointers(rows, n, shape)
    for j, c0 in enumerate(range(0, n, per)):
        c1 = min(n, c0 + per)
        buf, ev = _staging(j % 2)
        ev.synchronize()                  # the copy that last read this buffer is done
        src = buf[:(c1 - c0) * row].view(torch.float64).view((c1 - c0,) + shape)
        host = src.numpy()

        def fill(a, b, host=host, c0=c0):
            # one call per slice: numpy copies without the GIL, and few
            # Python steps keep the workers from contending for it
            np.concatenate([np.reshape(r, (1,) + shape) for r in rows[a:b]], axis=0,
                           out=host[a - c0:b - c0], casting='unsafe')
        if ptrs is not None:
            _hip.check(_hip.lib().rr_gather_rows(ptrs[c0:].ctypes.data, c1 - c0, row, host.ctypes.data,
                                                 GATHER_THREADS), 'rr_gather_rows')
        else:
            cuts = np.linspace(c0, c1, min(nthr, c1 - c0) + 1).astype(np.int64)
            list(pool.map(lambda ab: fill(*ab), zip(cuts[:-1], cuts[1:])))
        dst[c0:c1].copy_(src, non_blocking=True)
        ev.record(stream)


def download(tensors, stream=None):
    """New numpy arrays (ordinary pageable memory) holding contiguous device
    tensors, copied after the work already on `stream` (default: the current
    stream).  The copies go through the pinned staging buffers in chunks: the
    device-to-host copy of one chunk runs while the CPU copies the previous
    one out (2 threads).  The CPU's first touch of freshly allocated
    page-locked memory costs about 0.4 ms per MB here, four times a copy out
    of a staging buffer that is reused.  Returns a list in input order."""
    import torch
    s = torch.cuda.current_stream() if stream is None else stream
    outs = [np.empty(tuple(t.shape), dtype=torch.empty(0, dtype=t.dtype).numpy().dtype)
            for t in tensors]
    chunks = []
    for ti, t in enumerate(tensors):
        nbytes = t.numel() * t.element_size()
        chunks += [(ti, off, min(STAGE_BYTES, nbytes - off)) for off in range(0, nbytes, STAGE_BYTES)]
    if not chunks:
        return outs
    pool, nthr = _copy_pool(DOWNLOAD_THREADS)

    def finish(j, ti, off, nb):
        buf, ev = _staging(j % 2)
        ev.synchronize()
        src = buf[:nb].numpy()
        dst = outs[ti].reshape(-1).view(np.uint8)
        cuts = (np.linspace(0, nb // 8, min(nthr, max(1, nb // (1 << 20))) + 1).astype(np.int64) * 8)
        cuts[-1] = nb
        list(pool.map(lambda ab: np.copyto(dst[off + ab[0]:off + ab[1]], src[ab[0]:ab[1]]),
                      zip(cuts[:-1], cuts[1:])))

    with _STAGE_LOCK:
        prev = None
        for j, (ti, off, nb) in enumerate(chunks):
            # buffer j % 2 last held chunk j - 2, already copied out by the CPU
            buf, ev = _staging(j % 2)
            src = tensors[ti].reshape(-1).view(torch.uint8)[off:off + nb]
            with torch.cuda.stream(s):
                buf[:nb].copy_(src, non_blocking=True)
            ev.record(s)
            if prev is not None:
                finish(*prev)
            prev = (j, ti, off, nb)
        finish(*prev)
    return outs


def _upload_bands(out, spectra):
    """Row bands (nt, n, 11) of one camera's resolution matrices into the device
    view `out`.

    Spectra whose R are dia_matrix with the same offsets and data shape (the
    reference's layout, every DESI coadd) are uploaded as their diagonal data
    and turned into row bands on the GPU, one slice per diagonal for all
    targets: row p of the diagonal at offset `off` holds R[p, p + off] =
    data[d, p + off] (column-indexed).  Others go one by one through
    Spectrum.rband() on the host.
    """
    import scipy.sparse
    import torch
    from .targets import RBAND_HALF
    Rs = [s.R for s in spectra]
    first = Rs[0]
    same = (all(scipy.sparse.issparse(R) and R.format == 'dia' for R in Rs)
            and all(R.shape == first.shape and R.data.shape == first.data.shape for R in Rs))
    if same:            # every offsets array as one (nt, ndiag) comparison
        same = bool(np.all(np.array([R.offsets for R in Rs]) == np.asarray(first.offsets)[None, :]))
    nt, n = out.shape[0], out.shape[1]
    if not same:
        hn = np.empty((nt, n, 11))
        for i, s in enumerate(spectra):
            hn[i] = s.rband()
        out.copy_(torch.from_numpy(hn))
        return
    nd, L = first.data.shape
    d = torch.empty((nt, nd, L), dtype=torch.float64, device=out.device)
    _upload_rows(d, [R.data for R in Rs])
    nrow, ncol = first.shape
    L = min(ncol, L)
    for di, o in enumerate(np.asarray(first.offsets).astype(np.int64)):
        p0, p1 = max(0, -o), min(n, nrow, L - o)
        if p1 <= p0:
            continue
        vals = d[:, di, p0 + o:p1 + o]
        if abs(o) > RBAND_HALF:
            if bool(torch.any(vals != 0)):
                raise NotImplementedError('resolution matrix wider than 11 diagonals')
            continue
        out[:, p0:p1, o + RBAND_HALF] += vals


def _dev(a, dtype):
    """A host array as a new device tensor, copied without waiting for the
    stream's earlier work: staged in page-locked memory (torch's pinned
    allocator keeps the block until the copy has run)."""
    import torch
    t = torch.from_numpy(np.ascontiguousarray(a, dtype=dtype))
    if t.numel() == 0:
        return t.to('cuda')
    return t.pin_memory().to('cuda', non_blocking=True)


def target_runs(item_target):
    """Consecutive entries of one target as one scan item.

    Returns (targets, z0, nz): entry k of a run is redshift z0 + k of the
    cache, so per-entry redshift lists become multi-z items (the FP64 kernel
    puts a target's redshifts on the lanes of one wavefront) and the output
    rows keep the entry order.
    """
    t = np.asarray(item_target, dtype=np.int64)
    if t.size == 0:
        return t.astype(np.int32), np.zeros(0, np.int32), np.zeros(0, np.int32)
    start = np.flatnonzero(np.r_[True, t[1:] != t[:-1]])
    nz = np.diff(np.r_[start, t.size])
    return t[start].astype(np.int32), start.astype(np.int32), nz.astype(np.int32)


class ScanPlan(object):
    """A prepared rr_zscan launch: descriptor, outputs and workspace in HBM.

    items: (target_slot, cache_z0, nz) per item; output rows of item i are
    item_out[i] .. item_out[i]+nz-1 in zchi2/zcoeff/penalty/status.
    """

    def __init__(self, cache, batch, item_target, item_z0, item_nz, precision='f64',
                 prior=None, oiisum=None, tile=0, polish=None, polish_windows=POLISH_WINDOWS,
                 solve='PCA', bounds=None, lanes=None, resid=None):
        """polish: FP64 re-evaluation of the rows near each item's minimum after a
        mixed-precision scan (rr_zscan_polish), see polish_tol(); it needs the
        FP64 template cache as well.  solve: the template's method, 'PCA' or
        'NMF' / 'NNLS' (non-negative coefficients, no polish), or 'BVLS' with
        `bounds` (solve_spec).  lanes: FP64 only, scan with one item per lane
        (items sharing one redshift range); None = when there are at least
        LANE_MIN_ITEMS such items ('f16r' always needs them).  resid ('f16r'):
        inputs of the explicit-residual stage to build, 'mfma' (the
        matrix-core kernel, the default), 'lanes' (the FP32 lane kernel) or
        'both' (every configuration of rr_zscan_resid_tiles can run)."""
        import torch
        if precision not in _hip.PRECISIONS:
            raise ValueError(f'unknown precision {precision!r}')
        self.tile = tile
        dev = torch.device('cuda')
        nb = cache.nb
        if nb > _hip.RR_MAX_NBASIS:
            raise NotImplementedError(f'nbasis={nb} > {_hip.RR_MAX_NBASIS}')
        layout = batch.layout
        item_target = np.asarray(item_target, dtype=np.int32)
        item_z0 = np.asarray(item_z0, dtype=np.int32)
        item_nz = np.asarray(item_nz, dtype=np.int32)
        uniform = bool(len(item_nz) > 0 and np.all(item_z0 == item_z0[0])
                       and np.all(item_nz == item_nz[0]))
        if precision == 'f16r':
            if len(item_nz) and not uniform:
                raise NotImplementedError("precision 'f16r' needs items that share one redshift range")
            lanes = uniform
        elif precision == 'f64':
            if lanes is None:
                lanes = uniform and len(item_nz) >= LANE_MIN_ITEMS
            elif lanes and not uniform:
                raise ValueError('the lane-per-item FP64 scan needs items sharing one redshift range')
        else:
            lanes = False
        self.lanes = bool(lanes)
        if precision == 'f64' and self.lanes:
            tbuf, ldz = None, 0
        else:
            tbuf, ldz = cache.get(layout, 'f16' if precision == 'f16r' else precision)
        if np.any(item_z0 < 0) or np.any(item_z0 + item_nz > cache.nz):
            raise ValueError('item redshift range outside the template cache')
        if np.any(item_target < 0) or np.any(item_target >= batch.nt):
            raise ValueError('item target outside the batch')
        item_out = np.concatenate([[0], np.cumsum(item_nz.astype(np.int64))])[:-1]
        rows = int(item_nz.sum())
        self.nitems = len(item_nz)
        self.rows = rows
        self.nb = nb
        self.item_out = item_out
        self.item_nz = item_nz
        self._keep = [tbuf, batch]
        self.t_item_target = _dev(item_target, np.int32)
        self.t_item_z0 = _dev(item_z0, np.int32)
        self.t_item_nz = _dev(item_nz, np.int32)
        self.t_item_out = _dev(item_out, np.int64)
        self.zchi2 = torch.empty(rows, dtype=torch.float64, device=dev)
        self.zcoeff = torch.empty((rows, nb), dtype=torch.float64, device=dev)
        self.penalty = torch.empty(rows, dtype=torch.float64, device=dev)
        self.status = torch.empty(rows, dtype=torch.uint8, device=dev)
        self.prior = None if prior is None else _dev(prior, np.float64)
        self.oiisum = None if oiisum is None else _dev(oiisum, np.float64)
        d = _hip.ZscanDesc()
        d.nitems = self.nitems
        d.npix = layout.npix
        d.npix_pad = layout.npix_pad
        d.nb = nb
        d.precision = _hip.PRECISIONS[precision]
        d.halo = layout.halo
        d.d_tcache = None if tbuf is None else tbuf.data_ptr()
        d.cache_rows = layout.cache_rows
        d.cache_ldz = ldz
        d.cache_nz = cache.nz
        d.tscale = cache.tscale
        d.d_rband = batch.rband.data_ptr()
        d.d_w = batch.w.data_ptr()
        d.d_wf = batch.wf.data_ptr()
        d.d_f = batch.f.data_ptr()
        if precision != 'f64':
            rtile, sfcol, scales = batch.mixed(precision)
            d.d_rtile, d.d_scales = rtile.data_ptr(), scales.data_ptr()
            d.d_sfcol = None if sfcol is None else sfcol.data_ptr()
            self._keep += [rtile, sfcol, scales]
        resid = RESID_DEFAULT if resid is None else resid
        if resid not in ('mfma', 'lanes', 'both'):
            raise ValueError(f"resid must be 'mfma', 'lanes' or 'both', not {resid!r}")
        self.resid = resid if precision == 'f16r' else None
        if precision == 'f16r' and resid in ('mfma', 'both') and self.nitems > 0:
            tu, _ = cache.get(layout, 'f16u')
            # the kernel reads (cache_nz, cache_rows, KS) fp16 elements
            need = cache.nz * layout.cache_rows * _hip.resid_u_slots(nb)[0]
            if tu.dtype != torch.float16 or tu.numel() != need:
                raise ValueError(f'template operand cache holds {tu.numel()} {tu.dtype} elements, '
                                 f'the residual kernel reads {need} fp16')
            d.d_tcache_u = tu.data_ptr()
            self._keep.append(tu)
        lane_in = self.lanes and (precision == 'f64' or resid in ('lanes', 'both'))
        if lane_in and self.nitems > 0:
            tr, _ = cache.get(layout, 'f64r' if precision == 'f64' else 'f32r')
            identity = np.array_equal(item_target, np.arange(len(item_target)))
            rl = batch.lanes(precision, None if identity else item_target)
            d.d_tcache_r, d.d_rlane = tr.data_ptr(), rl.data_ptr()
            self._keep += [tr, rl]
        d.tile = int(tile)
        d.uniform_z = int(uniform)
        d.d_item_target = self.t_item_target.data_ptr()
        d.d_item_z0 = self.t_item_z0.data_ptr()
        d.d_item_nz = self.t_item_nz.data_ptr()
        d.d_item_out = self.t_item_out.data_ptr()
        d.d_prior = None if self.prior is None else self.prior.data_ptr()
        d.d_oiisum = None if self.oiisum is None else self.oiisum.data_ptr()
        d.d_zchi2 = self.zchi2.data_ptr()
        d.d_zcoeff = self.zcoeff.data_ptr()
        d.d_penalty = self.penalty.data_ptr()
        d.d_status = self.status.data_ptr()
        d.max_item_nz = int(item_nz.max()) if self.nitems else 0
        d.total_rows = rows
        d.solve, d.nonneg_mask, bnd = solve_spec(solve, nb, bounds)
        if bnd is not None:
            self._bounds = _dev(bnd, np.float64)
            d.d_bounds = self._bounds.data_ptr()
        nbytes = _hip.lib().rr_zscan_workspace_bytes(ctypes.byref(d))
        if nbytes < 0:
            _hip.check(int(nbytes), 'rr_zscan_workspace_bytes')
        self.work = torch.empty(max(int(nbytes), 256), dtype=torch.uint8, device=dev)
        d.d_work = self.work.data_ptr()
        d.work_bytes = self.work.numel()
        self.desc = d
        self.precision = precision
        self.polish = None
        # why a mixed-precision scan runs without its FP64 polish (None: it has
        # one, or the scan is FP64 or the polish was turned off)
        self.unpolished = None
        tol = polish_tol(precision, polish)
        if tol is not None and d.solve != _hip.RR_SOLVE_PCA:
            self.unpolished = 'the FP64 polish re-solves M c = y: not for bounded (NNLS / BVLS) solves'
        elif tol is not None and self.nitems > 0 and int(item_nz.max()) > _hip.RR_POLISH_MAX_NZ:
            self.unpolished = f'items of more than {_hip.RR_POLISH_MAX_NZ} redshifts'
        if self.unpolished is not None:
            if polish is not None:           # asked for explicitly
                warnings.warn(f'ScanPlan: no FP64 polish ({self.unpolished})', stacklevel=2)
            tol = None
        if tol is not None and self.nitems > 0:
            if not 1 <= polish_windows <= _hip.RR_POLISH_MAX_WINDOWS:
                raise ValueError(f'polish_windows must be in 1..{_hip.RR_POLISH_MAX_WINDOWS}')
            t64, ldz64 = cache.get(layout, 'f64')
            self.windows = torch.zeros((self.nitems, polish_windows), dtype=torch.int32, device=dev)
            self.nwindows = torch.zeros(self.nitems, dtype=torch.int32, device=dev)
            self.queue = torch.zeros(1 + self.nitems * polish_windows, dtype=torch.int32, device=dev)
            pd = _hip.PolishDesc()
            pd.d_tcache64 = t64.data_ptr()
            pd.cache_ldz = ldz64
            pd.tol = tol
            pd.max_windows = polish_windows
            pd.d_windows = self.windows.data_ptr()
            pd.d_nwindows = self.nwindows.data_ptr()
            pd.d_queue = self.queue.data_ptr()
            self._keep.append(t64)
            self.polish = pd

    def run(self, stream=None, stages=0):
        """Enqueue the scan on `stream` (default: torch's current stream).

        stages: RR_STAGE_* and STAGE_POLISH bits (0 = stats, main, finish, the
        explicit residual of 'f16r' and, when the plan has one, the polish).
        """
        scan = stages & (_hip.RR_STAGE_STATS | _hip.RR_STAGE_MAIN | _hip.RR_STAGE_FINISH
                         | _hip.RR_STAGE_RESID)
        h = _hip.stream_handle(stream)
        if stages == 0 or scan:
            self.desc.stages = scan
            rc = _hip.lib().rr_zscan(ctypes.byref(self.desc), h)
            _hip.check(rc, 'rr_zscan')
        if self.polish is not None and (stages == 0 or stages & STAGE_POLISH):
            rc = _hip.lib().rr_zscan_polish(ctypes.byref(self.desc), ctypes.byref(self.polish), h)
            _hip.check(rc, 'rr_zscan_polish')
        return self

    def host(self, stream=None):
        """(zchi2, zcoeff, penalty, status) copied to new numpy arrays
        (download(): ordered after the work already on `stream`, default the
        current stream)."""
        return tuple(download((self.zchi2, self.zcoeff, self.penalty, self.status), stream))


# Inputs a 'f16r' plan builds for its explicit-residual stage (ScanPlan resid=):
# the matrix-core kernel's (split-fp16 template operand cache and R tiles)
RESID_DEFAULT = 'mfma'

# Smallest item count for which an FP64 plan of items sharing one redshift
# range takes the lane-per-item kernel (fewer items would leave lanes idle;
# the redshift-per-lane kernel serves them).
LANE_MIN_ITEMS = 32


def _method(cache):
    """solve_matrices_algorithm of the template behind a cache ('PCA' if none)."""
    tm = getattr(cache, 'template', None)
    return getattr(tm, 'solve_matrices_algorithm', 'PCA') if tm is not None else 'PCA'


def _oiisum(template):
    if hasattr(template, 'OIItemplate'):
        return np.asarray(template.OIItemplate, dtype=np.float64).sum(axis=1)
    return None


# template full_type -> why its last mixed-precision scan ran without the FP64
# polish (scan_targets fills it; zfind reports it in last_timings)
unpolished = dict()


def scan_targets(cache, targets, precision=None, oiisum=None, batches=None, polish=None,
                 solve=None):
    """Scan every target over every redshift of `cache`.

    Targets are grouped by wavelength layout (one launch per layout).
    polish: FP64 polish of a mixed-precision scan (ScanPlan); solve: the
    template's method (default: cache.template's).
    Returns numpy (zchi2 (nt, nz), zcoeff (nt, nz, nb), penalty (nt, nz)).
    """
    precision = precision or default_precision()
    nt, nz, nb = len(targets), cache.nz, cache.nb
    zchi2 = np.zeros((nt, nz))
    zcoeff = np.zeros((nt, nz, nb))
    penalty = np.zeros((nt, nz))
    groups = dict()
    for i, tg in enumerate(targets):
        groups.setdefault(PixelLayout.signature_of(tg.spectra), []).append(i)
    for sig, idx in groups.items():
        batch = None if batches is None else batches.get(sig)
        if batch is None:
            batch = TargetBatch([targets[i] for i in idx])
        n = len(idx)
        plan = ScanPlan(cache, batch, np.arange(n), np.zeros(n), np.full(n, nz), precision,
                        oiisum=oiisum, polish=polish, solve=solve or _method(cache))
        if plan.unpolished is not None:
            tm = getattr(cache, 'template', None)
            unpolished[getattr(tm, 'full_type', '?')] = plan.unpolished
        plan.run()
        c2, cf, pen, _ = plan.host()
        if len(groups) == 1:                 # one layout: the plan's rows are the targets in order
            return c2.reshape(n, nz), cf.reshape(n, nz, nb), pen.reshape(n, nz)
        zchi2[idx] = c2.reshape(n, nz)
        zcoeff[idx] = cf.reshape(n, nz, nb)
        penalty[idx] = pen.reshape(n, nz)
    return zchi2, zcoeff, penalty


class ScanGraph(object):
    """Every target x every redshift x every template class in one graph launch.

    One ScanPlan per template (its own nbasis, redshift grid and tile
    configuration; the kernels are specialised per nbasis) over one
    TargetBatch.  capture() records the stats / main / finish stages of all
    plans into a single HIP graph; replay() re-runs the whole batched scan with
    one launch and no host work.  Targets with masked pixels need nothing
    special: a zero weight removes the pixel from every sum.

    This is the batched form of calc_zchi2_targets (zscan.py:910-1105) for
    targets that share one wavelength layout.
    """

    def __init__(self, caches, batch, precision=None, tune=False, polish=None):
        """caches: list of (TemplateCache, oiisum or None); tune: autotune each plan's
        tile; polish: FP64 polish of each plan (ScanPlan), captured with it."""
        precision = precision or default_precision()
        self.precision = precision
        self.batch = batch
        nt = batch.nt
        self.plans = []
        for cache, oii in caches:
            plan = ScanPlan(cache, batch, np.arange(nt), np.zeros(nt), np.full(nt, cache.nz),
                            precision, oiisum=oii, polish=polish, solve=_method(cache))
            if tune:
                from . import autotune
                autotune.choose(plan)
            self.plans.append(plan)
        self.graph = None
        self.stream = None

    def run(self):
        """Eager: enqueue every plan on the current stream."""
        for p in self.plans:
            p.run()
        return self

    def capture(self):
        """Record all plans into one HIP graph (after one eager warm-up run)."""
        import torch
        self.run()
        torch.cuda.synchronize()
        self.stream = torch.cuda.Stream()
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph, stream=self.stream):
            for p in self.plans:
                p.run()
        torch.cuda.synchronize()
        return self

    def replay(self):
        if self.graph is None:
            raise RuntimeError('ScanGraph.replay() before capture()')
        self.graph.replay()
        return self

    def results(self):
        """[(zchi2 (nt, nz), zcoeff (nt, nz, nb), penalty (nt, nz)), ...] per template, numpy."""
        nt = self.batch.nt
        out = []
        for p in self.plans:
            c2, cf, pen, _ = p.host()
            nz = len(c2) // nt if nt else 0
            out.append((c2.reshape(nt, nz), cf.reshape(nt, nz, p.nb), pen.reshape(nt, nz)))
        return out


def calc_zchi2(target_ids, target_data, dtemplate, progress=None, use_gpu=False, precision=None,
               batches=None, polish=None):
    """chi2 vs redshift of every target for one template (zscan.py:807-889).

    batches: {layout signature: TargetBatch} of these targets (target_batches),
    reused instead of uploading the targets again.  polish: FP64 polish of a
    mixed-precision scan (ScanPlan; on by default for mixed precisions).
    Returns (zchi2[nt, nz], zcoeff[nt, nz, nbasis], zchi2penalty[nt, nz]).
    """
    cache = dtemplate.cache
    oii = _oiisum(dtemplate.template)
    zchi2, zcoeff, penalty = scan_targets(cache, list(target_data), precision, oii, batches=batches,
                                          polish=polish)
    if progress is not None and dtemplate.comm is None:
        for _ in range(len(target_data)):
            progress.put(1)
    return zchi2, zcoeff, penalty


def calc_zchi2_batch(spectra, tdata, weights, flux, wflux, nz, nbasis, solve_matrices_algorithm=None,
                     solver_args=None, use_gpu=False, fullprecision=True, prior=None,
                     precision='f64', polish=None):
    """chi2 and coefficients of one target at every redshift of `tdata` (zscan.py:626-781).

    tdata: {wavehash: (nz, nwave, nbasis)} numpy arrays or device tensors.
    polish: FP64 polish of a mixed-precision scan (ScanPlan).
    solve_matrices_algorithm: 'PCA' (default), 'NMF' / 'NNLS' (coefficients
    >= 0, scipy.optimize.nnls(M, y) in the reference, zscan.py:1153-1172) or
    'BVLS' with solver_args['bounds'] (any bounds, scipy lsq_linear(M, y,
    bounds, 'bvls') in the reference, zscan.py:1175-1201; solve_spec).
    Returns numpy (zchi2 (nz,), zcoeff (nz, nbasis)).
    """
    import torch
    w = _host(weights)
    zchi2 = np.zeros(nz)
    if w.sum() == 0:
        zchi2[:] = HUGE_CHI2
        return zchi2, np.zeros((nz, nbasis))
    algo = 'PCA' if solve_matrices_algorithm is None else solve_matrices_algorithm.upper()
    bounds = None
    if algo == 'BVLS' and solver_args is not None and 'bounds' in solver_args:
        bounds = solver_args['bounds']                          # zscan.py:1175-1181
    if algo not in SOLVES:
        # solve_matrices raises NotImplementedError, caught by the caller (zscan.py:721-724)
        zchi2[:] = HUGE_CHI2
        return zchi2, np.zeros((nz, nbasis))
    layout = PixelLayout.of(spectra)
    P = layout.npix_pad
    rband = np.zeros((1, P, 11))
    for s, off in zip(spectra, layout.offsets):
        rband[0, off:off + s.nwave] = s.rband()
    wa = np.zeros((1, P))
    fa = np.zeros((1, P))
    wfa = np.zeros((1, P))
    wa[0, :layout.npix] = w
    fa[0, :layout.npix] = _host(flux)
    wfa[0, :layout.npix] = _host(wflux)
    batch = TargetBatch(arrays=(layout, rband, wa, fa, wfa))
    cache = _ExplicitCache(tdata, layout, nz, nbasis)
    plan = ScanPlan(cache, batch, [0], [0], [nz], precision, prior=prior, polish=polish, solve=algo,
                    bounds=bounds)
    plan.run()
    c2, cf, _, _ = plan.host()
    return c2, cf


def _u_operand(full):
    """The template operand cache of the matrix-core residual from a float32
    (nz, cache_rows, nb) cache (values * tscale): flat fp16 (nz, cache_rows,
    KS), K slots hi | lo | hi [| lo] per basis vector (rr_resid_u_slots)."""
    import torch
    nz, rows, nb = full.shape
    ks, nprod = _hip.resid_u_slots(nb)
    hi = full.half()
    lo = (full - hi.float()).half()
    buf = torch.zeros((nz, rows, ks), dtype=torch.float16, device=full.device)
    for j, part in enumerate((hi, lo, hi, lo)[:nprod]):
        buf[:, :, j * nb:(j + 1) * nb] = part
    return buf.reshape(-1)


def _host(a):
    if hasattr(a, 'detach'):
        return a.detach().cpu().numpy()
    return np.asarray(a)


def _split_planes(full):
    """(nz, nb, rows) float32 -> fp16 [nz][hi, lo][nb][rows] as int16, + 64 zeros."""
    import torch
    hi = full.half()
    lo = (full - hi.float()).half()
    planes = torch.stack([hi, lo], dim=1).view(torch.int16).reshape(-1)
    return torch.cat([planes, torch.zeros(64, dtype=torch.int16, device=full.device)])


class _ExplicitCache(object):
    """Template cache assembled from an explicit tdata dict (calc_zchi2_batch)."""

    def __init__(self, tdata, layout, nz, nb):
        self.tdata = tdata
        self.nz = nz
        self.nb = nb
        vals = [_host(v) for v in tdata.values()]
        fmax = max(float(np.max(np.abs(v))) if v.size else 0.0 for v in vals)
        self.tscale = 2.0 ** (8 - math.ceil(math.log2(fmax))) if fmax > 0 else 1.0
        self._bufs = dict()

    def get(self, layout, precision):
        import torch
        key = (layout.signature, precision)
        if key in self._bufs:
            return self._bufs[key]
        nb, nz, rows = self.nb, self.nz, layout.cache_rows
        if precision == 'f16u':
            ldz = 0
            full = torch.zeros((nz, rows, nb), dtype=torch.float32, device='cuda')
            for k, off, n in zip(layout.keys, layout.offsets, layout.sizes):
                t = torch.as_tensor(self.tdata[k], dtype=torch.float64, device='cuda')
                full[:, layout.halo + off:layout.halo + off + n, :] = (t * self.tscale).float()
            buf = _u_operand(full)
        elif precision in ('f64r', 'f32r'):
            ldz = 0
            dt = torch.float64 if precision == 'f64r' else torch.float32
            buf = torch.zeros((nz, rows, nb), dtype=dt, device='cuda')
            for k, off, n in zip(layout.keys, layout.offsets, layout.sizes):
                t = torch.as_tensor(self.tdata[k], dtype=torch.float64, device='cuda')
                buf[:, layout.halo + off:layout.halo + off + n, :] = t.to(dt)
            buf = buf.reshape(-1)
        elif precision == 'f64':
            ldz = _roundup(max(nz, 1), 64)
            buf = torch.zeros((nb, rows, ldz), dtype=torch.float64, device='cuda')
            for k, off, n in zip(layout.keys, layout.offsets, layout.sizes):
                t = torch.as_tensor(self.tdata[k], dtype=torch.float64, device='cuda')
                buf[:, layout.halo + off:layout.halo + off + n, :nz] = t.permute(2, 1, 0)
            buf = buf.reshape(-1)
        else:
            ldz = 0
            buf = torch.zeros((nz, nb, rows), dtype=torch.float32, device='cuda')
            for k, off, n in zip(layout.keys, layout.offsets, layout.sizes):
                t = torch.as_tensor(self.tdata[k], dtype=torch.float64, device='cuda')
                buf[:, :, layout.halo + off:layout.halo + off + n] = (
                    t.permute(0, 2, 1) * self.tscale).float()
            if precision == 'f32':
                buf = torch.cat([buf.reshape(-1), torch.zeros(64, dtype=torch.float32, device='cuda')])
            elif precision == 'f16x3':
                buf = _split_planes(buf)
            elif precision in ('f16', 'bf16'):
                dt = torch.float16 if precision == 'f16' else torch.bfloat16
                buf = torch.cat([buf.to(dt).view(torch.int16).reshape(-1),
                                 torch.zeros(64, dtype=torch.int16, device='cuda')])
            else:
                raise ValueError(f'no template cache layout {precision!r}')
        self._bufs[key] = (buf, ldz)
        return self._bufs[key]


def target_batches(targets):
    """{layout signature: TargetBatch} of a target list, in list order per layout
    (the grouping scan_targets uses)."""
    groups = dict()
    for tg in targets:
        groups.setdefault(PixelLayout.signature_of(tg.spectra), []).append(tg)
    return {sig: TargetBatch(tgs, layout=PixelLayout(*sig), grouped=True) for sig, tgs in groups.items()}


def calc_zchi2_targets(targets, templates, mp_procs=1, use_gpu=False, precision=None, batches=None,
                       polish=None):
    """Coarse chi2 scan of every local target for every template (zscan.py:910-1105).

    batches: target_batches(targets.local()) to reuse (built here otherwise,
    once for all templates).
    Returns results[targetid][fulltype] = dict(redshifts, zchi2, penalty, zcoeff).
    """
    return _scan_templates(targets, templates, precision, batches, polish)[0]


def _scan_templates(targets, templates, precision=None, batches=None, polish=None):
    """calc_zchi2_targets' results and, per template full_type, the (nt, nz)
    arrays zchi2 + penalty of the local targets in order (what zfind's fine
    fits start from), without splitting and re-stacking them per target.

    With one wavelength layout every template's scan is enqueued first and
    each one's results are copied to the host while the later scans run
    (scan_templates_iter)."""
    results = {tid: dict() for tid in targets.local_target_ids()}
    full = dict()
    local = targets.local()
    if batches is None and local:
        batches = target_batches(local)
    if local and len(batches) == 1:
        return _scan_templates_pipelined(results, local, templates, precision,
                                         next(iter(batches.values())), polish)
    for t in templates:
        ft = t.template.full_type
        zchi2, zcoeff, penalty = calc_zchi2(targets.local_target_ids(), local, t,
                                            precision=precision, batches=batches, polish=polish)
        full[ft] = zchi2 + penalty
        for i, tg in enumerate(local):
            results[tg.id][ft] = dict(redshifts=t.template.redshifts, zchi2=zchi2[i],
                                      penalty=penalty[i], zcoeff=zcoeff[i])
    return results, full


_PIPE = []


def pipeline_streams():
    """(scan, copy, fit) streams of the pipelined template scans and zfind's
    fine fits, taken once per process.  Streams share a few hardware queues,
    and work on two streams of one queue runs in order: three consecutive
    streams of torch's pool are on three queues (measured), so results are
    copied and fine fits run beside the scans.  Taken per call instead, a
    stream can land on the queue of the default stream or of another."""
    import torch
    if not _PIPE:
        _PIPE.extend(torch.cuda.Stream() for _ in range(3))
    return tuple(_PIPE)


def _plan_bytes(rows, nb):
    """Device bytes a full-grid ScanPlan of `rows` (target, z) rows holds:
    chi2, coefficients, penalty, status and the normal-equation workspace
    (upper bound: the mixed-precision column count)."""
    ne = nb * (nb + 2) - nb * (nb - 1) // 2
    return int(rows) * (8 * (nb + 2) + 1 + 8 * ne)


def _scan_templates_pipelined(results, local, templates, precision, batch, polish):
    full = dict()
    for t, zchi2, zcoeff, penalty in scan_templates_iter(local, templates, precision, batch, polish):
        ft = t.template.full_type
        full[ft] = zchi2 + penalty
        for i, tg in enumerate(local):
            results[tg.id][ft] = dict(redshifts=t.template.redshifts, zchi2=zchi2[i],
                                      penalty=penalty[i], zcoeff=zcoeff[i])
    return results, full


def scan_templates_iter(local, templates, precision, batch, polish=None):
    """Scan the targets of one TargetBatch (`local`, in batch order) for every
    template; yields (template, zchi2 (nt, nz), zcoeff (nt, nz, nb),
    penalty (nt, nz)) in template order, as numpy arrays (download()).

    Every template's scan is enqueued on the scan stream of
    pipeline_streams() (after the work already on the current stream) before
    the first result is handed out (a template's plan is unpacked earlier
    only when the next one would not fit in free HBM); a scan's results are
    copied to the host on the copy stream when they are handed out, while the
    later templates' scans run.  When the iterator ends the current stream
    waits for both."""
    import torch
    precision = precision or default_precision()
    nt = len(local)
    caller = torch.cuda.current_stream()
    main, copy, _ = pipeline_streams()
    main.wait_stream(caller)          # the target upload and anything else before
    pending = []

    def unpack(entry):
        t, plan, ran, nz = entry
        copy.wait_event(ran)
        zchi2, zcoeff, penalty = download((plan.zchi2, plan.zcoeff, plan.penalty), copy)
        return (t, zchi2.reshape(nt, nz), zcoeff.reshape(nt, nz, plan.nb), penalty.reshape(nt, nz))

    for t in templates:
        cache = t.cache
        nz = cache.nz
        # a plan holds its outputs and normal-equation workspace until it is
        # unpacked: unpack the pending ones first when the next would not fit
        need = _plan_bytes(nt * nz, cache.nb)
        while pending and torch.cuda.mem_get_info()[0] < 2 * need:
            yield unpack(pending.pop(0))
        with torch.cuda.stream(main):
            plan = ScanPlan(cache, batch, np.arange(nt), np.zeros(nt), np.full(nt, nz), precision,
                            oiisum=_oiisum(t.template), polish=polish, solve=_method(cache))
            if plan.unpolished is not None:
                unpolished[t.template.full_type] = plan.unpolished
            plan.run(main)
        ran = torch.cuda.Event()
        ran.record(main)
        pending.append((t, plan, ran, nz))
    while pending:
        yield unpack(pending.pop(0))
    caller.wait_stream(main)
    caller.wait_stream(copy)
