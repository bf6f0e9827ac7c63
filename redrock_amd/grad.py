"""Gradients of chi2(z) and the Newton refinement that stands beside fitz's parabola.

chi2(z) = min_c |W^1/2 (f - R T(z) c)|^2 for one target and one template.
rr_zscan_grad (csrc/zscan_grad.hip) returns d chi2/dz and d chi2/d(template
flux) at the best-fit coefficients, recomputing the rebin instead of storing
it.  Chi2AtRedshift wraps forward (rr_rebin + rr_zscan in FP64) and backward
(rr_zscan_grad) as a torch.autograd.Function.  newton_refine() drives z to
the stationary point of chi2 using the analytic gradient and a central
difference of it for the curvature; fitz_targets(..., newton=True) reports its
result beside the reference's parabola fit (fitz.py:259-261).
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _hip
from .igm import igm_applies, lyman_table
from .rebin import centers2edges, lyman_array
from .zscan import PixelLayout, ScanPlan, TargetBatch, TemplateCache, target_runs

_vp, _i, _d = ctypes.c_void_p, ctypes.c_int, ctypes.c_double


class GradDesc(ctypes.Structure):
    _fields_ = [
        ('nitems', _i), ('npix', _i), ('npix_pad', _i), ('nb', _i),
        ('d_x', _vp), ('nx', _i), ('d_y', _vp), ('d_pe_lo', _vp), ('d_pe_hi', _vp),
        ('d_pwave', _vp), ('lines', ctypes.POINTER(_hip.LymanLine)), ('nlines', _i),
        ('d_rband', _vp), ('d_w', _vp), ('d_f', _vp),
        ('d_item_target', _vp), ('d_item_z', _vp), ('d_item_coeff', _vp),
        ('d_dchi2_dz', _vp), ('d_chi2', _vp), ('d_dchi2_dy', _vp), ('d_item_scale', _vp),
        ('d_work', _vp), ('work_bytes', ctypes.c_longlong),
    ]


def _lib():
    L = _hip.lib()
    if not hasattr(L, '_grad_ready'):
        L.rr_zscan_grad.restype = _i
        L.rr_zscan_grad.argtypes = [ctypes.POINTER(GradDesc), _vp]
        L.rr_grad_workspace_bytes.restype = ctypes.c_longlong
        L.rr_grad_workspace_bytes.argtypes = [ctypes.POINTER(GradDesc)]
        L._grad_ready = True
    return L


class GradProblem(object):
    """Targets x one template, evaluated at per-item redshifts.

    items: item_target[i] is the slot (in `targets`) of item i.
    """

    def __init__(self, template, targets, item_target, batch=None):
        """batch: an existing TargetBatch of exactly `targets` (reused, not rebuilt)."""
        import torch
        self.template = template
        self.targets = list(targets)
        self.item_target = np.asarray(item_target, dtype=np.int32)
        self.layout = PixelLayout.of(self.targets[0].spectra)
        if batch is not None and (batch.nt != len(self.targets)
                                  or batch.layout.signature != self.layout.signature):
            raise ValueError('batch does not hold these targets')
        self.batch = batch if batch is not None else TargetBatch(self.targets, self.layout)
        self.dwave = {s.wavehash: s.wave for s in self.targets[0].spectra}
        dev = torch.device('cuda')
        lo, hi, wc = [], [], []
        for k in self.layout.keys:
            e = centers2edges(self.dwave[k])
            lo.append(e[:-1])
            hi.append(e[1:])
            wc.append(np.asarray(self.dwave[k], dtype=np.float64))
        self.pe_lo = torch.as_tensor(np.concatenate(lo), device=dev).contiguous()
        self.pe_hi = torch.as_tensor(np.concatenate(hi), device=dev).contiguous()
        self.pwave = torch.as_tensor(np.concatenate(wc), device=dev).contiguous()
        self.x = torch.as_tensor(np.ascontiguousarray(template.wave, dtype=np.float64),
                                 device=dev).contiguous()
        self.t_item_target = torch.as_tensor(self.item_target, device=dev)
        model = template.igm_model
        self.lines = None if model is None or model == 'None' else lyman_table(model)

    def forward(self, z, flux=None):
        """FP64 chi2 and coefficients of every item at z (host or device array).

        Returns device tensors (chi2 (n,), coeff (n, nb)).
        """
        zh = z.detach().cpu().numpy() if hasattr(z, 'detach') else np.asarray(z, dtype=np.float64)
        cache = TemplateCache(self.template, self.dwave, redshifts=zh, flux=flux)
        # consecutive entries of one target share a scan item (one z per lane)
        plan = ScanPlan(cache, self.batch, *target_runs(self.item_target), 'f64',
                        solve=self.template.solve_matrices_algorithm)
        plan.run()
        return plan.zchi2, plan.zcoeff

    def grad(self, z, coeff, flux=None, want_dy=False, scale=None):
        """(dchi2/dz (n,), explicit chi2 (n,), dchi2/dflux (nb, nx) or None) as device tensors.

        scale (n,) weights the per-item template-flux gradients before summing.
        """
        import torch
        dev = torch.device('cuda')
        n = len(self.item_target)
        zt = torch.as_tensor(z, dtype=torch.float64, device=dev).contiguous()
        y = (flux.detach() if flux is not None else
             torch.as_tensor(np.ascontiguousarray(self.template.flux), device=dev)).contiguous()
        nb, nx = y.shape
        dz = torch.empty(n, dtype=torch.float64, device=dev)
        chi2 = torch.empty(n, dtype=torch.float64, device=dev)
        dy = torch.zeros((nb, nx), dtype=torch.float64, device=dev) if want_dy else None
        sc = None if scale is None else torch.as_tensor(scale, dtype=torch.float64,
                                                        device=dev).contiguous()
        coeff = coeff.contiguous()
        lines, nlines = lyman_array(self.lines)
        d = GradDesc()
        d.nitems, d.npix, d.npix_pad, d.nb = n, self.layout.npix, self.layout.npix_pad, nb
        d.d_x, d.nx, d.d_y = self.x.data_ptr(), nx, y.data_ptr()
        d.d_pe_lo, d.d_pe_hi, d.d_pwave = self.pe_lo.data_ptr(), self.pe_hi.data_ptr(), self.pwave.data_ptr()
        d.lines, d.nlines = lines, nlines
        d.d_rband, d.d_w, d.d_f = (self.batch.rband.data_ptr(), self.batch.w.data_ptr(),
                                   self.batch.f.data_ptr())
        d.d_item_target = self.t_item_target.data_ptr()
        d.d_item_z, d.d_item_coeff = zt.data_ptr(), coeff.data_ptr()
        d.d_dchi2_dz, d.d_chi2 = dz.data_ptr(), chi2.data_ptr()
        d.d_dchi2_dy = None if dy is None else dy.data_ptr()
        d.d_item_scale = None if sc is None else sc.data_ptr()
        work = None
        if dy is not None:
            nbytes = _lib().rr_grad_workspace_bytes(ctypes.byref(d))
            if nbytes < 0:
                _hip.check(int(nbytes), 'rr_grad_workspace_bytes')
            work = torch.empty(max(int(nbytes), 256), dtype=torch.uint8, device=dev)
            d.d_work, d.work_bytes = work.data_ptr(), work.numel()
        rc = _lib().rr_zscan_grad(ctypes.byref(d), _hip.stream_handle())
        _hip.check(rc, 'rr_zscan_grad')
        return dz, chi2, dy


def _autograd_function():
    import torch

    class Chi2AtRedshift(torch.autograd.Function):
        """chi2 of every item at its redshift; differentiable in z and template flux."""

        @staticmethod
        def forward(ctx, z, flux, problem):
            chi2, coeff = problem.forward(z, flux)
            ctx.problem = problem
            ctx.save_for_backward(z, flux, coeff)
            return chi2.clone()

        @staticmethod
        def backward(ctx, gout):
            z, flux, coeff = ctx.saved_tensors
            want_dy = ctx.needs_input_grad[1]
            dz, _, dy = ctx.problem.grad(z, coeff, flux, want_dy=want_dy, scale=gout)
            gz = gout * dz if ctx.needs_input_grad[0] else None
            return gz, dy, None

    return Chi2AtRedshift


_FN = None


def chi2_at_redshift(z, flux, problem):
    """Differentiable chi2(z) (torch float64 tensors z (n,), flux (nb, nx))."""
    global _FN
    if _FN is None:
        _FN = _autograd_function()
    return _FN.apply(z, flux, problem)


def _subproblem(problem, item_target):
    """The problem's targets and template with another item list (no new uploads)."""
    import torch
    sub = GradProblem.__new__(GradProblem)
    sub.__dict__.update(problem.__dict__)
    sub.item_target = np.asarray(item_target, dtype=np.int32)
    sub.t_item_target = torch.as_tensor(sub.item_target, device='cuda')
    return sub


def newton_refine(problem, z0, zlo, zhi, iters=12, dz_fd=1e-6, rtol=1e-15, zref=None, chi2ref=None,
                  info=None):
    """Safeguarded Newton iterations on dchi2/dz = 0 for every item, inside [zlo, zhi].

    One evaluation at z runs the FP64 forward at (z - h, z, z + h) -- chi2(z)
    is the middle point -- and the adjoint gradient at the three; the
    curvature is the central difference of the analytic gradient.  From the
    current point the step is -g/H when H > 0, else a quarter of the bracket
    downhill, clipped to the bracket.  A trial point whose chi2 exceeds the
    current one is rejected and the step halved (backtracking); an accepted
    trial becomes the current point, so chi2 never rises above chi2(z0).  An
    item stops when the predicted decrease |g step| / 2 falls below
    rtol * chi2, when its step falls below 1e-12 (1 + z), or after `iters`
    evaluations.

    zref, chi2ref: an optional reference point per item (the parabola fit,
    fitz.py:259-261); an item whose final chi2 exceeds chi2ref returns zref
    (its chi2 and gradient evaluated there).  info: a dict that receives
    'converged' (bool per item), 'evaluations' and 'kept_reference' (bool
    per item).  Returns (z, chi2, dchi2/dz) as numpy arrays.
    """
    z = np.clip(np.array(z0, dtype=np.float64), zlo, zhi)
    zlo = np.asarray(zlo, dtype=np.float64)
    zhi = np.asarray(zhi, dtype=np.float64)
    n = len(z)

    def eval3(idx, zc):
        # (z - h, z, z + h) of an item are consecutive: one multi-redshift scan item
        sub = _subproblem(problem, np.repeat(problem.item_target[idx], 3))
        zz = np.stack([zc - dz_fd, zc, zc + dz_fd], axis=1).ravel()
        chi2, coeff = sub.forward(zz)
        g, _, _ = sub.grad(zz, coeff)
        g = g.cpu().numpy().reshape(-1, 3)
        c = chi2.cpu().numpy().reshape(-1, 3)
        return c[:, 1], g[:, 1], (g[:, 2] - g[:, 0]) / (2 * dz_fd)

    def newton_step(idx):
        gi, hi_ = g[idx], H[idx]
        s = np.where(hi_ > 0, -gi / np.where(hi_ > 0, hi_, 1.0), -np.sign(gi) * 0.25 * (zhi[idx] - zlo[idx]))
        return np.clip(z[idx] + s, zlo[idx], zhi[idx]) - z[idx]

    def done():
        return ((np.abs(g * step) <= 2 * rtol * np.abs(f)) | (np.abs(step) <= 1e-12 * (1 + np.abs(z)))
                | ~np.isfinite(step))

    allidx = np.arange(n)
    f, g, H = eval3(allidx, z)
    step = newton_step(allidx)
    evals = 1
    conv = done()
    while evals < iters and not conv.all():
        idx = np.flatnonzero(~conv)
        zt = np.clip(z[idx] + step[idx], zlo[idx], zhi[idx])
        ft, gt, ht = eval3(idx, zt)
        evals += 1
        acc = ft <= f[idx]
        ia, ir = idx[acc], idx[~acc]
        z[ia], f[ia], g[ia], H[ia] = zt[acc], ft[acc], gt[acc], ht[acc]
        step[ia] = newton_step(ia)
        step[ir] *= 0.5
        conv = done()
    kept = np.zeros(n, dtype=bool)
    if zref is not None and chi2ref is not None:
        kept = np.asarray(chi2ref, dtype=np.float64) < f
        if kept.any():
            ik = np.flatnonzero(kept)
            z[ik] = np.asarray(zref, dtype=np.float64)[ik]
            f[ik], g[ik], _ = eval3(ik, z[ik])
    if info is not None:
        info.update(converged=conv.copy(), evaluations=evals, kept_reference=kept)
    return z, f, g
