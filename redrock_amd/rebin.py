"""Trapezoidal rebinning on the GPU (py/redrock/rebin.py:100-517).

``trapz_rebin`` and ``rebin_template`` keep the reference signatures and
shape rules.  The arithmetic runs in the rr_rebin HIP kernel (FP64, the same
operations in the same order as the reference's sweep, rebin.py:126-176).
With use_gpu=True results are torch tensors on the GPU, otherwise numpy
arrays copied back from the GPU.  The optional ``lyman`` argument (a list of
(line, A, B)) fuses the IGM transmission multiply into the kernel.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _hip


def centers2edges(centers):
    """Bin edges from bin centres (rebin.py:100-118)."""
    centers = np.asarray(centers)
    edges = np.zeros(len(centers) + 1)
    edges[1:-1] = (centers[0:-1] + centers[1:]) / 2.0
    edges[0] = centers[0] - (centers[1] - edges[1])
    edges[-1] = centers[-1] + (centers[-1] - edges[-2])
    return edges


def _as_host(a):
    if hasattr(a, 'detach'):
        return a.detach().cpu().numpy()
    return np.asarray(a)


def lyman_array(lines):
    if not lines:
        return None, 0
    if len(lines) > _hip.RR_MAX_LYMAN_LINES:
        raise ValueError('too many Lyman lines')
    arr = (_hip.LymanLine * len(lines))(*[_hip.LymanLine(*l) for l in lines])
    return arr, len(lines)


def rebin_into(out, x, y, edges, z, s_z, s_pix, s_b, out_type=_hip.RR_F64, scale=1.0,
               lyman=None, wave=None, stream=None):
    """Launch rr_rebin writing element (iz, ibin, ib) at out + iz*s_z + ibin*s_pix + ib*s_b.

    x (nx), y (nb, nx), edges, z, wave are contiguous float64 device tensors;
    out is a contiguous device tensor (or a flat view) of the element type
    `out_type`.  No range check here.
    """
    lines, nlines = lyman_array(lyman)
    nb, nx = y.shape
    if x.numel() != nx:
        raise ValueError('x and y have different lengths')
    rc = _hip.lib().rr_rebin(_hip.ptr(x), nx, _hip.ptr(y), nb, nx,
                             _hip.ptr(edges), edges.numel() - 1,
                             _hip.ptr(wave) if nlines else None,
                             _hip.ptr(z), z.numel(), lines, nlines, _hip.ptr(out), out_type,
                             s_z, s_pix, s_b, float(scale), _hip.stream_handle(stream))
    _hip.check(rc, 'rr_rebin')


def trapz_rebin(x, y, xnew=None, edges=None, myz=None, use_gpu=False, xmin=None, xmax=None,
                edge_min=None, edge_max=None):
    """Rebin y(x) flux density between bin edges (rebin.py:212-378).

    Shapes: (nz, nbin, nbasis) with the nz axis dropped for scalar/None myz
    and the basis axis dropped for 1-d y.  Raises ValueError when the edges
    leave the redshifted x range (rebin.py:332-341).
    """
    import torch
    if edges is None:
        edges = centers2edges(_as_host(xnew))
    edges = np.asarray(_as_host(edges), dtype=np.float64)
    nbins = len(edges) - 1
    if edge_min is None:
        edge_min = edges[0]
    if edge_max is None:
        edge_max = edges[-1]
    scalar_z = False
    if myz is None:
        myz = np.array([0], dtype=np.float64)
        scalar_z = True
    elif np.isscalar(myz):
        myz = np.array([myz], dtype=np.float64)
        scalar_z = True
    else:
        myz = np.asarray(_as_host(myz), dtype=np.float64)
    nz = len(myz)
    if nz == 0:
        out = np.zeros((0, nbins, 1), dtype=np.float64)
        return torch.as_tensor(out, device='cuda') if use_gpu else out
    xh = np.asarray(_as_host(x), dtype=np.float64)
    if xmin is None:
        xmin = xh[0] * (1 + myz.max())
    if xmax is None:
        xmax = xh[-1] * (1 + myz.min())
    if edge_min < xmin or edge_max > xmax:
        raise ValueError('edges must be within input x range')
    yh = _as_host(y)
    scalar_basis = (np.ndim(yh) == 1)
    y2 = np.ascontiguousarray(np.atleast_2d(yh), dtype=np.float64)
    nbasis = y2.shape[0]
    result_shape = (nbins,)
    if not scalar_basis:
        result_shape += (nbasis,)
    if not scalar_z:
        result_shape = (nz,) + result_shape
    dev = torch.device('cuda')
    xt = torch.as_tensor(np.ascontiguousarray(xh), device=dev).contiguous()
    yt = torch.as_tensor(y2, device=dev).contiguous()
    et = torch.as_tensor(np.ascontiguousarray(edges), device=dev).contiguous()
    zt = torch.as_tensor(np.ascontiguousarray(myz), device=dev).contiguous()
    out = torch.empty((nz, nbins, nbasis), dtype=torch.float64, device=dev)
    rebin_into(out, xt, yt, et, zt, nbins * nbasis, nbasis, 1)
    out = out.reshape(result_shape)
    if use_gpu:
        return out
    return out.cpu().numpy()


def rebin_template(template, myz, dwave=None, dedges=None, use_gpu=False, xmin=None, xmax=None):
    """Rebin a template to every wavelength grid (rebin.py:453-517)."""
    myz = np.asarray(_as_host(myz), dtype=np.float64)
    if xmin is None:
        xmin = template.minwave * (1 + myz.max())
    if xmax is None:
        xmax = template.maxwave * (1 + myz.min())
    result = dict()
    if dwave is not None:
        for hs, wave in dwave.items():
            result[hs] = trapz_rebin(template.wave, template.flux, xnew=wave, myz=myz,
                                     use_gpu=use_gpu, xmin=xmin, xmax=xmax)
    elif dedges is not None:
        for hs, edges in dedges.items():
            minedge = maxedge = None
            if type(edges) is tuple:
                edges, minedge, maxedge = edges
            result[hs] = trapz_rebin(template.wave, template.flux, edges=edges, myz=myz,
                                     use_gpu=use_gpu, xmin=xmin, xmax=xmax,
                                     edge_min=minedge, edge_max=maxedge)
    return result
