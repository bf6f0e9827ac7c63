"""Archetypes: redrock's archetype mode, with the final fit on the GPU.

Same classes and functions as the reference's archetypes module
(py/redrock/archetypes.py): Archetype (a redrock-format archetype file),
All_archetypes, find_archetypes, split_archetype_coeff.  The fit itself --
every archetype at every (target, redshift) pair, bounded least squares
(archetype coefficients >= 0) with Legendre polynomials per camera and an
optional prior, then the best archetype or the combination of the n nearest
-- runs as rr_arch_prep / rr_arch_fit (csrc/archfit.hip) over all pairs at
once (ArchetypeFit); get_best_archetype keeps the reference's one-target
signature on top of it.
"""
from __future__ import annotations

import os
from glob import glob

import numpy as np

from . import _hip, constants
from .templates import make_fulltype
from .rebin import centers2edges, rebin_into, trapz_rebin
from .utils import reduced_wavelength, legendre_rows

HUGE_CHI2 = 9999999.0                     # zscan.py:18


def _lines_of(model):
    from .igm import lyman_table
    return lyman_table(model) or []


def _igm_array(z, wave, model):
    """Transmission T(z, wave) of the Lyman series, 1 where no line applies.

    The Inoue14 model is outside the scan path: it is accepted only where no
    Lyman line reaches the pixels (lambda/(1+z) above Lyman alpha), where
    every model gives T = 1.
    """
    wave = np.asarray(wave, dtype=np.float64)
    if model is None or model == 'None':
        return np.ones_like(wave)
    if np.min(wave) / (1. + z) > constants.LyA_wavelength:
        return np.ones_like(wave)
    from .igm import transmission_Lyman
    return np.asarray(transmission_Lyman(np.array([z]), wave, model=model))[0]


class Archetype(object):
    """All archetypes of one spectral type, from a redrock-format archetype
    file (archetypes.py:26-75): HDU ARCHETYPES with columns ARCHETYPE
    (narch, nwave) and SUBTYPE, header RRTYPE, VERSION, CRVAL1, CDELT1,
    LOGLAM."""

    def __init__(self, filename):
        from . import fits
        self._filename = filename
        h = fits.read(os.path.expandvars(filename))
        hdr = h['ARCHETYPES'].header
        data = h['ARCHETYPES'].data
        self.flux = np.asarray(data['ARCHETYPE']).astype('float64')
        self._narch = self.flux.shape[0]
        self._nwave = self.flux.shape[1]
        self._rrtype = hdr['RRTYPE'].strip()
        sub = np.array(np.char.strip(np.asarray(data['SUBTYPE']).astype(str)))
        self._subtype = np.char.add(np.char.add(sub, '_'), np.arange(self._narch, dtype=int).astype(str))
        self._full_type = np.array([make_fulltype(self._rrtype, s) for s in self._subtype])
        self._version = hdr['VERSION']
        self.wave = np.asarray(hdr['CRVAL1'] + hdr['CDELT1'] * np.arange(self._nwave))
        if hdr['LOGLAM']:
            self.wave = 10 ** self.wave
        self.minwave = self.wave[0]
        self.maxwave = self.wave[-1]
        self.igm_model = 'Inoue14'     # archetypes.py:72
        self.method = 'ARCH'
        self._dev = None

    @classmethod
    def from_arrays(cls, rrtype, wave, flux, subtypes=None, version='test', igm_model='Inoue14'):
        """An archetype set held in memory (same attributes as from a file)."""
        self = cls.__new__(cls)
        self._filename = None
        self.flux = np.asarray(flux, dtype=np.float64)
        self._narch, self._nwave = self.flux.shape
        self._rrtype = rrtype
        sub = np.array(subtypes if subtypes is not None else ['ARCH'] * self._narch, dtype=str)
        self._subtype = np.char.add(np.char.add(sub, '_'), np.arange(self._narch, dtype=int).astype(str))
        self._full_type = np.array([make_fulltype(self._rrtype, s) for s in self._subtype])
        self._version = version
        self.wave = np.asarray(wave, dtype=np.float64)
        self.minwave = self.wave[0]
        self.maxwave = self.wave[-1]
        self.igm_model = igm_model
        self.method = 'ARCH'
        self._dev = None
        return self

    @property
    def version(self):
        return self._version

    @property
    def filename(self):
        return self._filename

    @property
    def template_type(self):
        return self._rrtype

    @property
    def sub_type(self):
        return self._subtype

    @property
    def full_type(self):
        return self._full_type

    @property
    def narch(self):
        return self._narch

    def device(self):
        """(wave, flux) as device tensors, copied once."""
        import torch
        if self._dev is None:
            self._dev = (torch.as_tensor(self.wave, device='cuda').contiguous(),
                         torch.as_tensor(self.flux, device='cuda').contiguous())
        return self._dev

    def rebin_template(self, index, z, dwave, trapz=True):
        """One archetype at redshift z on every grid (archetypes.py:125-131)."""
        if trapz:
            return {hs: trapz_rebin((1. + z) * self.wave, self.flux[index], w) for hs, w in dwave.items()}
        return {hs: _interp_extrapolate(w / (1. + z), self.wave, self.flux[index]) for hs, w in dwave.items()}

    def rebin_template_batch(self, z, dwave, trapz=True, dedges=None, indx=None, use_gpu=False):
        """Every archetype (or flux[indx]) at redshift z: {wavehash: (nwave, narch)}
        (archetypes.py:133-191)."""
        flux = self.flux if indx is None else self.flux[indx]
        out = dict()
        if trapz:
            grids = dwave if dwave is not None else {hs: e for hs, e in dedges.items()}
            for hs, g in grids.items():
                if dwave is not None:
                    r = trapz_rebin((1. + z) * self.wave, flux, g, use_gpu=use_gpu)
                else:
                    e = g[0] if isinstance(g, tuple) else g
                    r = trapz_rebin(self.wave, flux, edges=e, myz=np.array([z]), use_gpu=use_gpu)[0]
                out[hs] = r
            return out
        for hs, w in dwave.items():
            out[hs] = np.stack([_interp_extrapolate(w / (1. + z), self.wave, f) for f in flux], axis=1)
        return out

    def eval(self, subtype, coeff, wave, z, R=None, legcoeff=None):
        """Model flux of the given archetype(s) and coefficients (archetypes.py:193-238)."""
        subtypes = subtype.split(';')
        model = np.zeros(len(wave))
        T = _igm_array(z, wave, self.igm_model)
        for this_subtype, c in zip(subtypes, coeff[0:len(subtypes)]):
            index = np.where(self._subtype == this_subtype)[0][0]
            binned = trapz_rebin((1 + z) * self.wave, self.flux[index], wave)
            model += c * binned * T
        if legcoeff is not None:
            legcoeff = np.asarray(legcoeff)
            deg = len(legcoeff) if legcoeff.ndim == 1 else legcoeff.shape[1]
            leg = legendre_rows(wave, deg)
            model += leg.T.dot(legcoeff).T
        if R is not None:
            model = R.dot(model)
        return model

    def get_best_archetype(self, target, weights, flux, wflux, dwave, z, per_camera, n_nearest,
                           trans=None, solve_method='bvls', prior=None, use_gpu=False):
        """(chi2, coeff, fulltype) of the best archetype, or of the n_nearest
        combination, for one target at redshift z (archetypes.py:316-400).

        trans: {wavehash: transmission} (or None entries) as fitz passes it;
        None computes the transmission of this archetype set's IGM model.
        """
        if solve_method.upper() != 'BVLS':
            raise NotImplementedError(f'archetype solve method {solve_method}')
        nleg = target.nleg
        ncam = len(dwave) if per_camera else 1
        fit = ArchetypeFit(self, [target], per_camera=per_camera, nleg=nleg, prior=prior,
                           n_nearest=n_nearest, ncamera=ncam)
        tpix = None
        if trans is None:
            tpix = [np.concatenate([_igm_array(z, s.wave, self.igm_model) for s in target.spectra])]
        else:
            parts = []
            for s in target.spectra:
                T = trans.get(s.wavehash)
                parts.append(np.ones(s.nwave) if T is None else np.asarray(T, dtype=np.float64).reshape(-1))
            tpix = [np.concatenate(parts)]
        chi2, coeff, ftype = fit.run(np.array([0]), np.array([z], dtype=np.float64), tpix=tpix)
        return chi2[0], coeff[0], ftype[0]


def _interp_extrapolate(xnew, x, y):
    """Linear interpolation with linear extrapolation at both ends
    (scipy interp1d(kind='linear', fill_value='extrapolate'))."""
    out = np.interp(xnew, x, y)
    lo, hi = xnew < x[0], xnew > x[-1]
    out[lo] = y[0] + (xnew[lo] - x[0]) * (y[1] - y[0]) / (x[1] - x[0])
    out[hi] = y[-1] + (xnew[hi] - x[-1]) * (y[-1] - y[-2]) / (x[-1] - x[-2])
    return out


def prior_on_coeffs(n_nbh, deg_legendre, sigma, ncamera):
    """(nbasis, nbasis) prior: 1/sigma^2 on the Legendre coefficients, 0 on the
    archetypes (fitz.py:113-130)."""
    nbasis = n_nbh + deg_legendre * ncamera
    prior = np.zeros((nbasis, nbasis), dtype='float64')
    np.fill_diagonal(prior, 1 / (sigma ** 2))
    for i in range(n_nbh):
        prior[i][i] = 0.
    return prior


def split_archetype_coeff(subtype, coeff, nbands, nleg):
    """Archetype coefficients and per-band Legendre coefficients
    (archetypes.py:428-443)."""
    narchetypes = len(subtype.split(';'))
    archcoeff = coeff[0:narchetypes]
    all_legcoeff = coeff[narchetypes:]
    legcoeff = [all_legcoeff[i * nleg:(i + 1) * nleg] for i in range(nbands)]
    return archcoeff, legcoeff


def find_archetypes(archetypes_dir=None):
    """rrarchetype-*.fits files of archetypes_dir, else $RR_ARCHETYPE_DIR
    (archetypes.py:445-467)."""
    if archetypes_dir is None:
        archetypes_dir = os.environ.get('RR_ARCHETYPE_DIR')
        if archetypes_dir is None:
            raise IOError('no archetype directory: pass one or set $RR_ARCHETYPE_DIR')
    if os.path.isfile(archetypes_dir):
        return [archetypes_dir]
    return sorted(glob(os.path.join(archetypes_dir, 'rrarchetype-*.fits')))


class All_archetypes(object):
    """Archetype objects of every spectral type, keyed by type (archetypes.py:403-426)."""

    def __init__(self, lstfilename=None, archetypes_dir=None, verbose=False):
        if lstfilename is None:
            lstfilename = find_archetypes(archetypes_dir)
        self.archetypes = dict()
        for f in lstfilename:
            a = Archetype(f)
            if verbose:
                print(f'DEBUG: Found {a._narch} archetypes for SPECTYPE {a._rrtype} in file {f}')
            self.archetypes[a._rrtype] = a


class ArchetypeFit(object):
    """The archetype fit of many (target, redshift) pairs on the GPU.

    targets share one pixel layout (their wavelength grids).  Per target the
    Legendre terms are prepared once (rr_arch_prep); run() rebins every
    archetype at the pairs' redshifts into a cache (npair, rows, narch) in
    HBM, in chunks that fit the device, and calls rr_arch_fit: first every
    archetype alone (archetype coefficient >= 0), then, with n_nearest, the
    n_nearest best archetypes of each pair together.

    per_camera: one set of nleg Legendre polynomials per camera (wavelength
    grid), ordered by band name in the result (zscan.py:336-347); else one
    set over the whole spectrum.  prior: (n + nleg*ncamera)^2 matrix added to
    M (prior_on_coeffs), n = n_nearest or 1; ignored without per_camera, as
    in the reference (archetypes.py:377-385).
    """

    HALO = 8

    def __init__(self, archetype, targets, per_camera=True, nleg=2, prior=None, n_nearest=None,
                 ncamera=None, batch=None):
        import torch
        from .zscan import TargetBatch, PixelLayout
        self.arch = archetype
        self.targets = list(targets)
        self.per_camera = bool(per_camera)
        self.nleg = int(nleg)
        self.n_nearest = n_nearest
        if n_nearest is not None and not (1 <= n_nearest <= _hip.RR_ARCH_MAX_NA):
            raise NotImplementedError(f'n_nearest={n_nearest}: 1..{_hip.RR_ARCH_MAX_NA} supported')
        sp0 = self.targets[0].spectra
        self.layout = PixelLayout.of(sp0)
        self.dwave = {s.wavehash: s.wave for s in sp0}
        self.hashes = list(self.dwave.keys())
        self.ncam = len(self.hashes) if self.per_camera else 1
        if ncamera is not None and ncamera != self.ncam:
            raise ValueError(f'ncamera={ncamera} but the targets have {self.ncam} wavelength grids')
        self.NL = self.nleg * self.ncam
        if self.NL > _hip.RR_ARCH_MAX_NLEG:
            raise NotImplementedError(f'{self.NL} Legendre columns: at most {_hip.RR_ARCH_MAX_NLEG}')
        if batch is not None and batch.nt == len(self.targets) and batch.layout.signature == self.layout.signature:
            self.batch = batch
        else:
            self.batch = TargetBatch(self.targets, layout=self.layout)
        P = self.layout.npix_pad
        self.P = P
        # Legendre columns on the pixel axis (targets.py:237-254 per grid)
        leg = np.zeros((max(self.NL, 1), P))
        rows = {hs: legendre_rows(w, self.nleg) for hs, w in self.dwave.items()} if self.nleg else {}
        for s, off in zip(sp0, self.layout.offsets):
            if self.nleg == 0:
                break
            i = self.hashes.index(s.wavehash) if self.per_camera else 0
            leg[i * self.nleg:(i + 1) * self.nleg, off:off + s.nwave] = rows[s.wavehash]
        dev = torch.device('cuda')
        self.leg = torch.as_tensor(leg, device=dev).contiguous()
        nt = len(self.targets)
        nl1 = max(self.NL, 1)
        self.wrl = torch.empty((nt, nl1, P), dtype=torch.float64, device=dev)
        self.gll = torch.empty((nt, nl1, nl1), dtype=torch.float64, device=dev)
        self.yl = torch.empty((nt, nl1), dtype=torch.float64, device=dev)
        self.fwf = torch.empty(nt, dtype=torch.float64, device=dev)
        rc = _hip.lib().rr_arch_prep(_hip.ptr(self.batch.rband), _hip.ptr(self.batch.w), _hip.ptr(self.batch.f),
                                     nt, P, _hip.ptr(self.leg), self.NL, _hip.ptr(self.wrl), _hip.ptr(self.gll),
                                     _hip.ptr(self.yl), _hip.ptr(self.fwf), _hip.stream_handle())
        _hip.check(rc, 'rr_arch_prep')
        self.wsum = np.array([sum(float(np.sum(s.ivar)) for s in t.spectra) for t in self.targets])
        self.wave_pix = torch.as_tensor(np.concatenate(
            [np.concatenate([s.wave for s in sp0]), np.zeros(P - self.layout.npix)]), device=dev).contiguous()
        self.edges = [torch.as_tensor(centers2edges(s.wave), device=dev).contiguous() for s in sp0]
        # fused rebin inputs: archetype fluxes wavelength-major, each pixel's bin edges
        pe = [centers2edges(s.wave) for s in sp0]
        pad = np.zeros(P - self.layout.npix)
        self.pe_lo = torch.as_tensor(np.concatenate([e[:-1] for e in pe] + [pad]), device=dev).contiguous()
        self.pe_hi = torch.as_tensor(np.concatenate([e[1:] for e in pe] + [pad]), device=dev).contiguous()
        self.yt = None
        self.edges_host = [(float(centers2edges(s.wave)[0]), float(centers2edges(s.wave)[-1])) for s in sp0]
        # Legendre blocks in band order (zscan.py:336-347) when the bands name the grids
        self.perm = np.arange(1 + self.NL)
        # without per_camera the reference solves with no prior (archetypes.py:377-385)
        self.prior_full = None if (prior is None or not self.per_camera) else np.asarray(prior, dtype=np.float64)
        nn = 1 if n_nearest is None else n_nearest
        if self.prior_full is not None and self.prior_full.shape != (nn + self.NL, nn + self.NL):
            raise ValueError(f'prior of shape {self.prior_full.shape}, expected {(nn + self.NL,) * 2}')

    def _band_order(self, t, na):
        """Column order of the result: archetypes, then Legendre blocks by band."""
        order = list(range(na + self.NL))
        if not self.per_camera or self.nleg == 0:
            return order
        bands = self.targets[t].bands
        if len(bands) != self.ncam or any(b is None for b in bands) or len(set(bands)) != len(bands):
            return order
        blocks = [list(range(na + i * self.nleg, na + (i + 1) * self.nleg)) for i in range(self.ncam)]
        byband = {b: blocks[i] for i, b in enumerate(bands)}
        return list(range(na)) + [c for b in sorted(bands) for c in byband[b]]

    def _orders(self, tidx, na):
        """(len(tidx), na + NL) result column order of every pair's target."""
        cache = dict()
        rows = []
        for t in tidx:
            t = int(t)
            if t not in cache:
                cache[t] = np.array(self._band_order(t, na))
            rows.append(cache[t])
        return np.stack(rows) if rows else np.zeros((0, na + self.NL), dtype=int)

    def _chunk(self, fused):
        import torch
        free, _ = torch.cuda.mem_get_info()
        # per pair: the archetype cache, or the fused path's per-pixel sample
        # geometry and weights (8 + 6 x 8 bytes a row) and normal equations
        per = ((self.P + 8) * 56 + self.arch.narch * 24 * 8 if fused
               else (self.P + 2 * self.HALO) * self.arch.narch * 8)
        budget = int(min(max(free // 8, 256 << 20), 8 << 30))
        return int(min(max(1, budget // per), 4 * 65535))      # 4 pairs per grid row

    def _fused_inputs(self):
        import torch
        if self.yt is None:
            x, y = self.arch.device()
            self.yt = y.t().contiguous()
        return self.arch.device()[0], self.yt

    def _fit(self, cache, npair, tgt, zigm, lines, na, nprob, columns, prior, nbound, zrebin=None):
        import torch
        dev = torch.device('cuda')
        nb = na + self.NL
        coeff = torch.empty((npair, nprob, nb), dtype=torch.float64, device=dev)
        chi2 = torch.empty((npair, nprob), dtype=torch.float64, device=dev)
        d = _hip.ArchDesc()
        d.npair, d.nprob, d.na, d.nleg, d.nbound = npair, nprob, na, self.NL, nbound
        d.npix_pad, d.halo = self.P, self.HALO
        d.d_pair_target = _hip.ptr(tgt)
        d.d_acache = _hip.ptr(cache)
        if cache is None:
            x, yt = self._fused_inputs()
            d.d_x, d.nx, d.d_yt, d.ldy = _hip.ptr(x), int(x.numel()), _hip.ptr(yt), self.arch.narch
            d.d_pe_lo, d.d_pe_hi, d.npix = _hip.ptr(self.pe_lo), _hip.ptr(self.pe_hi), self.layout.npix
            d.d_zrebin = _hip.ptr(zrebin)
        d.ld = self.arch.narch
        d.narch = self.arch.narch
        d.d_columns = _hip.ptr(columns)
        d.d_rband, d.d_w, d.d_wf = _hip.ptr(self.batch.rband), _hip.ptr(self.batch.w), _hip.ptr(self.batch.wf)
        if self.NL:
            d.d_wrl, d.d_gll, d.d_yl = _hip.ptr(self.wrl), _hip.ptr(self.gll), _hip.ptr(self.yl)
        d.d_fwf = _hip.ptr(self.fwf)
        pr = None
        if prior is not None:
            pr = torch.as_tensor(np.ascontiguousarray(prior), device=dev)
            d.d_prior = _hip.ptr(pr)
        larr = None
        if lines:
            larr = (_hip.LymanLine * len(lines))(*[_hip.LymanLine(*l) for l in lines])
            d.lines = larr
            d.nlines = len(lines)
            d.d_wave = _hip.ptr(self.wave_pix)
            d.d_zigm = _hip.ptr(zigm)
        d.d_coeff, d.d_chi2 = _hip.ptr(coeff), _hip.ptr(chi2)
        nbytes = _hip.lib().rr_arch_workspace_bytes(d)
        work = torch.empty(max(int(nbytes), 8) // 8 + 1, dtype=torch.float64, device=dev)
        d.d_work, d.work_bytes = _hip.ptr(work), int(work.numel() * 8)
        rc = _hip.lib().rr_arch_fit(d, _hip.stream_handle())
        _hip.check(rc, 'rr_arch_fit')
        return chi2, coeff

    def run_all(self, tidx, z, zigm=None, lines=None, fused=True):
        """Every archetype alone at every pair: chi2 (npair, narch) and coeff
        (npair, narch, 1 + nleg columns in band order)."""
        return self.run(tidx, z, zigm=zigm, lines=lines, keep_all=True, fused=fused)

    def run(self, tidx, z, zigm=None, lines=None, tpix=None, keep_all=False, fused=True):
        """Fit every pair (targets[tidx[k]], z[k]).

        zigm[k]: redshift of the Lyman-series factor of pair k (NaN: none);
        lines: (line, A, B) list of that factor.  tpix: per pair a pixel
        transmission array multiplied into the archetypes instead (the
        one-target API; it uses the cache path).  fused=False rebins into an
        HBM cache first (rr_rebin) instead of inside the fit kernel, where each
        bin mean is the same trapezoid written as weights of its samples
        (equal to a few ulp).
        Returns (chi2 (npair,), [coeff], [fulltype]) with
        coeff = archetype coefficient(s) then the Legendre coefficients.
        Raises ValueError when a redshifted archetype does not cover the
        pixels (rebin.py:332-341).
        """
        import torch
        tidx = np.asarray(tidx, dtype=np.int64)
        z = np.asarray(z, dtype=np.float64)
        npair = len(tidx)
        lines = list(lines or [])
        if zigm is None:
            zigm = np.full(npair, np.nan)
        for zk in z:
            lo, hi = self.arch.wave[0] * (1 + zk), self.arch.wave[-1] * (1 + zk)
            for elo, ehi in self.edges_host:
                if elo < lo or ehi > hi:
                    raise ValueError('edges must be within input x range')
        dev = torch.device('cuda')
        x, y = self.arch.device()
        narch = self.arch.narch
        rows = self.P + 2 * self.HALO
        fused = bool(fused) and tpix is None          # the cache path applies tpix
        chunk = min(npair, self._chunk(fused))
        cache = None if fused else torch.zeros((max(chunk, 1), rows, narch), dtype=torch.float64, device=dev)
        nn = self.n_nearest
        na1 = 1
        prior1 = None
        if self.per_camera and self.prior_full is not None:
            # the single-archetype pass drops the rows of the other neighbours (archetypes.py:362-366)
            prior1 = self.prior_full[(nn or 1) - 1:, (nn or 1) - 1:]
        out_chi2 = np.zeros(npair)
        out_coeff = [None] * npair
        out_type = [None] * npair
        all_chi2 = np.zeros((npair, narch)) if keep_all else None
        all_coeff = np.zeros((npair, narch, 1 + self.NL)) if keep_all else None
        for c0 in range(0, npair, max(chunk, 1)):
            sl = slice(c0, min(npair, c0 + chunk))
            n = sl.stop - sl.start
            zt = torch.as_tensor(z[sl], device=dev).contiguous()
            if not fused:
                cache.zero_()
                flat = cache.view(-1)
                for e, off in zip(self.edges, self.layout.offsets):
                    base = (self.HALO + off) * narch
                    rebin_into(flat[base:], x, y, e, zt, rows * narch, narch, 1)
            if tpix is not None:
                T = torch.as_tensor(np.stack([np.concatenate([tp, np.ones(self.P - len(tp))])
                                              for tp in tpix[sl]]), device=dev)
                cache[:n, self.HALO:self.HALO + self.P, :] *= T[:, :, None]
            tgt = torch.as_tensor(tidx[sl].astype(np.int32), device=dev).contiguous()
            zi = torch.as_tensor(np.asarray(zigm[sl], dtype=np.float64), device=dev).contiguous()
            chi2, coeff = self._fit(cache, n, tgt, zi, lines, na1, narch, None, prior1, 1, zrebin=zt)
            zero = self.wsum[tidx[sl]] == 0
            chi2h = chi2.cpu().numpy()
            chi2h[zero] = HUGE_CHI2                      # zscan.py:656-660
            if keep_all:
                coeff = coeff.cpu().numpy()
                coeff[zero] = 0.0
                for k in range(n):
                    all_chi2[c0 + k] = chi2h[k]
                    all_coeff[c0 + k] = coeff[k][:, self._band_order(int(tidx[sl][k]), 1)]
                continue
            if nn is None:
                # only the best archetype's coefficients come back to the host
                best = np.argmin(chi2h, axis=1)
                cb = coeff[torch.arange(n, device=dev), torch.as_tensor(best, device=dev)].cpu().numpy()
                cb[zero] = 0.0
                cb = np.take_along_axis(cb, self._orders(tidx[sl], 1), axis=1)
                out_chi2[sl] = chi2h[np.arange(n), best]
                out_coeff[c0:c0 + n] = list(cb)
                out_type[c0:c0 + n] = list(self.arch.full_type[best])
                continue
            ibest = np.argsort(chi2h, axis=1)[:, :nn]
            cols = torch.as_tensor(ibest.astype(np.int32).reshape(n, 1, nn), device=dev).contiguous()
            prior2 = self.prior_full if self.per_camera else None
            chi2b, coeffb = self._fit(cache, n, tgt, zi, lines, nn, 1, cols, prior2, nn, zrebin=zt)
            chi2b = chi2b.cpu().numpy()[:, 0]
            coeffb = coeffb.cpu().numpy()[:, 0]
            chi2b[zero] = HUGE_CHI2
            coeffb[zero] = 0.0
            for k in range(n):
                t = int(tidx[sl][k])
                order = self._band_order(t, nn)
                out_chi2[c0 + k] = chi2b[k]
                out_coeff[c0 + k] = coeffb[k][order]
                out_type[c0 + k] = make_fulltype(self.arch._rrtype,
                                                 ';'.join(str(self.arch._subtype[i]) for i in ibest[k]))
        if keep_all:
            return all_chi2, all_coeff
        return out_chi2, out_coeff, out_type
