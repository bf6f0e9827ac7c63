"""Synthetic spectra and templates for the scan workloads (seeded, reproducible).

The workloads are the ones the chi2 scan is measured and tested on:

* ``galaxy`` — 3 cameras with linear grids b 3600-5800 A (2751 px),
  r 5760-7620 A (2326 px), z 7520-9824 A (2881 px); a 10-component GALAXY
  basis on a rest grid 1200-11000 A at 0.5 A (power-law continua plus
  Gaussian emission lines of sigma 2 A at 3727, 4861, 4959, 5007, 6563 A,
  orthonormalised after a seeded random mixing); redshifts
  exp(arange(ln 1.005, ln 2.7, 5e-4)) - 1 (1977 z); an 11-diagonal Gaussian
  resolution matrix per camera (sigma 1 px, rows normalised); flux =
  R (template at z_true ~ U(0.01, 1.6), lognormal coefficients) + noise;
  ivar lognormal(0, 0.5) with 2 % of pixels set to 0.
* ``qso`` — 4-component QSO basis (power law with index ~ N(-1.5, 0.3) plus
  broad sigma 20 A lines at 1216, 1549, 1909, 2799 A) on a rest grid that
  starts at 500 A (not 600 A: b-camera edges need 3599.6/7 = 514 A at z = 6,
  and the reference rebin raises ValueError otherwise, rebin.py:340-341),
  redshifts exp(arange(ln 1.05, ln 7, 5e-4)) - 1 (3795 z), Lyman-alpha-only
  IGM exp(-0.0036 (1+z_abs)^3.46) registered as model 'LyA-3.46'.
* ``star`` — 5-component STAR basis, 501 redshifts linspace(-0.005, 0.005).

The signal amplitude is stated explicitly: coefficients are lognormal(0, 1)
times a scale chosen so that the median per-pixel |model| sqrt(ivar) equals
``snr`` (default 1), i.e. a realistic per-pixel S/N for DESI galaxies.
The template at z_true is obtained by the exact antiderivative of the
piecewise-linear template (an independent evaluation of the same bin mean
the scan's rebin computes).
"""
from __future__ import annotations

import numpy as np
import scipy.sparse

from .targets import Spectrum, Target, RBAND_HALF
from .templates import Template
from . import igm

CAMERAS = (('b', 3600.0, 5800.0, 2751), ('r', 5760.0, 7620.0, 2326),
           ('z', 7520.0, 9824.0, 2881))

igm.register_lyman_model('LyA-3.46', [(1215.67, 0.0036, 3.46)])


def camera_grids():
    return {name: np.linspace(lo, hi, n) for name, lo, hi, n in CAMERAS}


def resolution_band(n, sigma=1.0):
    """(n, 11) band of a Gaussian LSF with rows normalised over in-range columns."""
    k = np.arange(-RBAND_HALF, RBAND_HALF + 1)
    kern = np.exp(-0.5 * (k / sigma) ** 2)
    band = np.tile(kern, (n, 1))
    p = np.arange(n)[:, None] + k[None, :]
    band[(p < 0) | (p >= n)] = 0.0
    band /= band.sum(axis=1, keepdims=True)
    return band


def band_to_dia(band):
    """Band (n, 11) -> scipy dia_matrix in the reference's column-indexed layout."""
    n, nd = band.shape
    data = np.zeros((nd, n))
    for kk in range(nd):
        off = kk - RBAND_HALF
        j = np.arange(n)
        p = j - off
        ok = (p >= 0) & (p < n)
        data[kk, j[ok]] = band[p[ok], kk]
    return scipy.sparse.dia_matrix((data, np.arange(nd) - RBAND_HALF), shape=(n, n))


def antiderivative_rebin(x, y, edges):
    """Exact bin means of the piecewise-linear y(x) over [edges[i], edges[i+1]].

    y may be (n,) or (nb, n).  Independent of the two-pointer sweep used by
    the scan; edges must lie inside [x[0], x[-1]].
    """
    x = np.asarray(x, dtype=np.float64)
    y2 = np.atleast_2d(np.asarray(y, dtype=np.float64))
    seg = 0.5 * (y2[:, 1:] + y2[:, :-1]) * np.diff(x)[None, :]
    F = np.concatenate([np.zeros((y2.shape[0], 1)), np.cumsum(seg, axis=1)], axis=1)
    e = np.asarray(edges, dtype=np.float64)
    j = np.clip(np.searchsorted(x, e, side='right') - 1, 0, len(x) - 2)
    t = e - x[j]
    slope = (y2[:, j + 1] - y2[:, j]) / (x[j + 1] - x[j])
    Fe = F[:, j] + y2[:, j] * t + 0.5 * slope * t * t
    out = np.diff(Fe, axis=1) / np.diff(e)[None, :]
    return out[0] if np.ndim(y) == 1 else out


def _orthonormalise(comps, rng, mix=True):
    c = np.array(comps, dtype=np.float64)
    c /= np.linalg.norm(c, axis=1, keepdims=True)
    if mix:
        c = rng.normal(size=(c.shape[0], c.shape[0])) @ c
    q, _ = np.linalg.qr(c.T)
    return q.T.copy()


def _gauss(w, mu, sig):
    return np.exp(-0.5 * ((w - mu) / sig) ** 2)


def _texture(wave, rng, corr=15.0):
    """Unit-rms smooth random field (Gaussian correlation length `corr` A).

    Stands in for the stellar absorption structure of real templates: it keeps
    every basis vector independent inside any observed window, so the normal
    matrix stays full rank at every trial redshift (lines alone leave the
    window at high z and make M exactly singular).
    """
    from scipy.ndimage import gaussian_filter1d
    step = wave[1] - wave[0]
    t = gaussian_filter1d(rng.normal(size=wave.size), corr / step, mode='nearest')
    return t / np.sqrt(np.mean(t * t))


def _with_texture(comps, wave, rng, amp):
    out = []
    for c in comps:
        c = np.asarray(c, dtype=np.float64)
        out.append(c + amp * np.sqrt(np.mean(c * c)) * _texture(wave, rng))
    return out


def galaxy_template(seed=1):
    rng = np.random.default_rng(seed)
    wave = np.linspace(1200.0, 11000.0, 19601)
    comps = [(wave / 5000.0) ** a for a in (-2.0, -1.0, 0.0, 1.0, 2.0)]
    comps += [_gauss(wave, mu, 2.0) for mu in (3727.0, 4861.0, 4959.0, 5007.0, 6563.0)]
    comps = _with_texture(comps, wave, rng, 0.3)
    flux = _orthonormalise(comps, rng)
    z = np.exp(np.arange(np.log(1.005), np.log(2.7), 5e-4)) - 1
    return Template(spectype='GALAXY', redshifts=z, wave=wave, flux=flux,
                    subtype='', igm_model='Calura12')


def qso_template(seed=2):
    rng = np.random.default_rng(seed)
    wave = np.linspace(500.0, 10000.0, 19001)
    alpha = rng.normal(-1.5, 0.3)
    comps = [(wave / 1450.0) ** alpha,
             _gauss(wave, 1216.0, 20.0) + 0.3 * _gauss(wave, 1549.0, 20.0),
             _gauss(wave, 1549.0, 20.0) + 0.5 * _gauss(wave, 1909.0, 20.0),
             _gauss(wave, 1909.0, 20.0) + _gauss(wave, 2799.0, 20.0)]
    comps = _with_texture(comps, wave, rng, 0.3)
    flux = _orthonormalise(comps, rng, mix=False)
    z = np.exp(np.arange(np.log(1.05), np.log(7.0), 5e-4)) - 1
    return Template(spectype='QSO', redshifts=z, wave=wave, flux=flux,
                    subtype='', igm_model='LyA-3.46')


def star_template(seed=3):
    rng = np.random.default_rng(seed)
    wave = np.linspace(3000.0, 11000.0, 16001)
    hc_k = 1.4388e8   # h c / k in A K
    comps = [wave ** -5 / np.expm1(hc_k / (wave * T)) for T in (3500.0, 5000.0, 7000.0, 10000.0)]
    comps = [c / c.max() for c in comps]
    lines = np.zeros_like(wave)
    for mu in (3934.8, 3969.6, 4102.9, 4341.7, 4862.7, 6564.6, 8500.4, 8544.4, 8664.5):
        lines -= _gauss(wave, mu, 3.0)
    comps.append(lines)
    comps = _with_texture(comps, wave, rng, 0.1)
    flux = _orthonormalise(comps, rng)
    z = np.linspace(-0.005, 0.005, 501)
    return Template(spectype='STAR', redshifts=z, wave=wave, flux=flux,
                    subtype='', igm_model='None')


def make_targets(template, ntarget, seed=0, snr=1.0, zrange=None, mask_frac=None,
                 nodata_frac=0.02, first_id=0):
    """Synthetic targets for `template`.

    Returns (targets, z_true).  mask_frac = (lo, hi) draws a masked fraction
    U(lo, hi) per target, masked in contiguous runs of 50-500 px.
    """
    rng = np.random.default_rng(seed)
    grids = camera_grids()
    bands = {k: resolution_band(len(w)) for k, w in grids.items()}
    dias = {k: band_to_dia(b) for k, b in bands.items()}
    if zrange is None:
        zr = template.redshifts
        if template.template_type == 'GALAXY':
            zrange = (0.01, 1.6)
        else:
            zrange = (zr[1], zr[-2])
    from .igm import transmission_Lyman
    targets, ztrue = [], []
    for it in range(ntarget):
        z = rng.uniform(*zrange)
        coeff = rng.lognormal(0.0, 1.0, size=template.nbasis)
        models, ivars = {}, {}
        for cam, wave in grids.items():
            from .rebin import centers2edges
            e = centers2edges(wave)
            tb = antiderivative_rebin(template.wave * (1 + z), template.flux, e)
            T = transmission_Lyman(np.array([z]), wave, model=template.igm_model,
                                   always_return_array=True)
            m = coeff @ (tb * T[0][None, :])
            models[cam] = dias[cam].dot(m)
            iv = rng.lognormal(0.0, 0.5, size=len(wave))
            iv[rng.random(len(wave)) < nodata_frac] = 0.0
            ivars[cam] = iv
        allm = np.concatenate([models[c] for c in grids])
        alliv = np.concatenate([ivars[c] for c in grids])
        ok = alliv > 0
        scale = snr / max(np.median(np.abs(allm[ok]) * np.sqrt(alliv[ok])), 1e-30)
        if mask_frac is not None:
            ntot = len(alliv)
            target_frac = rng.uniform(*mask_frac)
            mask = np.zeros(ntot, dtype=bool)
            while mask.mean() < target_frac:
                L = int(rng.integers(50, 501))
                s0 = int(rng.integers(0, ntot - L))
                mask[s0:s0 + L] = True
            off = 0
            for cam in grids:
                n = len(grids[cam])
                ivars[cam][mask[off:off + n]] = 0.0
                off += n
        spectra = []
        for cam, wave in grids.items():
            iv = ivars[cam]
            sig = np.where(iv > 0, 1.0 / np.sqrt(np.where(iv > 0, iv, 1.0)), 1.0)
            flux = models[cam] * scale + rng.normal(size=len(wave)) * sig
            spectra.append(Spectrum(wave, flux, iv.copy(), dias[cam], band=cam))
        targets.append(Target(first_id + it, spectra))
        ztrue.append(z)
    return targets, np.array(ztrue)


def workload(name, ntarget=None, seed=0, snr=1.0):
    """(template, targets, z_true) for 'galaxy', 'qso' or 'star'."""
    if name == 'galaxy':
        t = galaxy_template()
    elif name == 'qso':
        t = qso_template()
    elif name == 'star':
        t = star_template()
    else:
        raise ValueError(name)
    targets, zt = make_targets(t, 10 if ntarget is None else ntarget, seed=seed, snr=snr)
    return t, targets, zt


def mixed_targets(ntarget, seed=0, snr=1.0, mask_frac=(0.0, 0.3), templates=None):
    """A shuffled mix of GALAXY / QSO / STAR targets with masked pixel runs.

    Thirds of the targets are drawn from each template (templates =
    (galaxy, qso, star) Template objects, built here when None); every target
    gets a masked fraction U(mask_frac) in contiguous runs of 50-500 pixels.
    Returns (targets, spectype per target, z_true); target ids are 0..n-1 in
    the returned order.
    """
    if templates is None:
        templates = (galaxy_template(), qso_template(), star_template())
    counts = [ntarget // 3 + (1 if i < ntarget % 3 else 0) for i in range(3)]
    tg, types, zt = [], [], []
    for k, (t, n) in enumerate(zip(templates, counts)):
        if n == 0:
            continue
        ts, z = make_targets(t, n, seed=seed + 101 * k, snr=snr, mask_frac=mask_frac)
        tg += ts
        types += [t.template_type] * n
        zt += list(z)
    order = np.random.default_rng(seed).permutation(len(tg))
    targets = []
    for new_id, i in enumerate(order):
        targets.append(Target(new_id, tg[i].spectra))
    return targets, [types[i] for i in order], np.array([zt[i] for i in order])


def archetypes_from(template, narch, seed=7, subtypes=None):
    """narch positive archetype spectra on the template's wavelength grid:
    |random mixtures of its basis| (an Archetype held in memory)."""
    from .archetypes import Archetype
    rng = np.random.default_rng(seed)
    c = rng.normal(size=(narch, template.nbasis))
    flux = np.abs(c @ template.flux)
    if subtypes is None:
        subtypes = [('ELG', 'LRG', 'BGS')[i % 3] for i in range(narch)]
    return Archetype.from_arrays(template.template_type, template.wave, flux, subtypes=subtypes,
                                 igm_model=template.igm_model)
