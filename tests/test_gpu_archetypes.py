"""Archetype fit on the GPU (rr_arch_prep / rr_arch_fit behind
redrock_amd.archetypes) against the oracle's restatement of
get_best_archetype (archetypes.py:241-400) with scipy's lsq_linear BVLS."""
import os

import numpy as np
import pytest

from oracle import rr_oracle as orc

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), 'golden')


def _oracle(arch, tg, z, nleg, per_camera, n_nearest, prior, trans=None):
    dwave = {s.wavehash: s.wave for s in tg.spectra}
    bands = tg.bands if len(tg.bands) == len(dwave) else list(dwave.keys())
    return orc.get_best_archetype(tg.spectra, bands, arch.wave, arch.flux, list(arch.full_type), z,
                                  nleg, per_camera, n_nearest=n_nearest, prior=prior, trans=trans)


def _compare(got, want, tol_chi2=1e-9):
    gc2, gcf, gty = got
    wc2, wcf, wty = want
    assert gty == wty
    assert abs(gc2 - wc2) <= tol_chi2 * abs(wc2), (gc2, wc2)
    assert gcf.shape == wcf.shape
    # coefficient error ~ cond(M) eps of the solve, relative to the vector
    assert np.linalg.norm(gcf - wcf) <= 1e-7 * np.linalg.norm(wcf), (gcf, wcf)


def _galaxy_setup(ntarget=3, seed=61, narch=20, reorder=False):
    from redrock_amd import synth
    from redrock_amd.targets import Target
    g = synth.galaxy_template()
    targets, zt = synth.make_targets(g, ntarget, seed=seed, zrange=(0.3, 0.9), snr=2.0)
    if reorder:      # spectra r, z, b: the per-camera blocks come back in band order
        targets = [Target(t.id, [t.spectra[1], t.spectra[2], t.spectra[0]]) for t in targets]
    arch = synth.archetypes_from(g, narch, seed=seed)
    return g, targets, zt, arch


@pytest.mark.parametrize('per_camera,nleg,n_nearest,sigma,reorder', [
    (True, 2, None, 0.1, False),
    (True, 2, None, None, True),
    (True, 3, 3, 0.1, True),
    (False, 3, None, None, False),
    (False, 0, None, None, False),
    (False, 2, 2, None, False),
    (True, 1, 2, 0.5, False),
])
def test_fit_matches_oracle(gpu, per_camera, nleg, n_nearest, sigma, reorder):
    from redrock_amd.archetypes import ArchetypeFit, prior_on_coeffs
    g, targets, zt, arch = _galaxy_setup(reorder=reorder)
    ncam = 3 if per_camera else 1
    prior = prior_on_coeffs(n_nearest or 1, nleg, sigma, ncam) if sigma else None
    print(f'case: archetype fit per_camera={per_camera} nleg={nleg} n_nearest={n_nearest} '
          f'sigma={sigma} reorder={reorder}', flush=True)
    fit = ArchetypeFit(arch, targets, per_camera=per_camera, nleg=nleg, prior=prior, n_nearest=n_nearest)
    tidx = np.repeat(np.arange(len(targets)), 3)
    z = np.concatenate([[zt[i] - 0.01, zt[i], zt[i] + 0.02] for i in range(len(targets))])
    chi2, coeff, ftype = fit.run(tidx, z)
    for k, (t, zk) in enumerate(zip(tidx, z)):
        want = _oracle(arch, targets[t], zk, nleg, per_camera, n_nearest, prior)
        _compare((chi2[k], coeff[k], ftype[k]), want)


def test_bounds_active(gpu):
    """Targets whose flux is minus an archetype: the unconstrained archetype
    coefficient is negative, the bound is active for most archetypes."""
    from redrock_amd import synth
    from redrock_amd.archetypes import ArchetypeFit, prior_on_coeffs
    from redrock_amd.targets import Target, Spectrum
    g, targets, zt, arch = _galaxy_setup(ntarget=2, seed=62)
    neg = []
    for t, z in zip(targets, zt):
        sp = []
        for s in t.spectra:
            from redrock_amd.rebin import trapz_rebin
            m = s.R.dot(trapz_rebin((1 + z) * arch.wave, arch.flux[0], s.wave))
            sp.append(Spectrum(s.wave, -5 * m / np.abs(m).mean() + s.flux, s.ivar, s.R, band=s.band))
        neg.append(Target(t.id, sp))
    prior = prior_on_coeffs(1, 2, 0.1, 3)
    fit = ArchetypeFit(arch, neg, per_camera=True, nleg=2, prior=prior)
    chi2, coeff, ftype = fit.run(np.arange(2), zt)
    for k in range(2):
        want = _oracle(arch, neg[k], zt[k], 2, True, None, prior)
        _compare((chi2[k], coeff[k], ftype[k]), want)
    # every archetype alone, against the oracle; many sit at the bound
    ac2, acf = fit.run_all(np.arange(2), zt)
    at_bound = 0
    for k in range(2):
        wc2, wcf, _ = orc.archetype_scan(neg[k].spectra, neg[k].bands, arch.wave, arch.flux, zt[k], 2, True,
                                         prior)
        assert np.allclose(ac2[k], wc2, rtol=1e-9, atol=0)
        err = np.linalg.norm(acf[k] - wcf, axis=1) / np.linalg.norm(wcf, axis=1)
        assert err.max() <= 1e-7
        assert np.array_equal(acf[k][:, 0] == 0.0, wcf[:, 0] == 0.0)
        at_bound += int(np.sum(wcf[:, 0] == 0.0))
    print('archetype coefficients at the bound:', at_bound, 'of', 2 * arch.narch)
    assert at_bound >= 5


def test_igm_at_its_own_redshift(gpu):
    """QSO-like archetypes at z ~ 2.5 with the Lyman-alpha factor taken at a
    different redshift (fitz passes the transmission of the parabola vertex)."""
    from redrock_amd import synth, igm
    from redrock_amd.archetypes import ArchetypeFit
    q = synth.qso_template()
    targets, zt = synth.make_targets(q, 2, seed=63, zrange=(2.3, 2.7), snr=2.0)
    arch = synth.archetypes_from(q, 12, seed=63)
    lines = igm.lyman_table(q.igm_model)
    fit = ArchetypeFit(arch, targets, per_camera=True, nleg=2)
    zigm = np.array([zt[0] + 0.001, np.nan])
    chi2, coeff, ftype = fit.run(np.arange(2), zt, zigm=zigm, lines=lines)
    for k in range(2):
        trans = None
        if zigm[k] == zigm[k]:
            trans = {s.wavehash: orc.transmission_Lyman(np.array([zigm[k]]), s.wave, lines=lines)[0]
                     for s in targets[k].spectra}
        want = _oracle(arch, targets[k], zt[k], 2, True, None, None, trans=trans)
        _compare((chi2[k], coeff[k], ftype[k]), want)


def test_reference_archetype_file_and_api(gpu):
    """The reference's archetype test file and get_best_archetype's one-target
    signature (test_archetypes.py:59-95 of the reference: two cameras with two
    exposures each, bands b and r): coefficient counts as there, values as
    the oracle."""
    from redrock_amd import synth
    from redrock_amd.archetypes import Archetype, prior_on_coeffs
    from redrock_amd.targets import Target, Spectrum
    from redrock_amd.zscan import spectral_data
    arch = Archetype(os.path.join(GOLDEN, 'rrarchetype-galaxy-test.fits'))
    g = synth.galaxy_template()
    base, zt = synth.make_targets(g, 1, seed=64, zrange=(0.5, 0.5), snr=3.0)
    s0, s1 = base[0].spectra[0], base[0].spectra[1]
    rng = np.random.default_rng(0)
    sp = [Spectrum(s.wave, s.flux + rng.normal(size=s.nwave) * 0.1, s.ivar, s.R, band=s.band)
          for s in (s0, s0, s1, s1)]
    tg = Target(123, sp)
    tg.bands = ['b', 'r']
    tg.nleg = 2
    dwave = {s.wavehash: s.wave for s in tg.spectra}
    w, f, wf = spectral_data(tg.spectra)
    prior = prior_on_coeffs(1, tg.nleg, 0.1, len(dwave))
    print('case: reference archetype file, one-target API', flush=True)
    for per_camera, ncoeff in ((True, 1 + 2 * 2), (False, 1 + 2)):
        got = arch.get_best_archetype(tg, w, f, wf, dwave, z=0.5, per_camera=per_camera, n_nearest=None,
                                      trans=None, solve_method='bvls', prior=prior)
        assert len(got[1]) == ncoeff
        want = orc.get_best_archetype(tg.spectra, ['b', 'r'], arch.wave, arch.flux, list(arch.full_type),
                                      0.5, 2, per_camera, prior=prior if per_camera else None)
        _compare(got, want)
    tg.nleg = 0
    got = arch.get_best_archetype(tg, w, f, wf, dwave, z=0.5, per_camera=False, n_nearest=None)
    assert len(got[1]) == 1
    want = orc.get_best_archetype(tg.spectra, ['b', 'r'], arch.wave, arch.flux, list(arch.full_type),
                                  0.5, 0, False)
    _compare(got, want)


def test_zero_weight_target(gpu):
    from redrock_amd.archetypes import ArchetypeFit
    from redrock_amd.targets import Target, Spectrum
    g, targets, zt, arch = _galaxy_setup(ntarget=2, seed=65)
    t1 = targets[1]
    targets[1] = Target(t1.id, [Spectrum(s.wave, s.flux, np.zeros_like(s.ivar), s.R, band=s.band)
                                for s in t1.spectra])
    fit = ArchetypeFit(arch, targets, per_camera=True, nleg=2)
    chi2, coeff, ftype = fit.run(np.arange(2), zt)
    assert chi2[1] == 9999999.0 and np.all(coeff[1] == 0)
    want = _oracle(arch, targets[0], zt[0], 2, True, None, None)
    _compare((chi2[0], coeff[0], ftype[0]), want)


def test_zfind_archetype_mode(gpu, tmp_path):
    """zfind with GALAXY archetypes (zfind.py:330-367, 476-693): every GALAXY
    minimum carries the archetype fit at its redshift (oracle), the other
    types keep their PCA fits, and the archetype-mode columns follow the
    reference's rules."""
    from redrock_amd import synth, zfind as zf, results
    from redrock_amd.archetypes import prior_on_coeffs
    from redrock_amd.templates import Template, DistTemplate
    from redrock_amd.targets import DistTargetsCopy
    from redrock_amd.zwarning import ZWarningMask as ZW
    g, q, s = synth.galaxy_template(), synth.qso_template(), synth.star_template()
    tg = Template(spectype='GALAXY', redshifts=g.redshifts[600:900], wave=g.wave, flux=g.flux,
                  igm_model=g.igm_model)
    tq = Template(spectype='QSO', redshifts=q.redshifts[300:700], wave=q.wave, flux=q.flux,
                  igm_model=q.igm_model)
    ts = Template(spectype='STAR', subtype='K', redshifts=s.redshifts, wave=s.wave, flux=s.flux,
                  igm_model='None')
    targets, _ = synth.make_targets(g, 3, seed=66, zrange=(0.38, 0.58), snr=1.5)
    dwave = {sp.wavehash: sp.wave for sp in targets[0].spectra}
    dts = [DistTemplate(t, dwave) for t in (tg, tq, ts)]
    arch = synth.archetypes_from(g, 16, seed=66)
    print('case: zfind archetype mode, 3 targets, 16 GALAXY archetypes', flush=True)
    _, zt = zf.zfind(DistTargetsCopy(targets), dts, nminima=3, archetypes={'GALAXY': arch},
                     per_camera=True, deg_legendre=2, prior_sigma=0.1, precision='f64')
    n = len(zt['z'])
    gal = zt['spectype'] == 'GALAXY'
    assert gal.sum() >= 3 and (~gal).sum() >= 3
    assert np.all(zt['fitmethod'][gal] == 'ARCH') and np.all(zt['fitmethod'][~gal] == 'PCA')
    assert zt['legcoeff'].shape == (n, 3, 2) and np.all(zt['legcoeff'][~gal] == 0)
    assert zt['pca_coeff'].shape == (n, 10)
    assert list(zt.keys())[-3:] == ['pca_coeff', 'pca_spectype', 'pca_subtype']
    assert np.all(zt['pca_spectype'][gal] == 'GALAXY')
    assert np.all(zt['pca_spectype'][~gal] == zt['spectype'][~gal])
    # COEFF width: max(nbasis 10, ncamera * deg + 1 = 7) -> 10
    assert zt['coeff'].shape == (n, 10) and np.all(zt['ncoeff'][gal] == 1)
    prior = prior_on_coeffs(1, 2, 0.1, 3)
    tmap = {t.id: t for t in targets}
    for k in np.where(gal)[0]:
        t = tmap[zt['targetid'][k]]
        wc2, wcf, wty = orc.get_best_archetype(t.spectra, t.bands, arch.wave, arch.flux, list(arch.full_type),
                                               zt['z'][k], 2, True, prior=prior)
        assert abs(zt['chi2'][k] - wc2) <= 1e-9 * wc2
        assert 'GALAXY:::' + zt['subtype'][k] == wty
        assert abs(zt['coeff'][k, 0] - wcf[0]) <= 1e-7 * np.linalg.norm(wcf)
        assert np.allclose(zt['legcoeff'][k].ravel(), wcf[1:], rtol=1e-6, atol=1e-7 * np.linalg.norm(wcf))
        assert not (zt['zwarn'][k] & ZW.NEGATIVE_MODEL)        # bounded: coeff >= 0
    for tid in np.unique(zt['targetid']):
        sel = zt['targetid'] == tid
        good = (zt['zwarn'][sel] & (ZW.BAD_MINFIT | ZW.Z_FITLIMIT | ZW.NEGATIVE_MODEL)) == 0
        c2 = zt['chi2'][sel][good]
        assert np.all(np.diff(c2) >= 0)
    path = str(tmp_path / 'rrdetails.h5')
    results.write_zscan(path, _, zt)
    back = results.read_zfit(path)
    assert np.array_equal(back['subtype'], zt['subtype']) and back['legcoeff'].shape == (n, 3, 2)


def test_fused_rebin_equals_cache(gpu):
    """The fit kernel's own bin means (the trapezoid as weights of the samples)
    against rr_rebin's cache: chi2 and coefficients equal to rounding, with
    IGM and with the n-nearest combination."""
    from redrock_amd import synth, igm
    from redrock_amd.archetypes import ArchetypeFit, prior_on_coeffs
    q = synth.qso_template()
    targets, zt = synth.make_targets(q, 3, seed=67, zrange=(2.2, 2.8), snr=2.0)
    arch = synth.archetypes_from(q, 70, seed=67)
    lines = igm.lyman_table(q.igm_model)
    print('case: fused rebin vs cache, 3 QSO targets, 70 archetypes', flush=True)
    for nn in (None, 3):
        prior = prior_on_coeffs(nn or 1, 2, 0.2, 3)
        fit = ArchetypeFit(arch, targets, per_camera=True, nleg=2, prior=prior, n_nearest=nn)
        tidx = np.repeat(np.arange(3), 2)
        z = np.concatenate([[zt[i], zt[i] + 0.013] for i in range(3)])
        zigm = z + 0.0004
        a = fit.run(tidx, z, zigm=zigm, lines=lines, fused=True)
        b = fit.run(tidx, z, zigm=zigm, lines=lines, fused=False)
        assert np.allclose(a[0], b[0], rtol=1e-12, atol=0) and a[2] == b[2]
        for ca, cb in zip(a[1], b[1]):
            assert np.linalg.norm(ca - cb) <= 1e-10 * np.linalg.norm(cb)
    ac2, acf = fit.run_all(tidx, z, zigm=zigm, lines=lines, fused=True)
    bc2, bcf = fit.run_all(tidx, z, zigm=zigm, lines=lines, fused=False)
    assert np.allclose(ac2, bc2, rtol=1e-12, atol=0)
    assert np.all(np.linalg.norm(acf - bcf, axis=2) <= 1e-10 * np.linalg.norm(bcf, axis=2))


def test_fused_rebin_many_samples_per_pixel(gpu):
    """An archetype set sampled at 0.05 A: more than six samples fall inside a
    pixel, so the fit kernel takes its exact-expression path for those bins;
    the result still matches the cached rebin and the oracle."""
    from redrock_amd import synth
    from redrock_amd.archetypes import Archetype, ArchetypeFit
    g = synth.galaxy_template()
    targets, zt = synth.make_targets(g, 2, seed=68, zrange=(0.4, 0.6), snr=2.0)
    fine = np.arange(g.wave[0], g.wave[-1], 0.05)
    base = synth.archetypes_from(g, 6, seed=68)
    arch = Archetype.from_arrays('GALAXY', fine, np.array([np.interp(fine, base.wave, f) for f in base.flux]))
    print('case: fused rebin exact path, 2 targets, 6 archetypes at 0.05 A', flush=True)
    fit = ArchetypeFit(arch, targets, per_camera=True, nleg=2)
    a = fit.run(np.arange(2), zt, fused=True)
    b = fit.run(np.arange(2), zt, fused=False)
    assert np.allclose(a[0], b[0], rtol=1e-12, atol=0) and a[2] == b[2]
    for k in range(2):
        _compare((a[0][k], a[1][k], a[2][k]), _oracle(arch, targets[k], zt[k], 2, True, None, None))
