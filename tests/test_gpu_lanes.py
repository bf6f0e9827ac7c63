"""Lane-per-item scan kernels (zscan_lanes.hip) against the oracle and the
redshift-per-lane FP64 kernel.

* FP64, one item per lane: the same doubles as the redshift-per-lane kernel
  (same operation order), so bit-identical rows; <= 1e-9 of the oracle with
  identical best-z.
* 'f16r' (fp16 normal equations for the coefficients, chi2 as the explicit
  FP32 residual at them): |delta chi2|/chi2 <= 1e-6 of the oracle at every
  (target, z) and identical best-z, GALAXY, QSO with the IGM, STAR, and at
  per-pixel S/N 10; HUGE_CHI2 conventions kept.
"""
import numpy as np
import pytest

from oracle import rr_oracle as orc

pytestmark = pytest.mark.gpu

F64_RTOL = 1e-9
F16R_RTOL = 1e-6
LYA = [(1215.67, 0.0036, 3.46)]


def _setup(kind, nt, seed, snr=1.0, zsel=None):
    from redrock_amd import synth
    from redrock_amd.templates import DistTemplate
    tm = {'galaxy': synth.galaxy_template, 'qso': synth.qso_template,
          'star': synth.star_template}[kind]()
    if zsel is not None:
        tm._redshifts = tm.redshifts[zsel]
    targets, _ = synth.make_targets(tm, nt, seed=seed, snr=snr)
    dwave = {s.wavehash: s.wave for s in targets[0].spectra}
    return tm, targets, dwave, DistTemplate(tm, dwave)


def _plan(dt, targets, precision, items=None, **kw):
    from redrock_amd import zscan
    batch = zscan.TargetBatch(targets)
    nt, nz = len(targets), dt.cache.nz
    it = np.arange(nt) if items is None else np.asarray(items)
    plan = zscan.ScanPlan(dt.cache, batch, it, np.zeros(len(it)), np.full(len(it), nz), precision,
                          **kw).run()
    n = len(it)
    return (plan.zchi2.cpu().numpy().reshape(n, nz), plan.zcoeff.cpu().numpy().reshape(n, nz, -1),
            plan.penalty.cpu().numpy().reshape(n, nz), plan)


@pytest.fixture(scope='module')
def galaxy70(gpu):
    # 70 targets: two lane groups, the second partly idle; one target without weight
    tm, targets, dwave, dt = _setup('galaxy', 70, seed=21, zsel=slice(500, 820))
    for s in targets[5].spectra:
        s.ivar[:] = 0.0
    return tm, targets, dwave, dt


def test_f64_lanes_equal_zfast_kernel(galaxy70):
    tm, targets, dwave, dt = galaxy70
    print('case: f64 lanes vs z-per-lane, 70 galaxy targets x 320 z', flush=True)
    a = _plan(dt, targets, 'f64', lanes=True)
    b = _plan(dt, targets, 'f64', lanes=False)
    assert a[3].lanes and not b[3].lanes
    for x, y in zip(a[:3], b[:3]):
        assert np.array_equal(x, y)
    assert np.all(a[0][5] == 9999999.0) and np.all(a[1][5] == 0.0)


def test_f64_lanes_shuffled_items(galaxy70):
    """Items in any target order (lanes then gather their per-target data)."""
    tm, targets, dwave, dt = galaxy70
    rng = np.random.default_rng(4)
    items = rng.permutation(np.r_[np.arange(70), np.arange(0, 70, 3)])
    print(f'case: f64 lanes, {items.size} shuffled items', flush=True)
    full = _plan(dt, targets, 'f64', lanes=True)
    a = _plan(dt, targets, 'f64', items=items, lanes=True)
    assert np.array_equal(a[0], full[0][items])
    assert np.array_equal(a[1], full[1][items])


def test_f64_lanes_match_oracle(galaxy70):
    tm, targets, dwave, dt = galaxy70
    sel = [0, 33, 64, 69]
    print('case: f64 lanes vs oracle, targets', sel, flush=True)
    chi2, coeff, pen, _ = _plan(dt, targets, 'f64', lanes=True)
    tdata = orc.template_cache(tm, dwave)
    rchi2, rcoeff, rpen = orc.calc_zchi2([targets[i] for i in sel], tdata, tm.nbasis, tm.OIItemplate)
    rel = np.abs(chi2[sel] - rchi2) / np.abs(rchi2)
    print('max rel chi2 err', rel.max())
    assert rel.max() <= F64_RTOL
    assert np.array_equal(chi2[sel].argmin(axis=1), rchi2.argmin(axis=1))
    assert np.allclose(coeff[sel], rcoeff, rtol=1e-7, atol=1e-9 * np.abs(rcoeff).max())


@pytest.mark.parametrize('kind,snr,zsel', [('galaxy', 1.0, slice(0, None, 6)),
                                           ('galaxy', 10.0, slice(0, None, 6)),
                                           ('qso', 1.0, slice(0, None, 10)),
                                           ('star', 3.0, slice(None))])
def test_f16r_matches_oracle(gpu, kind, snr, zsel):
    tm, targets, dwave, dt = _setup(kind, 3, seed=31, snr=snr, zsel=zsel)
    print(f'case: f16r {kind} S/N {snr}, 3 targets x {tm.redshifts.size} z', flush=True)
    chi2, coeff, pen, plan = _plan(dt, targets, 'f16r', polish=False)
    tdata = orc.template_cache(tm, dwave, lines=LYA if kind == 'qso' else None)
    rchi2, rcoeff, rpen = orc.calc_zchi2(targets, tdata, tm.nbasis,
                                         getattr(tm, 'OIItemplate', None))
    rel = np.abs(chi2 - rchi2) / np.abs(rchi2)
    print(f'f16r {kind} S/N {snr}: max rel err {rel.max():.3e}, median {np.median(rel):.3e}')
    assert rel.max() <= F16R_RTOL
    assert np.array_equal(chi2.argmin(axis=1), rchi2.argmin(axis=1))
    # chi2 at approximate coefficients is never below the least-squares minimum
    # (beyond FP32 rounding of the residual sum)
    assert np.all(chi2 >= rchi2 * (1 - 2e-7))


def test_f16r_conventions_and_polish(galaxy70):
    """Zero weight -> HUGE_CHI2 and zero coefficients; the FP64 polish leaves
    the best redshift of every target equal to the FP64 scan's."""
    tm, targets, dwave, dt = galaxy70
    print('case: f16r 70 targets, polished vs f64', flush=True)
    chi2, coeff, pen, plan = _plan(dt, targets, 'f16r')
    ref = _plan(dt, targets, 'f64')
    assert plan.polish is not None
    assert np.all(chi2[5] == 9999999.0) and np.all(coeff[5] == 0.0)
    ok = np.arange(70) != 5
    rel = np.abs(chi2[ok] - ref[0][ok]) / ref[0][ok]
    print('f16r vs f64 max rel', rel.max())
    assert rel.max() <= F16R_RTOL
    assert np.array_equal(chi2.argmin(axis=1), ref[0].argmin(axis=1))
    assert np.array_equal((chi2 + pen).argmin(axis=1), (ref[0] + ref[2]).argmin(axis=1))


@pytest.mark.parametrize('kind', ['galaxy', 'qso'])
def test_every_lane_configuration(galaxy70, kind):
    """Every configuration the autotuner may pick: the FP64 lane kernel's
    (waves per workgroup) give identical bits; the residual stage's -- the
    matrix-core kernel's (default; redshift blocks and waves per workgroup,
    identical bits to each other) and the FP32 lane kernel's (redshifts per
    wave, waves, pixels per LDS block) -- differ only by rounding and stay
    within the f16r bound of the FP64 scan, HUGE_CHI2 rows kept."""
    import torch
    from redrock_amd import _hip, autotune
    if kind == 'galaxy':
        tm, targets, dwave, dt = galaxy70
    else:
        tm, targets, dwave, dt = _setup('qso', 70, seed=22, zsel=slice(0, None, 12))
        for s in targets[5].spectra:
            s.ivar[:] = 0.0
    nz = dt.cache.nz
    c64, _, _, p64 = _plan(dt, targets, 'f64', lanes=True)
    for t in range(1, len(autotune.tile_table('f64'))):
        p64.desc.tile = t
        p64.run()
        torch.cuda.synchronize()
        print(f'case: {kind} f64 lane tile {t}', flush=True)
        assert np.array_equal(p64.zchi2.cpu().numpy().reshape(70, nz), c64)
    c0, _, _, plan = _plan(dt, targets, 'f16r', polish=False, resid='both')
    ok = np.arange(70) != 5
    table = autotune.resid_table()
    assert table[-1][0] == 2
    for t in range(1, len(table)):
        plan.desc.resid_tile = t
        plan.run(stages=_hip.RR_STAGE_RESID)
        torch.cuda.synchronize()
        c = plan.zchi2.cpu().numpy().reshape(70, nz)
        rel0 = np.abs(c[ok] - c0[ok]) / c0[ok]
        rel = np.abs(c[ok] - c64[ok]) / c64[ok]
        print(f'case: {kind} residual tile {t} {table[t]}: vs default {rel0.max():.2e}, '
              f'vs f64 {rel.max():.2e}', flush=True)
        if table[t][0] == 2:            # matrix-core kernel: same arithmetic as the default
            assert np.array_equal(c, c0)
        assert rel0.max() <= 1e-6       # FP32 lane kernel: its own rounding
        assert rel.max() <= F16R_RTOL
        assert np.all(c[5] == 9999999.0)


def _basis(nb):
    """A GALAXY-like basis of nb vectors: the synthetic galaxy template's first
    nb, or for nb > 10 those 10 plus extra emission-line components."""
    from redrock_amd import synth
    from redrock_amd.templates import Template
    g = synth.galaxy_template()
    if nb <= g.nbasis:
        flux = g.flux[:nb]
    else:
        rng = np.random.default_rng(nb)
        extra = [synth._gauss(g.wave, mu, 2.0) for mu in (3869.0, 6717.0)][:nb - g.nbasis]
        extra = synth._with_texture(extra, g.wave, rng, 0.3)
        flux = synth._orthonormalise(list(g.flux) + list(extra), rng, mix=False)
    return Template(spectype='GALAXY', redshifts=g.redshifts[700:745], wave=g.wave, flux=flux,
                    igm_model=g.igm_model)


@pytest.mark.parametrize('nb', [1, 2, 3, 7, 8, 9, 11, 12])
def test_f16r_every_basis_size(gpu, nb):
    """The matrix-core residual at every basis size it instantiates: the
    products u = T c take 3 or 4 hi/lo terms in one or two MFMAs per tile
    (rr_resid_u_slots), items in 16-wide tiles (34 targets: the last tile
    partial), 45 redshifts (the last 32-redshift block partial); within the
    f16r bound of the FP64 scan (itself within 1e-9 of the oracle) and of the
    oracle on three targets, same best redshift."""
    from redrock_amd import _hip, synth
    from redrock_amd.templates import DistTemplate
    tm = _basis(nb)
    assert tm.nbasis == nb
    targets, _ = synth.make_targets(synth.galaxy_template(), 34, seed=60 + nb, zrange=(0.45, 0.6))
    dwave = {s.wavehash: s.wave for s in targets[0].spectra}
    dt = DistTemplate(tm, dwave)
    ks, nprod = _hip.resid_u_slots(nb)
    print(f'case: f16r nb {nb} (K slots {ks}, {nprod} products), 34 targets x 45 z', flush=True)
    chi2, _, _, plan = _plan(dt, targets, 'f16r', polish=False)
    assert plan.desc.d_tcache_u
    c64 = _plan(dt, targets, 'f64')[0]
    rel = np.abs(chi2 - c64) / c64
    print(f'nb {nb}: f16r vs f64 max rel {rel.max():.2e}')
    assert rel.max() <= F16R_RTOL
    assert np.array_equal(chi2.argmin(axis=1), c64.argmin(axis=1))
    sel = [0, 17, 33]
    rchi2, _, _ = orc.calc_zchi2([targets[i] for i in sel], orc.template_cache(tm, dwave), nb)
    assert (np.abs(chi2[sel] - rchi2) / rchi2).max() <= F16R_RTOL
    assert np.all(chi2 >= c64 * (1 - 2e-7))
