"""GPU chi2 scan (rr_zscan) against the float64 oracle.

Plain FP64: |delta chi2|/chi2 <= 1e-9 and identical per-target best-z index.
Mixed precision (fp16 / bf16 MFMA): tolerances stated per test.
"""
import numpy as np
import pytest

from oracle import rr_oracle as orc

pytestmark = pytest.mark.gpu

F64_RTOL = 1e-9


@pytest.fixture(scope='module')
def galaxy(gpu):
    from redrock_amd import synth
    from redrock_amd.templates import DistTemplate
    tm = synth.galaxy_template()
    targets, zt = synth.make_targets(tm, 3, seed=11)
    dwave = {s.wavehash: s.wave for s in targets[0].spectra}
    dt = DistTemplate(tm, dwave)
    tdata = orc.template_cache(tm, dwave)
    ref = orc.calc_zchi2(targets, tdata, tm.nbasis, tm.OIItemplate)
    return tm, targets, dt, tdata, ref


def test_f64_matches_oracle(galaxy):
    from redrock_amd import zscan
    tm, targets, dt, tdata, (rchi2, rcoeff, rpen) = galaxy
    print('case: galaxy f64, 3 targets x', tm.redshifts.size, 'z', flush=True)
    chi2, coeff, pen = zscan.calc_zchi2([t.id for t in targets], targets, dt, precision='f64')
    rel = np.abs(chi2 - rchi2) / np.abs(rchi2)
    print('max rel chi2 err', rel.max())
    assert rel.max() <= F64_RTOL
    assert np.array_equal(chi2.argmin(axis=1), rchi2.argmin(axis=1))
    assert np.allclose(coeff, rcoeff, rtol=1e-7, atol=1e-9 * np.abs(rcoeff).max())
    assert np.allclose(pen, rpen, rtol=1e-7, atol=1e-9 * max(np.abs(rpen).max(), 1.0))


def test_f64_short_items_match_full_grid(galaxy):
    """Items of few redshifts (4 or 16 items per wavefront, pixel quarters
    summed in a fixed order) and ragged item lengths give the full-grid scan's
    rows to rounding (1e-12), with identical rows for 4 and 16 items per
    wavefront; items of more than 16 redshifts run the full-grid form and give
    its rows bit for bit."""
    from redrock_amd import zscan
    tm, targets, dt, *_ = galaxy
    batch = zscan.TargetBatch(targets)
    nt, nz = len(targets), dt.cache.nz
    full = zscan.ScanPlan(dt.cache, batch, np.arange(nt), np.zeros(nt), np.full(nt, nz), 'f64').run()
    c_full = full.zchi2.cpu().numpy().reshape(nt, nz)
    f_full = full.zcoeff.cpu().numpy().reshape(nt, nz, -1)
    rng = np.random.default_rng(3)
    rows = dict()
    for case in ('ragged', 3, 15, 40, 'g16', 'g4'):
        it = np.repeat(np.arange(nt), 23)
        rng.shuffle(it)
        if case in ('g16', 'g4'):            # the same items, 16 and then 4 per wavefront
            it, L, z0 = rows['g'] if case == 'g4' else (it, rng.integers(1, 5, size=it.size), None)
            if case == 'g16':
                z0 = rng.integers(0, nz - 4, size=it.size)  # room for the 5-redshift item below
                rows['g'] = (it, L.copy(), z0)
            else:
                L = L.copy()
                L[-1] = 5                    # one item of 5 redshifts: 4 items per wavefront
        else:
            L = rng.integers(1, 17, size=it.size) if case == 'ragged' else np.full(it.size, case)
            z0 = rng.integers(0, nz - L + 1)
        print(f'case: f64 short items {case}, {it.size} items', flush=True)
        plan = zscan.ScanPlan(dt.cache, batch, it, z0, L, 'f64').run()
        c, f = plan.zchi2.cpu().numpy(), plan.zcoeff.cpu().numpy()
        want_c = np.concatenate([c_full[t, z:z + n] for t, z, n in zip(it, z0, L)])
        want_f = np.concatenate([f_full[t, z:z + n] for t, z, n in zip(it, z0, L)])
        if case == 40:
            assert np.array_equal(c, want_c), case
            assert np.array_equal(f, want_f), case
        else:
            np.testing.assert_allclose(c, want_c, rtol=1e-12, atol=0, err_msg=str(case))
            np.testing.assert_allclose(f, want_f, rtol=1e-7, atol=1e-9 * np.abs(f_full).max(),
                                       err_msg=str(case))
        rows[case] = (c, f)
    n16 = rows['g16'][0].size - rows['g'][1][-1]      # rows before the last item
    assert np.array_equal(rows['g4'][0][:n16], rows['g16'][0][:n16])
    assert np.array_equal(rows['g4'][1][:n16], rows['g16'][1][:n16])


def test_f64_small_launch_equals_large(galaxy):
    """Short-item launches of fewer than 256 workgroups take a variant that
    loads every basis vector's template window before using the first; rows
    equal the large launch's bit for bit (16 and 4 items per wavefront)."""
    from redrock_amd import zscan
    tm, targets, dt, *_ = galaxy
    batch = zscan.TargetBatch(targets)
    nt, nz = len(targets), dt.cache.nz
    rng = np.random.default_rng(11)
    for lo, hi, n in ((1, 5, 4200), (5, 17, 1100)):     # >= 256 workgroups of 16 / 4 items
        it = rng.integers(0, nt, size=n)
        L = rng.integers(lo, hi, size=n)
        z0 = rng.integers(0, nz - L + 1)
        print(f'case: f64 short items of {lo}-{hi - 1} redshifts, {n} items vs the first 100', flush=True)
        big = zscan.ScanPlan(dt.cache, batch, it, z0, L, 'f64').run()
        small = zscan.ScanPlan(dt.cache, batch, it[:100], z0[:100], L[:100], 'f64').run()
        m = int(L[:100].sum())
        assert np.array_equal(small.zchi2.cpu().numpy(), big.zchi2.cpu().numpy()[:m])
        assert np.array_equal(small.zcoeff.cpu().numpy(), big.zcoeff.cpu().numpy()[:m])


def test_f64_deterministic(galaxy):
    """Run-to-run bit identity (the reference checks mp_procs=1 vs 2, test_zscan.py:113-140)."""
    from redrock_amd import zscan
    tm, targets, dt, *_ = galaxy
    a = zscan.calc_zchi2([0], targets[:2], dt, precision='f64')
    b = zscan.calc_zchi2([0], targets[:2], dt, precision='f64')
    for x, y in zip(a, b):
        assert np.array_equal(x, y)


@pytest.mark.parametrize('precision,tol_max,tol_med', [('f16', 3e-4, 1e-5), ('bf16', 3e-3, 1e-4),
                                                       ('f32', 1e-6, 1e-7), ('f16x3', 1e-6, 1e-7)])
def test_mixed_matches_oracle(galaxy, precision, tol_max, tol_med):
    """Mixed precision at per-pixel S/N 1 (a bright target).

    Single-pass 16-bit inputs give a first-order chi2 error -2 r'W dA c that
    grows with the model power; the stated tolerances are max and median of
    |delta chi2|/chi2 over every (target, z), with identical best-z.
    """
    from redrock_amd import zscan
    tm, targets, dt, tdata, (rchi2, rcoeff, rpen) = galaxy
    print(f'case: galaxy {precision}', flush=True)
    chi2, coeff, pen = zscan.calc_zchi2([t.id for t in targets], targets, dt, precision=precision)
    rel = np.abs(chi2 - rchi2) / np.abs(rchi2)
    ibest = rchi2.argmin(axis=1)
    print(f'{precision}: max rel err {rel.max():.3e}, median {np.median(rel):.3e}, '
          f'at best z {rel[np.arange(3), ibest].max():.3e}')
    assert rel.max() <= tol_max
    assert np.median(rel) <= tol_med
    assert np.array_equal(chi2.argmin(axis=1), ibest)


def test_calc_zchi2_batch_with_prior(gpu):
    """Explicit tdata + prior (zscan.py:711-712,761-762) on a multi-exposure layout."""
    from redrock_amd import synth, zscan
    from redrock_amd.targets import Spectrum, Target
    tm = synth.galaxy_template()
    tg, _ = synth.make_targets(tm, 1, seed=5)
    # two exposures per camera: repeated wavehash in the pixel axis
    spectra = []
    for s in tg[0].spectra:
        for e in range(2):
            spectra.append(Spectrum(s.wave, s.flux + 0.1 * e, s.ivar.copy(), s.R, band=s.band))
    dwave = {s.wavehash: s.wave for s in spectra}
    zsel = tm.redshifts[300:360]
    tdata = orc.template_cache(tm, dwave, myz=zsel)
    w, f, wf = orc.spectral_data(spectra)
    prior = np.diag(np.linspace(0.0, 2.0, tm.nbasis))
    rchi2, rcoeff = orc.calc_zchi2_batch(spectra, tdata, w, f, wf, len(zsel), tm.nbasis, prior=prior)
    chi2, coeff = zscan.calc_zchi2_batch(spectra, tdata, w, f, wf, len(zsel), tm.nbasis, 'PCA',
                                         prior=prior)
    assert np.max(np.abs(chi2 - rchi2) / rchi2) <= F64_RTOL
    assert np.allclose(coeff, rcoeff, rtol=1e-7, atol=1e-9 * np.abs(rcoeff).max())
    # a polish tolerance that takes in every row (15 windows of 4 redshifts)
    # turns a bf16 scan into the FP64 one, prior included
    print('case: bf16 + polish of every row, prior, 60 z', flush=True)
    chi2, coeff = zscan.calc_zchi2_batch(spectra, tdata, w, f, wf, len(zsel), tm.nbasis, 'PCA',
                                         prior=prior, precision='bf16', polish=10.0)
    assert np.max(np.abs(chi2 - rchi2) / rchi2) <= F64_RTOL
    assert np.allclose(coeff, rcoeff, rtol=1e-7, atol=1e-9 * np.abs(rcoeff).max())
    # f16r: coefficients from the fp16 normal equations with the prior, chi2
    # the explicit residual (which, as the reference's, leaves the prior out)
    print('case: f16r, prior, multi-exposure, 60 z', flush=True)
    chi2, coeff = zscan.calc_zchi2_batch(spectra, tdata, w, f, wf, len(zsel), tm.nbasis, 'PCA',
                                         prior=prior, precision='f16r', polish=False)
    assert np.max(np.abs(chi2 - rchi2) / rchi2) <= 1e-6
    assert np.array_equal(chi2.argmin(), rchi2.argmin())


def test_calc_zchi2_batch_bvls_archetype_bounds(gpu):
    """solve_matrices_algorithm 'BVLS' with the archetype bounds of
    zscan.py:289-294 (the first coefficients >= 0, the rest free):
    scipy lsq_linear(M, y, bounds, 'bvls') in the reference (zscan.py:1175-1201).
    FP64 within 1e-9 of the oracle with bounds active; split fp16 within 1e-5
    of f'Wf; bounds of another form are refused."""
    from redrock_amd import synth, zscan
    tm = synth.galaxy_template()
    tg, _ = synth.make_targets(tm, 1, seed=23)
    spectra = tg[0].spectra
    dwave = {s.wavehash: s.wave for s in spectra}
    zsel = tm.redshifts[500:580]
    tdata = orc.template_cache(tm, dwave, myz=zsel)
    w, f, wf = orc.spectral_data(spectra)
    nbound = 4
    bounds = np.zeros((2, tm.nbasis))
    bounds[0][nbound:] = -np.inf
    bounds[1] = np.inf
    rchi2, rcoeff = orc.calc_zchi2_batch(spectra, tdata, w, f, wf, len(zsel), tm.nbasis, bounds=bounds)
    active = np.mean(rcoeff[:, :nbound] == 0)
    print(f'case: BVLS, {nbound} bounded of {tm.nbasis}, {len(zsel)} z; active bounds {active:.2f}', flush=True)
    assert active > 0.05
    args = {'bounds': bounds}
    chi2, coeff = zscan.calc_zchi2_batch(spectra, tdata, w, f, wf, len(zsel), tm.nbasis, 'BVLS',
                                         solver_args=args)
    assert np.max(np.abs(chi2 - rchi2) / rchi2) <= F64_RTOL
    assert np.all(coeff[:, :nbound] >= 0)
    assert np.allclose(coeff, rcoeff, rtol=1e-6, atol=1e-8 * np.abs(rcoeff).max())
    chi2, coeff = zscan.calc_zchi2_batch(spectra, tdata, w, f, wf, len(zsel), tm.nbasis, 'BVLS',
                                         solver_args=args, precision='f16x3')
    assert np.max(np.abs(chi2 - rchi2)) / np.sum(wf * f) <= 1e-5
    assert chi2.argmin() == rchi2.argmin()
    # with a prior on the free coefficients, as the archetype fit passes one
    # (zscan.py:245-348: prior_on_coeffs on the Legendre terms)
    prior = np.diag(np.r_[np.zeros(nbound), np.full(tm.nbasis - nbound, 50.0)])
    rchi2, rcoeff = orc.calc_zchi2_batch(spectra, tdata, w, f, wf, len(zsel), tm.nbasis, prior=prior,
                                         bounds=bounds)
    chi2, coeff = zscan.calc_zchi2_batch(spectra, tdata, w, f, wf, len(zsel), tm.nbasis, 'BVLS',
                                         solver_args=args, prior=prior)
    assert np.max(np.abs(chi2 - rchi2) / rchi2) <= F64_RTOL
    assert np.allclose(coeff, rcoeff, rtol=1e-6, atol=1e-8 * np.abs(rcoeff).max())
    # no bounds: the unconstrained solve, as the PCA scan
    chi2, _ = zscan.calc_zchi2_batch(spectra, tdata, w, f, wf, len(zsel), tm.nbasis, 'BVLS')
    pchi2, _ = zscan.calc_zchi2_batch(spectra, tdata, w, f, wf, len(zsel), tm.nbasis, 'PCA')
    assert np.max(np.abs(chi2 - pchi2) / pchi2) <= F64_RTOL



def test_calc_zchi2_batch_bvls_general_bounds(gpu):
    """'BVLS' with two-sided and one-sided finite bounds on every kind of
    coefficient (lsq_linear(M, y, bounds, 'bvls'), zscan.py:1175-1201): FP64
    within 1e-9 of the oracle, coefficients inside the box with lower and upper
    bounds active, with and without a prior; split fp16 within 1e-5 of f'Wf."""
    from redrock_amd import synth, zscan
    tm = synth.galaxy_template()
    tg, _ = synth.make_targets(tm, 1, seed=29)
    spectra = tg[0].spectra
    dwave = {s.wavehash: s.wave for s in spectra}
    zsel = tm.redshifts[400:480]
    tdata = orc.template_cache(tm, dwave, myz=zsel)
    w, f, wf = orc.spectral_data(spectra)
    _, c0 = orc.calc_zchi2_batch(spectra, tdata, w, f, wf, len(zsel), tm.nbasis)
    a = np.abs(c0).mean(axis=0)                  # typical size of each coefficient
    nb = tm.nbasis
    lo = np.full(nb, -np.inf)
    hi = np.full(nb, np.inf)
    lo[0:3], hi[0:3] = -0.3 * a[0:3], 0.4 * a[0:3]     # boxes
    lo[3:5] = 0.2 * a[3:5]                             # lower only, above zero
    hi[5:7] = -0.1 * a[5:7]                            # upper only, below zero
    bounds = np.array([lo, hi])
    rchi2, rcoeff = orc.calc_zchi2_batch(spectra, tdata, w, f, wf, len(zsel), nb, bounds=bounds)
    at_lo = np.isclose(rcoeff, lo, rtol=0, atol=1e-12 * a.max())
    at_hi = np.isclose(rcoeff, hi, rtol=0, atol=1e-12 * a.max())
    print(f'case: BVLS general bounds, {len(zsel)} z; at lower {at_lo.mean():.2f}, '
          f'at upper {at_hi.mean():.2f}', flush=True)
    assert at_lo.any() and at_hi.any()
    args = {'bounds': bounds}
    chi2, coeff = zscan.calc_zchi2_batch(spectra, tdata, w, f, wf, len(zsel), nb, 'BVLS',
                                         solver_args=args)
    print('max rel chi2 err', np.max(np.abs(chi2 - rchi2) / rchi2))
    assert np.max(np.abs(chi2 - rchi2) / rchi2) <= F64_RTOL
    assert np.all(coeff >= lo) and np.all(coeff <= hi)
    assert np.allclose(coeff, rcoeff, rtol=1e-6, atol=1e-8 * np.abs(rcoeff).max())
    prior = np.diag(np.full(nb, 30.0))
    rchi2p, rcoeffp = orc.calc_zchi2_batch(spectra, tdata, w, f, wf, len(zsel), nb, prior=prior,
                                           bounds=bounds)
    chi2p, coeffp = zscan.calc_zchi2_batch(spectra, tdata, w, f, wf, len(zsel), nb, 'BVLS',
                                           solver_args=args, prior=prior)
    assert np.max(np.abs(chi2p - rchi2p) / rchi2p) <= F64_RTOL
    assert np.allclose(coeffp, rcoeffp, rtol=1e-6, atol=1e-8 * np.abs(rcoeffp).max())
    chi2m, coeffm = zscan.calc_zchi2_batch(spectra, tdata, w, f, wf, len(zsel), nb, 'BVLS',
                                           solver_args=args, precision='f16x3')
    assert np.max(np.abs(chi2m - rchi2)) / np.sum(wf * f) <= 1e-5
    assert np.all(coeffm >= lo) and np.all(coeffm <= hi)


def test_huge_chi2_conventions(gpu):
    """Zero weights -> HUGE everywhere; a vanishing basis -> singular -> HUGE (zscan.py:657-660,769-771)."""
    import torch
    from redrock_amd import synth, zscan
    from redrock_amd.templates import Template, DistTemplate
    base = synth.galaxy_template()
    flux = base.flux.copy()
    flux[3, :] = 0.0   # basis 3 identically zero -> M singular at every z
    tm = Template(spectype='GALAXY', redshifts=base.redshifts[500:530], wave=base.wave,
                  flux=flux, igm_model='None')
    targets, _ = synth.make_targets(base, 2, seed=3)
    for s in targets[1].spectra:
        s.ivar[:] = 0.0
    dwave = {s.wavehash: s.wave for s in targets[0].spectra}
    dt = DistTemplate(tm, dwave)
    for precision in ('f64', 'f16'):
        chi2, coeff, pen = zscan.calc_zchi2([0, 1], targets, dt, precision=precision)
        assert np.all(chi2 == zscan.HUGE_CHI2)
        assert np.all(coeff == 0) and np.all(pen == 0)
    tdata = orc.template_cache(tm, dwave)
    rchi2, _, _ = orc.calc_zchi2(targets, tdata, tm.nbasis)
    assert np.all(rchi2 == orc.HUGE_CHI2)


def test_qso_igm_f64(gpu):
    """QSO with the fused Lyman-alpha IGM (nb=4) against the oracle on a z subset."""
    from redrock_amd import synth, zscan
    from redrock_amd.templates import Template, DistTemplate
    q = synth.qso_template()
    tm = Template(spectype='QSO', redshifts=q.redshifts[2000:2400], wave=q.wave, flux=q.flux,
                  igm_model=q.igm_model)
    targets, _ = synth.make_targets(q, 2, seed=21, zrange=(2.3, 2.9))
    dwave = {s.wavehash: s.wave for s in targets[0].spectra}
    dt = DistTemplate(tm, dwave)
    chi2, coeff, _ = zscan.calc_zchi2([0, 1], targets, dt, precision='f64')
    tdata = orc.template_cache(tm, dwave, lines=[(1215.67, 0.0036, 3.46)])
    rchi2, rcoeff, _ = orc.calc_zchi2(targets, tdata, tm.nbasis)
    assert np.max(np.abs(chi2 - rchi2) / rchi2) <= F64_RTOL
    assert np.array_equal(chi2.argmin(axis=1), rchi2.argmin(axis=1))
    c16, _, _ = zscan.calc_zchi2([0, 1], targets, dt, precision='f16')
    assert np.max(np.abs(c16 - rchi2) / rchi2) <= 1e-4


def test_star_f64(gpu):
    from redrock_amd import synth, zscan
    from redrock_amd.templates import DistTemplate
    tm = synth.star_template()
    targets, _ = synth.make_targets(tm, 2, seed=8)
    dwave = {s.wavehash: s.wave for s in targets[0].spectra}
    dt = DistTemplate(tm, dwave)
    chi2, _, _ = zscan.calc_zchi2([0, 1], targets, dt, precision='f64')
    tdata = orc.template_cache(tm, dwave)
    rchi2, _, _ = orc.calc_zchi2(targets, tdata, tm.nbasis)
    assert np.max(np.abs(chi2 - rchi2) / rchi2) <= F64_RTOL
    assert np.array_equal(chi2.argmin(axis=1), rchi2.argmin(axis=1))


@pytest.mark.parametrize('kind', ['qso4', 'star5', 'gal3', 'gal2'])
@pytest.mark.parametrize('precision,tol_max', [('f16', 1e-3), ('f16x3', 1e-5), ('f32', 1e-5)])
def test_packed_small_basis_matches_oracle(gpu, kind, precision, tol_max):
    """nbasis <= 7: several redshifts share one MFMA tile (4 for nb <= 3, 3 for
    nb 4, 2 for nb 5-7); every redshift of a group must match the oracle, with
    grids whose lengths are not multiples of the group size."""
    from redrock_amd import synth, zscan
    from redrock_amd.templates import Template, DistTemplate
    lines = None
    if kind == 'qso4':
        q = synth.qso_template()
        tm = Template(spectype='QSO', redshifts=q.redshifts[1500:1617], wave=q.wave, flux=q.flux,
                      igm_model=q.igm_model)
        targets, _ = synth.make_targets(q, 3, seed=41, zrange=(1.2, 1.4))
        lines = [(1215.67, 0.0036, 3.46)]
    elif kind == 'star5':
        tm = synth.star_template()
        targets, _ = synth.make_targets(tm, 3, seed=42)
    else:
        g = synth.galaxy_template()
        nb = int(kind[-1])
        tm = Template(spectype='GALAXY', redshifts=g.redshifts[700:803], wave=g.wave,
                      flux=g.flux[:nb], igm_model=g.igm_model)
        targets, _ = synth.make_targets(tm, 3, seed=43, zrange=(0.45, 0.6))
    print(f'case: packed {kind} {precision}, nz={len(tm.redshifts)}', flush=True)
    dwave = {s.wavehash: s.wave for s in targets[0].spectra}
    dt = DistTemplate(tm, dwave)
    chi2, coeff, pen = zscan.calc_zchi2([t.id for t in targets], targets, dt, precision=precision)
    tdata = orc.template_cache(tm, dwave, lines=lines)
    rchi2, rcoeff, rpen = orc.calc_zchi2(targets, tdata, tm.nbasis, getattr(tm, 'OIItemplate', None))
    rel = np.abs(chi2 - rchi2) / np.abs(rchi2)
    print(f'{kind} {precision}: max rel err {rel.max():.3e}')
    assert rel.max() <= tol_max
    assert np.array_equal(chi2.argmin(axis=1), rchi2.argmin(axis=1))
    # every redshift of a group is filled (no stale rows)
    assert np.all(np.isfinite(chi2)) and np.all(chi2 > 0)


@pytest.mark.parametrize('kind', ['galaxy', 'qso'])
def test_ragged_items_every_precision(gpu, kind):
    """Items with their own (target, z0, nz) — the layout fitz and zfind use —
    in every precision: FP64 rows are bit-identical to the full scan's, mixed
    rows within the mode's tolerance of them; nz not a multiple of the
    redshift group (QSO packs 3 per MFMA tile) and single-z items included."""
    from redrock_amd import synth, zscan
    from redrock_amd.templates import DistTemplate
    tm = synth.galaxy_template() if kind == 'galaxy' else synth.qso_template()
    targets, _ = synth.make_targets(tm, 4, seed=51)
    dwave = {s.wavehash: s.wave for s in targets[0].spectra}
    dt = DistTemplate(tm, dwave)
    nz = len(tm.redshifts)
    full, _, _ = zscan.calc_zchi2([t.id for t in targets], targets, dt, precision='f64')
    items = [(0, 0, 37), (1, 100, 1), (2, 500, 64), (0, 1000, 5), (3, nz - 7, 7), (2, 7, 2)]
    it, z0, n = (np.array(c) for c in zip(*items))
    batch = zscan.TargetBatch(targets)
    want = np.concatenate([full[t, a:a + k] for t, a, k in items])
    tol = {'f64': 0.0, 'f16x3': 1e-5, 'f32': 1e-5, 'f16': 2e-3, 'bf16': 1e-2}
    for p, rtol in tol.items():
        print(f'case: ragged {kind} {p}', flush=True)
        plan = zscan.ScanPlan(dt.cache, batch, it, z0, n, precision=p,
                              oiisum=zscan._oiisum(tm))
        plan.run()
        got = plan.zchi2.cpu().numpy()
        rel = np.abs(got - want) / np.abs(want)
        print(f'{kind} {p}: max rel err {rel.max():.2e}')
        if p == 'f64':
            assert np.array_equal(got, want)
        else:
            assert rel.max() <= rtol
    # no items: nothing to do, no error
    empty = zscan.ScanPlan(dt.cache, batch, np.zeros(0, int), np.zeros(0, int), np.zeros(0, int), 'f16x3')
    empty.run()
    assert empty.zchi2.numel() == 0


@pytest.fixture(scope='module')
def nmf(gpu):
    """An NMF template (non-negative basis: |galaxy basis|) fit to targets made
    with the signed basis, so that coefficient bounds are often active."""
    from redrock_amd import synth
    from redrock_amd.templates import Template, DistTemplate
    base = synth.galaxy_template()
    tm = Template(spectype='GALAXY', redshifts=base.redshifts[600:664], wave=base.wave,
                  flux=np.abs(base.flux), method='NMF', igm_model='None')
    targets, _ = synth.make_targets(base, 4, seed=17, zrange=(0.37, 0.40))
    dwave = {s.wavehash: s.wave for s in targets[0].spectra}
    dt = DistTemplate(tm, dwave)
    tdata = orc.template_cache(tm, dwave)
    ref = orc.calc_zchi2(targets, tdata, tm.nbasis, algo='NMF')
    return tm, targets, dt, ref


def test_nnls_template_matches_oracle(nmf):
    """NMF templates (RRMETHOD = NMF) are solved as scipy.optimize.nnls(M, y)
    (zscan.py:1153-1172): FP64 chi2 within 1e-9 of the oracle, same best
    redshift, coefficients >= 0 and equal to the oracle's."""
    from redrock_amd import zscan
    tm, targets, dt, (rchi2, rcoeff, _) = nmf
    print('case: NMF template, 4 targets x 64 z, f64', flush=True)
    chi2, coeff, _ = zscan.calc_zchi2([t.id for t in targets], targets, dt, precision='f64')
    rel = np.abs(chi2 - rchi2) / np.abs(rchi2)
    print(f'nnls: max rel chi2 err {rel.max():.2e}; active bounds {np.mean(rcoeff == 0):.2f}')
    assert np.mean(rcoeff == 0) > 0.1            # the case exercises active bounds
    assert rel.max() <= F64_RTOL
    assert np.array_equal(chi2.argmin(axis=1), rchi2.argmin(axis=1))
    assert np.all(coeff >= 0)
    assert np.allclose(coeff, rcoeff, rtol=1e-6, atol=1e-8 * np.abs(rcoeff).max())


def test_nnls_template_mixed_precision(nmf):
    """The same NMF fit in split fp16 (no polish for NNLS): chi2 within
    1e-5 of the target's f'Wf of the oracle, same best redshift."""
    from redrock_amd import zscan
    tm, targets, dt, (rchi2, rcoeff, _) = nmf
    print('case: NMF template, f16x3', flush=True)
    chi2, coeff, _ = zscan.calc_zchi2([t.id for t in targets], targets, dt, precision='f16x3')
    fwf = np.array([np.sum(s_ivar * s_flux ** 2) for s_ivar, s_flux in
                    ((np.concatenate([s.ivar for s in t.spectra]),
                      np.concatenate([s.flux for s in t.spectra])) for t in targets)])
    err = np.abs(chi2 - rchi2) / fwf[:, None]
    print(f'nnls f16x3: max |dchi2|/fWf {err.max():.2e}')
    assert err.max() <= 1e-5
    assert np.array_equal(chi2.argmin(axis=1), rchi2.argmin(axis=1))
    assert np.all(coeff >= 0)
