"""Parity at the benchmark configurations' full sizes.

Per-(target, z) comparison with the oracle is only affordable for a few
targets; at full size the mixed modes are held to the FP64 GPU scan (itself
checked against the oracle on the configuration's CPU case and bit-identical
to the redshift-per-lane FP64 kernel) through size-independent properties:
the per-target best redshift (identical on >= 99.9 % of targets) and the
largest |delta chi2| / chi2 over every (target, z).

* galaxy petal: 500 GALAXY targets x 1977 z (nb 10), per-pixel S/N 1; the
  f16r per-S/N bound |delta chi2| <= 1e-6 f'Wf at S/N 3 and 10
* QSO petal: 500 QSO targets x 3795 z (nb 4), Lyman-alpha IGM fused
* mixed batch: 5000 GALAXY / QSO / STAR targets, 0-30 % masked in runs of
  50-500 px, all three templates in one captured HIP graph, tiles autotuned
"""
import numpy as np
import pytest

from oracle import rr_oracle as orc

pytestmark = pytest.mark.gpu

LYA = [(1215.67, 0.0036, 3.46)]
BESTZ_MIN = 0.999


def _scan(dt, batch, nt, precision, **kw):
    from redrock_amd import zscan
    nz = dt.cache.nz
    oii = zscan._oiisum(dt.template)
    plan = zscan.ScanPlan(dt.cache, batch, np.arange(nt), np.zeros(nt), np.full(nt, nz), precision,
                          oiisum=oii, **kw).run()
    return plan.zchi2.reshape(nt, nz).cpu().numpy(), plan


def _petal(kind, nt=500, snr=1.0):
    from redrock_amd import synth, zscan
    from redrock_amd.templates import DistTemplate
    tm = synth.galaxy_template() if kind == 'galaxy' else synth.qso_template()
    targets, _ = synth.make_targets(tm, nt, seed=0, snr=snr)
    dwave = {s.wavehash: s.wave for s in targets[0].spectra}
    return tm, targets, dwave, DistTemplate(tm, dwave), zscan.TargetBatch(targets)


def _agree(a, b):
    return float(np.mean(a.argmin(axis=1) == b.argmin(axis=1)))


@pytest.fixture(scope='module')
def galaxy_petal(gpu):
    return _petal('galaxy')


def test_galaxy_petal_f64_matches_oracle_on_cpu_case(galaxy_petal):
    """Config 1 (the reference's CPU case: 10 targets x 1977 z) inside the
    petal scan: <= 1e-9 of the oracle, identical best-z index."""
    tm, targets, dwave, dt, batch = galaxy_petal
    print('case: galaxy petal f64 (500 x 1977), oracle on 10 targets', flush=True)
    c64, plan = _scan(dt, batch, len(targets), 'f64')
    assert plan.lanes
    tdata = orc.template_cache(tm, dwave)
    rchi2, _, _ = orc.calc_zchi2(targets[:10], tdata, tm.nbasis, tm.OIItemplate)
    rel = np.abs(c64[:10] - rchi2) / rchi2
    print('max rel chi2 err vs oracle', rel.max())
    assert rel.max() <= 1e-9
    assert np.array_equal(c64[:10].argmin(axis=1), rchi2.argmin(axis=1))
    zl, _ = _scan(dt, batch, len(targets), 'f64', lanes=False)
    assert np.array_equal(c64, zl)


@pytest.mark.parametrize('precision', ['f16r', 'f16x3'])
def test_galaxy_petal_mixed(galaxy_petal, precision):
    tm, targets, dwave, dt, batch = galaxy_petal
    nt = len(targets)
    print(f'case: galaxy petal {precision} vs f64, {nt} x {dt.cache.nz}', flush=True)
    c64, _ = _scan(dt, batch, nt, 'f64')
    cm, _ = _scan(dt, batch, nt, precision)
    rel = np.abs(cm - c64) / c64
    agree = _agree(cm, c64)
    print(f'{precision}: max rel {rel.max():.2e}, best-z agreement {agree:.4f}')
    assert agree >= BESTZ_MIN
    assert rel.max() <= 1e-5


@pytest.mark.parametrize('snr', [3.0, 10.0])
def test_galaxy_petal_f16r_snr_bound(gpu, snr):
    """Bright targets: every mixed mode's absolute chi2 error scales with
    f'Wf, so the bound is stated per f'Wf: |delta chi2| <= 1e-6 f'Wf at every
    (target, z), which is <= 1e-5 of chi2 wherever f'Wf <= 10 chi2; the best
    redshift agrees on >= 99.9 % of targets."""
    tm, targets, dwave, dt, batch = _petal('galaxy', snr=snr)
    nt = len(targets)
    print(f'case: galaxy petal f16r vs f64 at per-pixel S/N {snr}, {nt} x {dt.cache.nz}', flush=True)
    c64, _ = _scan(dt, batch, nt, 'f64')
    cm, _ = _scan(dt, batch, nt, 'f16r')
    fwf = (batch.wf * batch.f).sum(dim=1).cpu().numpy()
    err = np.abs(cm - c64)
    agree = _agree(cm, c64)
    print(f'S/N {snr}: max |dchi2|/fWf {(err / fwf[:, None]).max():.2e}, '
          f'max rel {(err / c64).max():.2e}, best-z agreement {agree:.4f}')
    assert agree >= BESTZ_MIN
    assert np.all(err <= 1e-6 * fwf[:, None])
    ok = fwf[:, None] <= 10 * c64
    assert np.all(err[ok] <= 1e-5 * c64[ok])


def test_qso_petal_full_grid(gpu):
    """Config 3: 500 QSO x 3795 z with the fused IGM.  f16r <= 1e-5; the
    fp32-MFMA and split-fp16 forms stated at <= 2e-5 (their measured tail is
    1.1-1.3e-5 here); FP64 <= 1e-9 of the oracle on 2 targets."""
    tm, targets, dwave, dt, batch = _petal('qso')
    nt = len(targets)
    print(f'case: QSO petal, {nt} x {dt.cache.nz}', flush=True)
    c64, _ = _scan(dt, batch, nt, 'f64')
    tdata = orc.template_cache(tm, dwave, lines=LYA)
    rchi2, _, _ = orc.calc_zchi2(targets[:2], tdata, tm.nbasis)
    assert np.max(np.abs(c64[:2] - rchi2) / rchi2) <= 1e-9
    assert np.array_equal(c64[:2].argmin(axis=1), rchi2.argmin(axis=1))
    for precision, tol in (('f16r', 1e-5), ('f32', 2e-5), ('f16x3', 2e-5)):
        cm, _ = _scan(dt, batch, nt, precision)
        rel = np.abs(cm - c64) / c64
        agree = _agree(cm, c64)
        print(f'QSO {precision}: max rel {rel.max():.2e}, best-z agreement {agree:.4f}')
        assert agree >= BESTZ_MIN
        assert rel.max() <= tol


def test_mixed_batch_graph_5000(gpu):
    """Config 4: 5000 masked GALAXY/QSO/STAR targets, every template in one
    captured graph with autotuned tiles; replays with eager work between them
    equal the eager scan bit for bit; f16r within 1e-5 of the FP64 graph with
    the same best redshift per template and the same best template."""
    import torch
    from redrock_amd import synth, zscan
    from redrock_amd.templates import DistTemplate
    tms = (synth.galaxy_template(), synth.qso_template(), synth.star_template())
    targets, types, _ = synth.mixed_targets(5000, seed=1, templates=tms)
    dwave = {s.wavehash: s.wave for s in targets[0].spectra}
    dts = [DistTemplate(t, dwave) for t in tms]
    batch = zscan.TargetBatch(targets)
    caches = [(dt.cache, zscan._oiisum(dt.template)) for dt in dts]
    print('case: mixed batch 5000, f64 graph', flush=True)
    g64 = zscan.ScanGraph(caches, batch, 'f64').capture()
    g64.replay()
    r64 = g64.results()
    del g64
    print('case: mixed batch 5000, f16r graph (autotuned)', flush=True)
    gm = zscan.ScanGraph(caches, batch, 'f16r', tune=True)
    gm.run()
    eager = gm.results()
    gm.capture()
    for rep in range(3):
        gm.replay()
        torch.cuda.synchronize()
        got = gm.results()
        for (a, _, _), (b, _, _) in zip(got, eager):
            assert np.array_equal(a, b), f'replay {rep}'
        _ = torch.ones(1 << 20, device='cuda').sum()      # eager work between replays
    best_m, best_d = [], []
    for tm, (cm, _, pm), (cd, _, pd) in zip(tms, eager, r64):
        rel = np.abs(cm - cd) / cd
        agree = _agree(cm, cd)
        print(f'{tm.template_type}: f16r max rel {rel.max():.2e}, best-z agreement {agree:.4f}')
        assert agree >= BESTZ_MIN
        assert rel.max() <= 1e-5
        best_m.append(cm.min(axis=1))
        best_d.append(cd.min(axis=1))
    same_type = np.mean(np.argmin(best_m, axis=0) == np.argmin(best_d, axis=0))
    print('best template agreement', same_type)
    assert same_type >= BESTZ_MIN
