"""Batched scan of mixed GALAXY / QSO / STAR targets with masked pixel runs
over three template classes (uneven nbasis and redshift grids) captured in one
HIP graph (zscan.ScanGraph)."""
import numpy as np
import pytest

from oracle import rr_oracle as orc

pytestmark = pytest.mark.gpu

QSO_LINES = [(1215.67, 0.0036, 3.46)]


def _setup(n=12):
    from redrock_amd import synth
    from redrock_amd.templates import Template, DistTemplate
    g, q, s = synth.galaxy_template(), synth.qso_template(), synth.star_template()
    targets, types, zt = synth.mixed_targets(n, seed=5, mask_frac=(0.1, 0.3), templates=(g, q, s))
    # z subsets keep the oracle cross-check short; the GPU side is the same at full size
    tg = Template(spectype='GALAXY', redshifts=g.redshifts[::8], wave=g.wave, flux=g.flux,
                  igm_model=g.igm_model)
    tq = Template(spectype='QSO', redshifts=q.redshifts[::12], wave=q.wave, flux=q.flux,
                  igm_model=q.igm_model)
    ts = Template(spectype='STAR', subtype='K', redshifts=s.redshifts[::4], wave=s.wave,
                  flux=s.flux, igm_model='None')
    dwave = {sp.wavehash: sp.wave for sp in targets[0].spectra}
    dts = [DistTemplate(t, dwave) for t in (tg, tq, ts)]
    return targets, types, dts, dwave


@pytest.mark.parametrize('precision', ['f64', 'f16x3'])
def test_graph_replays_match_eager(gpu, precision):
    import torch
    from redrock_amd import zscan
    targets, types, dts, dwave = _setup()
    print(f'case: ScanGraph {precision}, {len(targets)} mixed masked targets x 3 templates', flush=True)
    batch = zscan.TargetBatch(targets)
    sg = zscan.ScanGraph([(dt.cache, zscan._oiisum(dt.template)) for dt in dts], batch, precision)
    eager = [zscan.scan_targets(dt.cache, targets, precision, zscan._oiisum(dt.template))
             for dt in dts]
    sg.capture()
    for rep in range(3):
        for p in sg.plans:                       # clobber the outputs between replays
            p.zchi2.fill_(-1.0)
            p.zcoeff.fill_(-1.0)
        sg.replay()
        torch.cuda.synchronize()
        # eager work between replays on the same device
        other = zscan.scan_targets(dts[2].cache, targets[:3], precision)
        for (c2, cf, pen), (e2, ef, ep) in zip(sg.results(), eager):
            assert np.array_equal(c2, e2), rep
            assert np.array_equal(cf, ef), rep
            assert np.array_equal(pen, ep), rep
        assert np.array_equal(other[0], eager[2][0][:3])


def test_graph_f64_matches_oracle(gpu):
    from redrock_amd import zscan
    targets, types, dts, dwave = _setup(6)
    batch = zscan.TargetBatch(targets)
    sg = zscan.ScanGraph([(dt.cache, zscan._oiisum(dt.template)) for dt in dts], batch, 'f64')
    sg.capture().replay()
    for dt, (c2, cf, pen) in zip(dts, sg.results()):
        tm = dt.template
        lines = QSO_LINES if tm.template_type == 'QSO' else None
        tdata = orc.template_cache(tm, dwave, lines=lines)
        rc2, rcf, rpen = orc.calc_zchi2(targets, tdata, tm.nbasis, getattr(tm, 'OIItemplate', None))
        assert np.max(np.abs(c2 - rc2) / np.abs(rc2)) <= 1e-9, tm.full_type
        assert np.array_equal(c2.argmin(axis=1), rc2.argmin(axis=1)), tm.full_type
        assert np.allclose(pen, rpen, rtol=1e-7, atol=1e-9 * max(np.abs(rpen).max(), 1.0))


@pytest.mark.parametrize('precision', ['f64', 'f16x3', 'f16', 'f32'])
def test_batch_invariance(gpu, precision):
    """A target's chi2 does not depend on which targets share its launch, their
    order or the tile shape: every (target, z) sum runs in a fixed order in its
    own lane or MFMA column (the GPU analogue of the reference's
    multi-process bit-identity test, test_zscan.py:113-140)."""
    from redrock_amd import zscan, autotune
    targets, types, dts, dwave = _setup(n=18)
    print(f'case: batch invariance {precision}, 18 targets', flush=True)
    for dt in dts[:2]:
        oii = zscan._oiisum(dt.template)
        full = zscan.scan_targets(dt.cache, targets, precision, oii)
        perm = np.random.default_rng(1).permutation(len(targets))
        shuffled = zscan.scan_targets(dt.cache, [targets[i] for i in perm], precision, oii)
        alone = [zscan.scan_targets(dt.cache, [targets[i]], precision, oii) for i in (0, 7, 17)]
        for k in range(3):
            assert np.array_equal(shuffled[k], full[k][perm])
        for i, one in zip((0, 7, 17), alone):
            for k in range(3):
                assert np.array_equal(one[k][0], full[k][i])
        if precision in ('f16x3', 'f16'):
            # every tile configuration of the kernel gives the same doubles
            batch = zscan.TargetBatch(targets)
            nt, nz = len(targets), dt.cache.nz
            ref = None
            for tile in range(len(autotune.tile_table(precision))):
                plan = zscan.ScanPlan(dt.cache, batch, np.arange(nt), np.zeros(nt), np.full(nt, nz),
                                      precision, oiisum=oii)
                plan.desc.tile = tile
                plan.run()
                c2 = plan.host()[0]
                if ref is None:
                    ref = c2
                assert np.array_equal(c2, ref), f'tile {tile}'
