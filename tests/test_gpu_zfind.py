"""zfind end to end on the GPU (scan + batched fitz + assembly) against the
oracle pipeline (float64 scan, fitz restatement, assembly restatement)."""
import numpy as np
import pytest

from oracle import rr_oracle as orc
from redrock_amd.zwarning import ZWarningMask as ZW

pytestmark = pytest.mark.gpu

QSO_LINES = [(1215.67, 0.0036, 3.46)]


def _setup():
    from redrock_amd import synth
    from redrock_amd.templates import Template, DistTemplate
    from redrock_amd.targets import DistTargetsCopy
    g = synth.galaxy_template()
    q = synth.qso_template()
    s = synth.star_template()
    tg = Template(spectype='GALAXY', redshifts=g.redshifts[600:900], wave=g.wave, flux=g.flux,
                  igm_model=g.igm_model)
    tq = Template(spectype='QSO', redshifts=q.redshifts[300:700], wave=q.wave, flux=q.flux,
                  igm_model=q.igm_model)
    ts = Template(spectype='STAR', subtype='K', redshifts=s.redshifts, wave=s.wave, flux=s.flux,
                  igm_model='None')
    targets, _ = synth.make_targets(g, 3, seed=31, zrange=(0.38, 0.58), snr=0.7)
    dwave = {sp.wavehash: sp.wave for sp in targets[0].spectra}
    dts = [DistTemplate(t, dwave) for t in (tg, tq, ts)]
    return targets, dts, dwave, DistTargetsCopy(targets)


def _oracle_zfind(targets, dts, dwave, nminima, priors=None):
    res = {t.id: dict() for t in targets}
    for dt in dts:
        tm = dt.template
        lines = QSO_LINES if tm.template_type == 'QSO' else None
        tdata = orc.template_cache(tm, dwave, lines=lines)
        oii = getattr(tm, 'OIItemplate', None)
        chi2, coeff, pen = orc.calc_zchi2(targets, tdata, tm.nbasis, oii,
                                          algo=tm.solve_matrices_algorithm)
        for i, t in enumerate(targets):
            c = chi2[i] + (priors.eval(t.id, tm.redshifts) if priors is not None else 0.)
            res[t.id][tm.full_type] = dict(zfit=orc.fitz(c + pen[i], tm.redshifts, t, tm,
                                                         nminima=nminima, lines=lines))
    return orc.zfind_assemble(res, [t.id for t in targets],
                              maxcoeff=max(dt.template.nbasis for dt in dts), nminima=nminima)


@pytest.mark.parametrize('with_prior', [False, True])
def test_zfind_matches_oracle(gpu, with_prior):
    from redrock_amd import zfind as zf
    targets, dts, dwave, dist = _setup()
    priors = None
    if with_prior:
        priors = zf.Priors([t.id for t in targets], [0.45] * len(targets), [0.05] * len(targets))
    print(f'case: zfind galaxy+qso+star, 3 targets, prior={with_prior}', flush=True)
    allresults, got = zf.zfind(dist, dts, nminima=3, priors=priors, precision='f64')
    want = _oracle_zfind(targets, dts, dwave, 3, priors)
    assert list(got.keys())[0] == 'targetid'
    assert set(got.keys()) == set(want.keys())
    for k in ('targetid', 'spectype', 'subtype', 'znum', 'ncoeff', 'npixels', 'fitmethod', 'zwarn'):
        assert np.array_equal(got[k].astype(str), want[k].astype(str)), k
    assert np.allclose(got['z'], want['z'], rtol=1e-9, atol=1e-12)
    assert np.allclose(got['chi2'], want['chi2'], rtol=1e-9, atol=0)
    assert np.allclose(got['deltachi2'], want['deltachi2'], rtol=1e-6, atol=1e-6)
    assert np.allclose(got['zerr'], want['zerr'], rtol=1e-6, atol=0)
    # coefficients are refit at the parabola vertex zmin (fitz.py:264-290); on
    # a BAD_MINFIT row the vertex of a nearly flat parabola lies outside the
    # fine grid and moves with the last bits of chi2, so only the rows with a
    # well-defined vertex are compared at 1e-6 (row-norm relative)
    dc = np.linalg.norm(got['coeff'] - want['coeff'], axis=1) / np.linalg.norm(want['coeff'], axis=1)
    good = (want['zwarn'] & ZW.BAD_MINFIT) == 0
    w = int(np.argmax(np.where(good, dc, -1)))
    print('worst coeff row', int(dc.argmax()), dc.max(), 'zwarn', want['zwarn'][dc.argmax()],
          '; worst good row', w, dc[w])
    assert good.sum() >= len(targets)
    assert dc[good].max() <= 1e-6
    assert np.all(np.isfinite(got['coeff']))
    # the scan results stay available per (target, fulltype), without 'zfit'
    for t in targets:
        for dt in dts:
            r = allresults[t.id][dt.template.full_type]
            assert set(r.keys()) == {'redshifts', 'zchi2', 'penalty', 'zcoeff'}


def test_zfind_nmf_template_matches_oracle(gpu):
    """A PCA and an NMF galaxy template side by side (RRMETHOD = NMF: the scan
    and the fine fits solve nnls(M, y), fitz.py:246-253,290-292)."""
    from redrock_amd import synth, zfind as zf
    from redrock_amd.templates import Template, DistTemplate
    from redrock_amd.targets import DistTargetsCopy
    g = synth.galaxy_template()
    tg = Template(spectype='GALAXY', redshifts=g.redshifts[600:900], wave=g.wave, flux=g.flux,
                  igm_model=g.igm_model)
    tn = Template(spectype='GALAXY', subtype='NMF', redshifts=g.redshifts[600:900], wave=g.wave,
                  flux=np.abs(g.flux), method='NMF', igm_model=g.igm_model)
    targets, _ = synth.make_targets(g, 3, seed=31, zrange=(0.38, 0.58), snr=0.7)
    dwave = {sp.wavehash: sp.wave for sp in targets[0].spectra}
    dts = [DistTemplate(t, dwave) for t in (tg, tn)]
    print('case: zfind PCA + NMF galaxy templates, 3 targets', flush=True)
    allresults, got = zf.zfind(DistTargetsCopy(targets), dts, nminima=3, precision='f64')
    want = _oracle_zfind(targets, dts, dwave, 3)
    for k in ('targetid', 'spectype', 'subtype', 'znum', 'fitmethod', 'zwarn'):
        assert np.array_equal(got[k].astype(str), want[k].astype(str)), k
    assert np.allclose(got['z'], want['z'], rtol=1e-9, atol=1e-12)
    assert np.allclose(got['chi2'], want['chi2'], rtol=1e-9, atol=0)
    # the NMF template's scan coefficients are non-negative (the table keeps
    # the 3 best minima per spectype, here all from the better-fitting PCA one)
    for t in targets:
        assert np.all(allresults[t.id][tn.full_type]['zcoeff'] >= 0)


@pytest.mark.parametrize('precision', ['f16x3', 'f16r'])
def test_zfind_polish_skips_nmf_template(gpu, precision):
    """polish=True with a PCA and an NMF template: the NMF (bounded) scan runs
    without the FP64 polish, with a warning, and zfind records it; the PCA
    template is polished; the table equals the FP64 run's in z and spectype."""
    from redrock_amd import synth, zfind as zf
    from redrock_amd.templates import Template, DistTemplate
    from redrock_amd.targets import DistTargetsCopy
    g = synth.galaxy_template()
    tg = Template(spectype='GALAXY', redshifts=g.redshifts[600:900], wave=g.wave, flux=g.flux,
                  igm_model=g.igm_model)
    tn = Template(spectype='GALAXY', subtype='NMF', redshifts=g.redshifts[600:900], wave=g.wave,
                  flux=np.abs(g.flux), method='NMF', igm_model=g.igm_model)
    targets, _ = synth.make_targets(g, 3, seed=31, zrange=(0.38, 0.58), snr=0.7)
    dwave = {sp.wavehash: sp.wave for sp in targets[0].spectra}
    dts = [DistTemplate(t, dwave) for t in (tg, tn)]
    print(f'case: zfind PCA + NMF, {precision}, polish=True', flush=True)
    with pytest.warns(UserWarning, match='no FP64 polish'):
        allresults, got = zf.zfind(DistTargetsCopy(targets), dts, nminima=3, precision=precision,
                                   polish=True)
    assert set(zf.last_timings['unpolished']) == {tn.full_type}
    _, want = zf.zfind(DistTargetsCopy(targets), dts, nminima=3, precision='f64')
    assert zf.last_timings['unpolished'] == {}
    assert np.array_equal(got['spectype'].astype(str), want['spectype'].astype(str))
    assert np.allclose(got['z'], want['z'], rtol=1e-9, atol=1e-12)
    for t in targets:
        assert np.all(allresults[t.id][tn.full_type]['zcoeff'] >= 0)


def test_zfind_mixed_precision_best_z(gpu):
    """The split-fp16 scan feeds the same fits: identical best redshift per target."""
    from redrock_amd import zfind as zf
    targets, dts, dwave, dist = _setup()
    _, a = zf.zfind(dist, dts, nminima=3, precision='f64')
    _, b = zf.zfind(dist, dts, nminima=3, precision='f16x3')
    best_a = {t: a['z'][(a['targetid'] == t) & (a['znum'] == 0)][0] for t in np.unique(a['targetid'])}
    best_b = {t: b['z'][(b['targetid'] == t) & (b['znum'] == 0)][0] for t in np.unique(b['targetid'])}
    assert best_a.keys() == best_b.keys()
    for t in best_a:
        assert abs(best_a[t] - best_b[t]) < 1e-9 * (1 + best_a[t])


def test_zfind_resumes_from_details_file(gpu, tmp_path):
    """zfind -> write_zscan -> zfind(chi2_scan=file) (zfind.py:289-292): the
    stored float32 scan selects the same minima, and the fine fits, which run
    again in FP64, give the same table."""
    from redrock_amd import zfind as zf, results
    targets, dts, dwave, dist = _setup()
    print('case: zfind resume from details file, 3 targets', flush=True)
    allresults, a = zf.zfind(dist, dts, nminima=3, precision='f64')
    path = str(tmp_path / 'rrdetails.h5')
    results.write_zscan(path, allresults, a)
    saved = results.read_zscan_redrock(path)
    for t in targets:
        for dt in dts:
            ft = dt.template.full_type
            assert np.array_equal(saved[t.id][ft]['zchi2'],
                                  allresults[t.id][ft]['zchi2'].astype(np.float32))
    _, b = zf.zfind(dist, dts, nminima=3, chi2_scan=path)
    for k in ('targetid', 'spectype', 'subtype', 'znum', 'zwarn', 'npixels'):
        assert np.array_equal(a[k], b[k]), k
    assert np.allclose(a['z'], b['z'], rtol=1e-12, atol=0)
    assert np.allclose(a['chi2'], b['chi2'], rtol=1e-12, atol=0)
    # read_zfit stacks the targets in the order of the zbest table: zfind's order
    zfit = results.read_zfit(path)
    assert np.array_equal(zfit['z'], a['z']) and np.array_equal(zfit['targetid'], a['targetid'])
    # test_io.py:45-81 of the reference: same columns; z and targetid keep their
    # precision, the other columns equal after the writer's conversions
    zscan2, zfit2 = results.read_zscan(path)
    assert list(zfit2.keys()) == list(a.keys())
    assert zfit2['z'].dtype == a['z'].dtype and zfit2['targetid'].dtype == a['targetid'].dtype
    for cn in a:
        d2 = zfit2[cn]
        assert np.array_equal(np.asarray(a[cn]).astype(d2.dtype), d2), cn
    for tid in allresults:
        for ft in allresults[tid]:
            for key, d1 in allresults[tid][ft].items():
                if key != 'redshifts':
                    d1 = d1.astype(np.float32)
                assert np.array_equal(d1, zscan2[tid][ft][key]), (tid, ft, key)


def test_scan_templates_pipelined_under_memory_pressure(gpu, monkeypatch):
    """The pipelined scan of every template unpacks pending scans first when
    the next plan would not fit; the results are the same either way."""
    from redrock_amd import zscan
    targets, dts, dwave, dist = _setup()
    print('case: _scan_templates with and without early unpacking', flush=True)
    res_a, full_a = zscan._scan_templates(dist, dts, precision='f16r')
    monkeypatch.setattr(zscan, '_plan_bytes', lambda rows, nb: 1 << 62)
    res_b, full_b = zscan._scan_templates(dist, dts, precision='f16r')
    assert list(full_a) == list(full_b)
    for ft in full_a:
        assert np.array_equal(full_a[ft], full_b[ft])
        for t in targets:
            for k in ('zchi2', 'penalty', 'zcoeff'):
                assert np.array_equal(res_a[t.id][ft][k], res_b[t.id][ft][k])


def test_zfind_two_layouts_matches_oracle(gpu):
    """Targets of two wavelength layouts (three cameras, and two of them): one
    target batch per layout, scanned and refined per layout (the path beside
    the pipelined one-layout scan), against the oracle pipeline."""
    from redrock_amd import synth, zfind as zf
    from redrock_amd.targets import Target, DistTargetsCopy
    targets, dts, dwave, _ = _setup()
    more, _ = synth.make_targets(synth.galaxy_template(), 2, seed=32, zrange=(0.40, 0.55), snr=0.7)
    two = [Target(100 + i, t.spectra[:2]) for i, t in enumerate(more)]
    mixed = [targets[0], two[0], targets[1], two[1], targets[2]]
    print('case: zfind, 5 targets in two wavelength layouts', flush=True)
    _, got = zf.zfind(DistTargetsCopy(mixed), dts, nminima=3, precision='f64')
    want = _oracle_zfind(mixed, dts, dwave, 3)
    for k in ('targetid', 'spectype', 'subtype', 'znum', 'npixels', 'zwarn'):
        assert np.array_equal(got[k].astype(str), want[k].astype(str)), k
    assert np.allclose(got['z'], want['z'], rtol=1e-9, atol=1e-12)
    assert np.allclose(got['chi2'], want['chi2'], rtol=1e-9, atol=0)
