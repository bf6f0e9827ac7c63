"""zfind assembly (redrock_amd.zfind) on the host: delta chi2, sorting, priors,
and the per-target result table against the oracle's restatement."""
import os

import numpy as np
import pytest

from oracle import rr_oracle as orc
from redrock_amd import zfind as zf
from redrock_amd.zwarning import ZWarningMask as ZW


def test_calc_deltachi2_reference_vectors():
    # test_zfind.py:141-170 and test_zscan.py:101-111
    chi2 = np.array([1.0, 2.0, 20.0, 40.0])
    d, s = zf.calc_deltachi2(chi2, np.array([3.0, 3.1, 3.2, 3.3]), np.array([0, 0, 0, 0]))
    assert np.all(d == [1, 18, 20, 0.0]) and np.all(s == [True, True, False, False])
    d, s = zf.calc_deltachi2(chi2, np.array([3.0, 3.0001, 4.0, 4.0001]), np.array([0, 0, 0, 0]))
    assert np.all(d == [19, 18, 0.0, 0.0]) and not np.any(s)
    d, s = zf.calc_deltachi2(chi2, np.array([3.0, 3.1, 3.2, 3.3]), np.array([0, 32, 0, 0]))
    assert np.all(d == [19, 18, 20, 0.0]) and np.all(s == [False, True, False, False])
    chi2 = np.array([1.0, 2.0, 4.0, 8.0])
    d, _ = zf.calc_deltachi2(chi2, np.array([3.0, 3.1, 3.2, 3.3]), np.zeros(4, int))
    assert np.all(d == [1, 2, 4, 0.0])
    d, _ = zf.calc_deltachi2(chi2, np.array([3.0, 3.0, 4.0, 4.0]), np.zeros(4, int))
    assert np.all(d == [3, 2, 0.0, 0.0])


@pytest.mark.parametrize('seed', range(20))
def test_calc_deltachi2_matches_oracle(seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(1, 12))
    chi2 = np.sort(rng.uniform(0, 40, n)).round(rng.integers(0, 3))   # ties included
    z = rng.choice([0.5, 0.5001, 0.51, 1.2, 1.2003, 2.0], n) + rng.normal(0, 1e-5, n) * rng.integers(0, 2)
    zwarn = rng.choice([0, 0, 0, 32, 1024, 256, 4], n)
    dv = rng.choice([100.0, 1000.0], n)
    d1, s1 = zf.calc_deltachi2(chi2, z, zwarn, dvlimit=dv)
    d0, s0 = orc.calc_deltachi2(chi2, z, zwarn, dvlimit=dv)
    assert np.array_equal(d1, d0) and np.array_equal(s1, s0)


def test_sort_zfit_dict_reference_vectors():
    # test_zfind.py:86-139
    cases = [([0, 0, 0], [2, 1, 0]), ([0, 0, 32], [1, 0, 2]), ([0, 0, 1024], [1, 0, 2]),
             ([0, 0, 256], [2, 1, 0]), ([0, 256, 288], [1, 0, 2])]
    for zwarn, rank in cases:
        z = dict(id=np.array([1, 2, 3]), zwarn=np.array(zwarn), chi2=np.array([30, 20, 10]),
                 rank=np.array(rank))
        zf.sort_zfit_dict(z)
        assert list(z['rank']) == [0, 1, 2]
        assert list(z.keys()) == ['id', 'zwarn', 'chi2', 'rank']


def test_sort_dict_by_cols_order():
    # test_zfind.py:180-206: zwarn first, then chi2
    d = dict(chi2=np.array([10.0, 3.0, 1.0, 5.0, 20.0]), zwarn=np.array([0, 4, 0, 1, 0]),
             z=np.array([2.0, 5.0, 1.0, 4.0, 3.0]))
    zf.sort_dict_by_cols(d, ('zwarn', 'chi2'))
    assert list(d['z']) == [1.0, 2.0, 3.0, 4.0, 5.0]
    d = dict(chi2=np.array([10.0, 3.0, 1.0, 5.0, 20.0]), zwarn=np.array([0, 4, 0, 1, 0]))
    with pytest.raises(KeyError):
        zf.sort_dict_by_cols(d, ('zwarn', 'nope'))


def test_priors_closed_forms():
    p = zf.Priors([7, 8], [1.0, 2.0], [0.1, 0.5], 'gaussian')
    z = np.array([0.9, 1.0, 1.1])
    assert np.allclose(p.eval(7, z), [1.0, 0.0, 1.0])
    assert p.eval(99, z) == 0.
    p = zf.Priors([7], [1.0], [0.1], 'lorentzien')
    assert np.allclose(p.eval(7, z), [2 * np.log(2), 0.0, 2 * np.log(2)])
    p = zf.Priors([7], [1.0], [0.25], 'tophat')
    z = np.linspace(0.7, 1.3, 13)             # step 0.05: |z - 1| < 0.125 -> 5 points
    v = p.eval(7, z)
    inside = np.abs(z - 1.0) < 0.125
    assert np.all(v[inside] == 0.0)
    assert np.isinf(v[3]) and np.isinf(v[9])
    assert np.all(np.isnan(np.delete(v, [3, 4, 5, 6, 7, 8, 9])))
    # tophat touching the grid edge: the edge point itself becomes inf
    v = p.eval(7, np.linspace(0.95, 1.5, 12))
    assert np.isinf(v[0])
    with pytest.raises(ValueError):
        zf.Priors([1], [1.0], [1.0], 'boxcar')


class _T(object):
    def __init__(self, nbasis):
        self.nbasis = nbasis


class _DT(object):
    def __init__(self, nbasis):
        self.template = _T(nbasis)


def _fake_zfit(rng, nmin, nb, npix, badbits=(0,)):
    z = np.sort(rng.uniform(0.1, 2.0, nmin))[:, None]
    return dict(z=z, zerr=rng.uniform(1e-5, 1e-4, (nmin, 1)),
                zwarn=rng.choice(badbits, (nmin, 1)).astype(np.int64),
                chi2=rng.uniform(7000, 7100, (nmin, 1)),
                zz=rng.uniform(0, 2, (nmin, 15)), zzchi2=rng.uniform(7000, 8000, (nmin, 15)),
                coeff=rng.normal(size=(nmin, nb)),
                fitmethod=np.full((nmin, 1), 'PCA'),
                npixels=np.full((nmin, 1), npix, dtype=np.int64))


def _fake_results(seed):
    rng = np.random.default_rng(seed)
    types = [('GALAXY', 10), ('QSO', 4), ('STAR:::K', 5), ('STAR:::M', 5)]
    res = dict()
    tids = [11, 12, 13, 14]
    for i, tid in enumerate(tids):
        res[tid] = dict()
        npix = [7958, 0, 30, 7958][i]           # NODATA, LITTLE_COVERAGE cases
        for ft, nb in types:
            nmin = int(rng.integers(0 if i == 3 and ft == 'QSO' else 1, 4))
            res[tid][ft] = dict(zfit=_fake_zfit(rng, nmin, nb, npix, badbits=(0, 0, 32, 1024, 256)))
    return res, tids, [_DT(nb) for _, nb in types]


@pytest.mark.parametrize('seed', range(6))
def test_assemble_matches_oracle(seed):
    res, tids, templates = _fake_results(seed)
    import copy
    want = orc.zfind_assemble(copy.deepcopy(res), tids, maxcoeff=10, nminima=3)
    got = zf._assemble(copy.deepcopy(res), tids, templates, nminima=3)
    assert list(got.keys())[0] == 'targetid'
    assert set(got.keys()) == set(want.keys())
    for k in want:
        assert got[k].shape == want[k].shape, k
        if got[k].dtype.kind in 'fc':
            assert np.array_equal(got[k], want[k]), k
        else:
            assert np.array_equal(got[k].astype(str), want[k].astype(str)), k


@pytest.mark.parametrize('seed', range(12))
def test_assemble_table_equals_per_target_assembly(seed):
    """The all-targets-at-once assembly (zfind without archetypes) gives the
    per-target assembly's table: same columns in the same order, dtypes, rows
    and row order, also with exact chi2 ties (HUGE_CHI2 rows) across templates
    and candidates, for every nminima, and with one target holding its
    templates in another order (the rank of a template is per target)."""
    import copy
    res, tids, templates = _fake_results(seed)
    rng = np.random.default_rng(100 + seed)
    if seed % 3 == 2:
        res[tids[1]] = dict(reversed(list(res[tids[1]].items())))
    if seed % 2:
        for tid in tids:
            for ft in res[tid]:
                c = res[tid][ft]['zfit']['chi2']
                c[rng.random(c.shape) < 0.4] = 9999999.0
                c[rng.random(c.shape) < 0.2] = 7050.0
    for nmin in (1, 2, 3):
        a = zf._assemble(copy.deepcopy(res), tids, templates, nminima=nmin)
        b = zf._assemble_table(copy.deepcopy(res), tids, templates, nminima=nmin)
        assert list(a.keys()) == list(b.keys())
        for k in a:
            assert a[k].dtype == b[k].dtype, k
            assert a[k].shape == b[k].shape, k
            assert np.array_equal(a[k], b[k]), k


@pytest.mark.parametrize('seed', range(6))
def test_assemble_table_from_columns(seed):
    """zfind hands the fine fits to the table as flat columns per template
    (fitz.fitz_columns: ids, per-target counts, columns target by target);
    the table equals the per-target assembly of the same fits."""
    import copy
    res, tids, templates = _fake_results(seed)
    rng = np.random.default_rng(200 + seed)
    if seed % 2:
        for tid in tids:
            for ft in res[tid]:
                c = res[tid][ft]['zfit']['chi2']
                c[rng.random(c.shape) < 0.4] = 9999999.0
    for nmin in (1, 3):
        want = zf._assemble(copy.deepcopy(res), tids, templates, nminima=nmin)
        blocks = dict()
        cols_res = copy.deepcopy(res)
        for ft in res[tids[0]]:
            per = [cols_res[tid][ft].pop('zfit') for tid in tids]
            nacc = np.array([len(z['chi2']) for z in per], dtype=np.int64)
            cols = {k: np.concatenate([z[k] for z in per]) for k in per[0]}
            cols = {k: (v.reshape(-1) if k != 'coeff' and v.ndim == 2 and v.shape[1] == 1 else v)
                    for k, v in cols.items()}
            blocks[ft] = (list(tids), nacc, cols)
        got = zf._assemble_table(cols_res, tids, templates, nminima=nmin, blocks=blocks)
        assert list(got.keys()) == list(want.keys())
        for k in want:
            assert got[k].dtype == want[k].dtype, k
            assert np.array_equal(got[k], want[k]), k


def test_assemble_known_answer():
    """Hand-computed: nminima trim per spectype, bad fits last, flags."""
    def zfit(zs, chi2s, zwarn, nb, npix):
        n = len(zs)
        return dict(z=np.array(zs, float)[:, None], zerr=np.full((n, 1), 1e-4),
                    zwarn=np.array(zwarn, np.int64)[:, None], chi2=np.array(chi2s, float)[:, None],
                    zz=np.zeros((n, 15)), zzchi2=np.zeros((n, 15)), coeff=np.ones((n, nb)),
                    fitmethod=np.full((n, 1), 'PCA'), npixels=np.full((n, 1), npix, np.int64))
    def fresh():
        return {5: {'GALAXY': dict(zfit=zfit([0.5, 1.0], [100.0, 105.0], [0, 0], 3, 100)),
                    'STAR:::K': dict(zfit=zfit([0.0, 0.001], [90.0, 200.0], [0, 0], 2, 100)),
                    'STAR:::M': dict(zfit=zfit([0.0002, 0.0005], [95.0, 50.0], [0, 32], 2, 100))}}
    res = fresh()
    got = zf._assemble(res, [5], [_DT(3), _DT(2), _DT(2)], nminima=2)
    # STAR keeps its 2 best good fits (90, 95); the Z_FITLIMIT one (50) sorts
    # after every good fit and is trimmed; GALAXY keeps both
    assert list(got['chi2']) == [90.0, 95.0, 100.0, 105.0]
    assert list(got['spectype']) == ['STAR', 'STAR', 'GALAXY', 'GALAXY']
    assert list(got['subtype']) == ['K', 'M', '', '']
    assert list(got['znum']) == [0, 1, 2, 3]
    # STAR rows: 0.0 vs 0.0002 is 60 km/s apart (< 100): same solution
    assert got['deltachi2'][0] == 10.0          # next good fit far enough: 100 (GALAXY z=0.5)
    assert got['deltachi2'][2] == 5.0 and got['deltachi2'][3] == 0.0
    # |100 - 105| < 9 and 1000 km/s apart: SMALL_DELTA_CHI2 on both GALAXY rows
    assert got['zwarn'][2] & ZW.SMALL_DELTA_CHI2 and got['zwarn'][3] & ZW.SMALL_DELTA_CHI2
    assert not (got['zwarn'][0] & ZW.SMALL_DELTA_CHI2)
    assert got['coeff'].shape == (4, 3)         # padded to the widest template
    assert np.array_equal(got['ncoeff'], [2, 2, 3, 3])
    assert got['zwarn'][1] & ZW.SMALL_DELTA_CHI2    # 95 vs GALAXY 100, far apart
    want = orc.zfind_assemble(fresh(), [5], maxcoeff=3, nminima=2)
    for k in want:
        assert np.array_equal(np.asarray(got[k]).astype(str), np.asarray(want[k]).astype(str)), k


def test_assemble_nodata_and_coverage_bits():
    res, tids, templates = _fake_results(1)
    got = zf._assemble(res, tids, templates, nminima=3)
    t12 = got['targetid'] == 12
    assert np.all(got['zwarn'][t12] & ZW.NODATA)
    t13 = got['targetid'] == 13
    assert np.all(got['zwarn'][t13] & ZW.LITTLE_COVERAGE)
    t11 = got['targetid'] == 11
    assert not np.any(got['zwarn'][t11] & (ZW.NODATA | ZW.LITTLE_COVERAGE))


def test_write_zbest_roundtrip(tmp_path):
    """Best rows in the REDSHIFTS layout of the reference's output file."""
    from redrock_amd import fits
    res, tids, templates = _fake_results(2)
    got = zf._assemble(res, tids, templates, nminima=3)
    path = str(tmp_path / 'zbest.fits')
    zf.write_zbest(path, got, header={'RRVER': 'redrock_amd'})
    t = fits.read(path)['REDSHIFTS'].data
    ref = fits.read(os.path.join(os.path.dirname(__file__), 'golden', 'redrock-test.fits'))['REDSHIFTS'].data
    assert list(t.keys()) == list(ref.keys())          # same columns, same order
    best = got['znum'] == 0
    assert np.array_equal(t['TARGETID'], got['targetid'][best])
    assert sorted(t['TARGETID']) == sorted(set(got['targetid']))
    for k in ('Z', 'ZERR', 'CHI2', 'DELTACHI2', 'COEFF', 'ZWARN', 'NPIXELS', 'NCOEFF'):
        assert np.array_equal(t[k], np.asarray(got[k.lower()])[best]), k
    assert list(t['SPECTYPE']) == list(got['spectype'][best])


def test_priors_from_fits(tmp_path):
    """PRIORS HDU with TARGETID, Z, SIGMA, FUNCTION (priors.py:26-34)."""
    from redrock_amd import fits
    path = str(tmp_path / 'priors.fits')
    fits.write(path, [fits.table_hdu({'TARGETID': np.array([10, 11]), 'Z': np.array([0.5, 1.0]),
                                      'SIGMA': np.array([0.1, 0.2]),
                                      'FUNCTION': np.array(['lorentzien', 'lorentzien'])}, 'PRIORS')])
    p = zf.Priors(path)
    z = np.array([0.4, 0.5, 0.6])
    assert np.allclose(p.eval(10, z), [2 * np.log(2), 0.0, 2 * np.log(2)])
    assert p.eval(12, z) == 0.


def _vertex_exact(x, y):
    """Vertex, curvature and value of the parabola through 3 points in exact
    rational arithmetic (divided differences)."""
    from fractions import Fraction as Fr
    X = [Fr(float(v)) for v in x]
    Y = [Fr(float(v)) for v in y]
    d1 = (Y[1] - Y[0]) / (X[1] - X[0])
    d2 = (Y[2] - Y[1]) / (X[2] - X[1])
    a = (d2 - d1) / (X[2] - X[0])
    b = d1 - a * (X[0] + X[1])
    c = Y[0] - a * X[0] ** 2 - b * X[0]
    return float(-b / (2 * a)), float(a), float(c - b * b / (4 * a))


def test_minfit_batch_is_the_exact_parabola():
    """minfit on 3 points (fitz.py:70-111) is the interpolating parabola: the
    batched closed form against exact rational arithmetic on fine-grid-like
    triples (dz ~ 1e-4, chi2 ~ 1e4), and the reference's BAD_MINFIT bits."""
    from redrock_amd.fitz import minfit_batch, minfit
    rng = np.random.default_rng(3)
    K = 300
    zc = rng.uniform(0.05, 3.5, K)
    h = rng.choice([2e-5, 1e-4, 4e-4], K) * (1 + zc)
    x = zc[:, None] + h[:, None] * np.array([-1.0, 0.0, 1.0])
    curv = rng.uniform(1e-3, 1e2, K) / h ** 2
    xt = zc + rng.uniform(-1.5, 1.5, K) * h          # some vertices outside the points
    y = rng.uniform(5e3, 5e4, K)[:, None] + curv[:, None] * (x - xt[:, None]) ** 2
    y[::7] *= -1                                     # negative curvature and value
    x0, xerr, y0, zw = minfit_batch(x, y)
    for k in range(K):
        ex0, ea, ey0 = _vertex_exact(x[k], y[k])
        assert abs(x0[k] - ex0) <= 4e-16 * (1 + abs(ex0)), k
        assert abs(xerr[k] - 1 / np.sqrt(abs(ea))) <= 1e-9 * xerr[k], k
        assert abs(y0[k] - ey0) <= 1e-12 * abs(ey0), k
        bad = (ex0 <= x[k, 0]) or (x[k, 2] <= ex0) or (ey0 <= 0) or (ea < 0)
        assert bool(zw[k] & ZW.BAD_MINFIT) == bad, k
        assert minfit(x[k], y[k]) == (x0[k], xerr[k], y0[k], zw[k])
        ox0, oxerr, oy0, ozw = orc.minfit(x[k], y[k])     # the oracle's restatement
        assert abs(ox0 - ex0) <= 4e-16 * (1 + abs(ex0)), k
        assert abs(oxerr - 1 / np.sqrt(abs(ea))) <= 1e-9 * oxerr, k
        assert abs(oy0 - ey0) <= 1e-12 * abs(ey0), k
        assert bool(ozw & ZW.BAD_MINFIT) == bad, k


def test_minfit_degenerate_rows_take_the_reference_path():
    """Rows the closed form does not cover give exactly what np.polyfit-based
    minfit (fitz.py:70-111) gives: a straight line (a == 0), NaN values,
    repeated abscissae; fewer than 3 points is BAD_MINFIT."""
    from redrock_amd.fitz import minfit_batch, _minfit_polyfit, minfit
    x = np.array([[0.5, 0.6, 0.7], [0.5, 0.6, 0.7], [0.5, 0.5, 0.7]])
    y = np.array([[1.0, 2.0, 3.0], [1.0, np.nan, 3.0], [3.0, 1.0, 2.0]])
    import warnings
    with warnings.catch_warnings():
        warnings.simplefilter('ignore')
        got = minfit_batch(x, y)
        for k in range(3):
            want = _minfit_polyfit(x[k], y[k])
            assert want[3] & ZW.BAD_MINFIT
            for g, w in zip((got[0][k], got[1][k], got[2][k], got[3][k]), want):
                assert g == w or (np.isnan(g) and np.isnan(w))
    assert minfit([0.1, 0.2], [1.0, 2.0]) == (-1, -1, -1, ZW.BAD_MINFIT)


def test_candidate_stacking_equals_vstack():
    """fitz stacks each result field over the candidates (fitz.py:359-395
    uses np.vstack): same shape, dtype and values for scalars, strings,
    equal-length 1-D arrays, mixed int / float and 2-D rows."""
    from redrock_amd.fitz import _vstack
    rng = np.random.default_rng(0)
    cases = [[1.5, 2.5, np.float64(3)], [np.int64(3), 1, 0], ['PCA', 'PCA'],
             [rng.normal(size=15) for _ in range(3)], [rng.normal(size=10)], [np.nan],
             [rng.normal(size=(1, 4)), rng.normal(size=(1, 4))], [1, 2.5],
             [np.int32(2), np.int32(5)], [rng.normal(size=4).astype(np.float32), rng.normal(size=4)]]
    for c in cases:
        got, want = _vstack(c), np.vstack(c)
        assert got.dtype == want.dtype and got.shape == want.shape, c
        assert np.array_equal(got, want, equal_nan=got.dtype.kind == 'f'), c
