"""select_minima (the candidate loop of fitz.py:198-327 over all targets at
once) against the loop run target by target, as the reference writes it.

A synthetic refiner stands in for the GPU fine-grid fits: every (target,
coarse minimum) gets a deterministic fine grid, chi2 values and a parabola
vertex that falls inside or outside its grid, so the BAD_MINFIT and
Z_FITLIMIT branches, both velocity skips, and targets that outrun the
speculative candidates (further refinement rounds) all occur.
"""
import numpy as np
import pytest

from redrock_amd import constants
from redrock_amd.fitz import find_minima, get_dv, select_minima
from redrock_amd.zwarning import ZWarningMask as ZW

NZF = 15


def _fake(zs, t, imin, bad=None):
    """The refined candidate (target t, coarse minimum imin)."""
    rng = np.random.default_rng(1000003 * t + imin)
    zz = np.linspace(zs[max(imin - 1, 0)], zs[min(imin + 1, zs.size - 1)], NZF)
    half = 0.5 * (zz[-1] - zz[0])
    zmin = 0.5 * (zz[0] + zz[-1]) + rng.uniform(-1.4, 1.4) * half
    return dict(zz=zz, zzchi2=rng.uniform(10, 20, NZF), zmin=zmin, sigma=rng.uniform(1e-4, 1e-3),
                chi2min=rng.uniform(5, 30), zwarn=int(rng.integers(0, 2)) * ZW.BAD_MINFIT,
                coeff=rng.normal(size=3), zigm=zmin, err=(bad == (t, imin)))


class _Refine:
    def __init__(self, zs, bad=None):
        self.zs, self.bad, self.calls = zs, bad, 0

    def __call__(self, pt, pimin):
        self.calls += 1
        rs = [_fake(self.zs, int(t), int(i), self.bad) for t, i in zip(pt, pimin)]
        return {k: np.array([r[k] for r in rs]) for k in rs[0]} if rs else None


def _sequential(zchi2_row, zs, t, nminima, mvd, bad=None):
    """fitz.py:198-327 for one target, on the synthetic refiner."""
    results = []
    for imin in find_minima(zchi2_row):
        if len(results) == nminima:
            break
        zprev = np.array([r['z'] for r in results])
        if np.any(np.abs(get_dv(z=zs[imin], zref=zprev)) < mvd):
            continue
        r = _fake(zs, t, int(imin), bad)
        if r['err']:
            raise ValueError('edges must be within input x range')
        zz, zzchi2 = r['zz'], r['zzchi2']
        zmin, zwarn, chi2min = r['zmin'], r['zwarn'], r['chi2min']
        zbest = zmin
        if zbest < zz[0] or zbest > zz[-1]:
            zwarn |= ZW.BAD_MINFIT
            i = np.where(zzchi2 == np.min(zzchi2))[0][0]
            zbest = zz[i]
            chi2min = zzchi2[i]
        if zbest < zs[1] or zbest > zs[-2]:
            zwarn |= ZW.Z_FITLIMIT
        if zmin < zs[1] or zmin > zs[-2]:
            zwarn |= ZW.Z_FITLIMIT
        zprev = np.array([r['z'] for r in results])
        if np.any(np.abs(get_dv(z=zbest, zref=zprev)) < mvd):
            continue
        results.append(dict(z=zbest, zerr=r['sigma'], zwarn=zwarn, chi2=chi2min, imin=int(imin)))
    return results


def _problem(nt=40, nz=300, seed=0):
    rng = np.random.default_rng(seed)
    zs = np.linspace(0.1, 1.0, nz)            # adjacent points ~600 km/s apart: many skips
    zchi2 = rng.normal(size=(nt, nz)).cumsum(axis=1) * 0.3 + rng.normal(size=(nt, nz))
    zchi2[3] = 5.0                            # flat: every point a minimum (non-strict)
    zchi2[4, :] = np.arange(nz)               # one minimum, at the first redshift
    zchi2[5, :] = -np.arange(nz)              # one minimum, at the last redshift
    return zs, zchi2


@pytest.mark.parametrize('nminima,speculate', [(1, 0), (3, 0), (3, 2), (5, 1), (8, 2)])
def test_select_minima_equals_sequential(nminima, speculate):
    zs, zchi2 = _problem()
    mvd = constants.max_velo_diff
    ref = _Refine(zs)
    nacc, acc = select_minima(zchi2, zs, nminima, mvd, ref, speculate)
    st = acc['store']
    for t in range(zchi2.shape[0]):
        want = _sequential(zchi2[t], zs, t, nminima, mvd)
        assert nacc[t] == len(want), t
        for k, w in enumerate(want):
            assert acc['z'][t, k] == w['z']
            assert acc['zerr'][t, k] == w['zerr']
            assert acc['zwarn'][t, k] == w['zwarn']
            assert acc['chi2'][t, k] == w['chi2']
            row = acc['row'][t, k]
            assert np.array_equal(st['zz'][row], _fake(zs, t, w['imin'])['zz'])
    if speculate == 0 and nminima > 1:
        assert ref.calls > 1                  # targets waited for more candidates


def test_select_minima_flags_and_skips_occur():
    """The synthetic problem exercises every branch the loop has."""
    zs, zchi2 = _problem()
    nacc, acc = select_minima(zchi2, zs, 5, constants.max_velo_diff, _Refine(zs), 1)
    zw = acc['zwarn'][np.arange(5)[None, :] < nacc[:, None]]
    assert np.any(zw & ZW.BAD_MINFIT) and np.any(zw & ZW.Z_FITLIMIT)
    assert np.any(nacc < np.array([find_minima(r).size for r in zchi2]))   # skips
    assert nacc[4] == 1 and nacc[5] == 1


def test_select_minima_raises_where_the_reference_does():
    """A candidate whose refinement failed raises only if the loop reaches it."""
    zs, zchi2 = _problem(nt=6)
    mvd = constants.max_velo_diff
    t = 1
    reached = [r['imin'] for r in _sequential(zchi2[t], zs, t, 3, mvd)]
    with pytest.raises(ValueError):
        select_minima(zchi2, zs, 3, mvd, _Refine(zs, bad=(t, reached[0])), 2)
    # a refined but unreached candidate (beyond nminima) does not raise
    unreached = [int(i) for i in find_minima(zchi2[t]) if int(i) not in reached][-1]
    select_minima(zchi2, zs, 3, mvd, _Refine(zs, bad=(t, unreached)), 2)


def test_select_minima_no_targets():
    zs = np.linspace(0.1, 1.0, 10)
    nacc, acc = select_minima(np.zeros((0, 10)), zs, 3, constants.max_velo_diff, _Refine(zs), 2)
    assert nacc.size == 0 and acc['store'] is None


def test_columns_and_dicts_equal_stacked_results():
    """fitz_targets' per-target dicts (flat columns viewed per target) equal
    the reference's stacking of each target's accepted minima (fitz.py:359-395:
    sorted by chi2, (n, 1) columns, the empty-result dict)."""
    from types import SimpleNamespace
    from redrock_amd.fitz import _as_dicts, _columns, _stack
    zs, zchi2 = _problem(nt=12)
    zchi2[7] = 0.0 * zchi2[7] - np.arange(zs.size)          # one minimum at the last redshift
    nacc, acc = select_minima(zchi2, zs, 3, constants.max_velo_diff, _Refine(zs), 1)
    st = acc['store']
    tm = SimpleNamespace(nbasis=3, method='PCA')
    npix = list(range(100, 112))
    got = _as_dicts(*_columns(acc, nacc, st, tm, npix, NZF))
    for t in range(zchi2.shape[0]):
        rows = [dict(z=acc['z'][t, k], zerr=acc['zerr'][t, k], zwarn=int(acc['zwarn'][t, k]),
                     chi2=acc['chi2'][t, k], zz=st['zz'][acc['row'][t, k]],
                     zzchi2=st['zzchi2'][acc['row'][t, k]], coeff=st['coeff'][acc['row'][t, k]],
                     fitmethod='PCA') for k in range(nacc[t])]
        want = _stack(rows, None, npix[t])
        assert list(got[t].keys()) == list(want.keys())
        for k in want:
            assert got[t][k].shape == want[k].shape, (t, k)
            assert got[t][k].dtype == want[k].dtype, (t, k)
            assert np.array_equal(got[t][k], want[k]), (t, k)
    # a target without minima gets the reference's empty dict
    empty = _as_dicts(np.zeros(1, dtype=np.int64), {k: v[:0] for k, v in
                                                    _columns(acc, nacc, st, tm, npix, NZF)[1].items()})[0]
    assert empty['z'].shape == (0, 1) and empty['fitmethod'].dtype.kind == 'U'


@pytest.mark.parametrize('nz', [1, 2, 7, 300])
def test_minima_head_equals_find_minima(nz):
    """The first k candidates of every row (rr_minima_head, a host function
    of the library) equal find_minima's, order included: distinct values,
    ties (flat rows, repeated values: np.argsort's order), NaN and infinite
    values, k beyond a row's candidates and beyond the library's 64."""
    from redrock_amd.fitz import minima_head
    rng = np.random.default_rng(nz)
    x = rng.normal(size=(60, nz)).cumsum(axis=1)
    x[1] = 3.0                                        # all tied
    x[2] = np.round(x[2])                             # many ties
    x[3, ::3] = np.nan
    x[4, ::5] = np.inf
    x[5] = -x[5]
    x[6, ::2] = -np.inf
    x[7] = np.round(rng.normal(size=nz) * 2) * 0.5    # ties between separate minima
    if nz >= 40:                                      # minima 5, 1, 3, 2, 3 between peaks:
        x[8] = 10.0                                   # a tie at the cut of a 3-long head
        x[8, 4:40:8] = [5.0, 1.0, 3.0, 2.0, 3.0]
        x[9, 17:33] = x[9, 17:33].min() - 1.0         # a flat run holding the lowest values
    for k in (0, 1, 3, 4, 9, 70):
        head, ncand = minima_head(x, k)
        assert head.shape == (60, k)
        for t in range(x.shape[0]):
            want = find_minima(x[t])
            assert ncand[t] == want.size, (k, t)
            n = min(k, want.size)
            assert np.array_equal(head[t, :n], want[:k]), (k, t)
            assert np.all(head[t, n:] == -1), (k, t)
    h0, n0 = minima_head(np.zeros((0, nz)), 4)
    assert h0.shape == (0, 4) and n0.size == 0
