"""HDF5 details file (redrock_amd.hdf5, redrock_amd.results).

tests/golden/rrdetails-test.h5 and redrock-test.fits are the data files of the
reference's own tests (py/redrock/test/data): one rrdesi run over 3 DESI
targets with the GALAXY and QSO:::HIZ templates wrote both (its README).  They
pin the readers against each other, and the fits stored in the details file
are reference output for find_minima, minfit, get_dv, calc_deltachi2 and the
candidate ordering, checked here on the product's host code and on the oracle.
"""
import os

import numpy as np
import pytest

from redrock_amd import hdf5, results, fits, fitz, zfind as rzfind
from redrock_amd.zwarning import ZWarningMask as ZW
from redrock_amd import constants
from oracle import rr_oracle as orc

GOLDEN = os.path.join(os.path.dirname(__file__), 'golden')
DETAILS = os.path.join(GOLDEN, 'rrdetails-test.h5')


def _equal_trees(a, b, path=''):
    assert set(a.keys()) == set(b.keys()), path
    for k in a:
        if isinstance(a[k], dict):
            _equal_trees(a[k], b[k], path + '/' + k)
        else:
            assert a[k].dtype == b[k].dtype, path + '/' + k
            assert a[k].shape == b[k].shape and np.array_equal(a[k], b[k]), path + '/' + k


def test_reference_details_layout():
    f = hdf5.read(DETAILS)
    tids = f['targetids']
    assert tids.dtype == np.int64 and len(tids) == 3
    assert sorted(f['zscan'].keys()) == ['GALAXY', 'QSO:::HIZ']
    for ft, nb in (('GALAXY', 10), ('QSO:::HIZ', 4)):
        g = f['zscan'][ft]
        nz = len(g['redshifts'])
        assert g['zchi2'].shape == (3, nz) and g['penalty'].shape == (3, nz)
        assert g['zcoeff'].shape == (3, nz, nb)
        assert np.all(np.diff(g['redshifts']) > 0)
        assert np.all(g['zchi2'] > 0) and np.all(g['penalty'] >= 0)
    assert set(f['zfit'].keys()) == {str(t) for t in tids}
    rec = f['zfit'][str(tids[0])]['zfit']
    assert rec.dtype.names[:5] == ('targetid', 'z', 'zerr', 'zwarn', 'chi2')
    assert rec['zz'].shape == (6, 15) and rec['coeff'].shape == (6, 10)


def test_details_agree_with_redshift_table():
    """Both outputs of the reference run: the znum == 0 rows of the details
    file are the REDSHIFTS table, column by column."""
    zfit = results.read_zfit(DETAILS)
    t = fits.read(os.path.join(GOLDEN, 'redrock-test.fits'))['REDSHIFTS'].data
    best = zfit['znum'] == 0
    order = [list(zfit['targetid'][best]).index(x) for x in t['TARGETID']]
    for col in ('Z', 'ZERR', 'ZWARN', 'CHI2', 'COEFF', 'NPIXELS', 'SPECTYPE', 'SUBTYPE', 'NCOEFF',
                'DELTACHI2'):
        assert np.array_equal(t[col], zfit[col.lower()][best][order]), col


def test_read_zscan_api():
    zscan, zfit = results.read_zscan(DETAILS)
    tid = results.read_zscan_redrock(DETAILS)
    for t in zscan:
        assert sorted(zscan[t].keys()) == ['GALAXY', 'QSO:::HIZ']
        for ft in zscan[t]:
            z = zscan[t][ft]
            assert np.array_equal(z['zchi2'], tid[t][ft]['zchi2'])
            assert len(z['zfit']['z']) == 3 and np.all(z['zfit']['targetid'] == t)
            spectype = ft.split(':::')[0]
            assert np.all(z['zfit']['spectype'] == spectype)
    one = results.read_zscan(DETAILS, select_targetids=39633393167631849, nozfit=True)
    assert list(one.keys()) == [39633393167631849]
    up = results.read_zfit(DETAILS, select_targetids=[616094145466860255], upper=True)
    assert 'TARGETID' in up and np.all(up['TARGETID'] == 616094145466860255) and len(up['Z']) == 6


@pytest.mark.parametrize('impl', ['product', 'oracle'])
def test_reference_fits_replayed(impl):
    """Replays fitz's candidate choice on the stored scan (fitz.py:198-213) and
    its parabola (fitz.py:259-313), and zfind's deltachi2 (zfind.py:103-155,
    602-611) on the stored fits: every stored fine grid, z, zerr, chi2, the
    minfit zwarn bits, deltachi2 and SMALL_DELTA_CHI2 must come out again."""
    m = fitz if impl == 'product' else orc
    mz = rzfind if impl == 'product' else orc
    zscan, zfit = results.read_zscan(DETAILS)
    for tid in zscan:
        for ft, d in zscan[tid].items():
            rows = d['zfit']
            redshifts = d['redshifts']
            chi2 = d['zchi2'] + d['penalty']
            accepted, used = [], []
            for imin in m.find_minima(chi2):
                if len(accepted) == 3:
                    break
                if accepted and np.any(np.abs(m.get_dv(redshifts[imin], np.array(accepted)))
                                       < constants.max_velo_diff):
                    continue
                ilo, ihi = max(0, imin - 1), min(imin + 1, len(chi2) - 1)
                zz = np.linspace(redshifts[ilo], redshifts[ihi], 15)
                hit = [r for r in range(len(rows['z'])) if np.array_equal(rows['zz'][r], zz)]
                assert len(hit) == 1, (tid, ft, imin)
                used.append(hit[0])
                accepted.append(rows['z'][hit[0]])
            assert sorted(used) == list(range(len(rows['z'])))
            # np.polyfit on 3 points ~1e-4 apart near z ~ 1 has condition
            # number ~ (z / dz)^2 ~ 1e8: the curvature (zerr) differs between
            # LAPACK builds by up to ~2e-8 relative; vertex and value by far less.
            for r in range(len(rows['z'])):
                zz, zzchi2 = rows['zz'][r], rows['zzchi2'][r]
                i = min(max(np.argmin(zzchi2), 1), len(zz) - 2)
                zmin, sigma, chi2min, zwarn = m.minfit(zz[i - 1:i + 2], zzchi2[i - 1:i + 2])
                assert abs(rows['zerr'][r] - sigma) <= 3e-8 * sigma
                if zz[0] <= zmin <= zz[-1]:
                    zbest = zmin
                    assert abs(rows['z'][r] - zmin) <= 1e-11 * (1 + zmin)
                    assert abs(rows['chi2'][r] - chi2min) <= 1e-10 * chi2min
                else:
                    k = np.argmin(zzchi2)
                    zbest = zz[k]
                    zwarn |= ZW.BAD_MINFIT
                    assert rows['z'][r] == zz[k] and rows['chi2'][r] == zzchi2[k]
                for zt in (zbest, zmin):
                    if zt < redshifts[1] or zt > redshifts[-2]:
                        zwarn |= ZW.Z_FITLIMIT
                bits = ZW.BAD_MINFIT | ZW.Z_FITLIMIT
                assert (rows['zwarn'][r] & bits) == (zwarn & bits), (tid, ft, r)
    # deltachi2 over each target's candidates, in the stored (znum) order
    for tid in np.unique(zfit['targetid']):
        sel = np.where(zfit['targetid'] == tid)[0]
        sel = sel[np.argsort(zfit['znum'][sel])]
        zw = zfit['zwarn'][sel] & ~ZW.SMALL_DELTA_CHI2
        dv = np.where(zfit['spectype'][sel] == 'STAR', constants.max_velo_diff_star,
                      constants.max_velo_diff)
        dchi2, setz = mz.calc_deltachi2(zfit['chi2'][sel], zfit['z'][sel], zw, dvlimit=dv)
        assert np.array_equal(dchi2, zfit['deltachi2'][sel])
        assert np.array_equal(setz, (zfit['zwarn'][sel] & ZW.SMALL_DELTA_CHI2) != 0)
        # stored order = zfind's sort: bad fits last, then by chi2
        tz = {k: zfit[k][sel] for k in ('chi2', 'zwarn')}
        tz['i'] = np.arange(len(sel))
        mz.sort_zfit_dict(tz)
        assert np.array_equal(tz['i'], np.arange(len(sel)))


def test_writer_encodes_like_h5py(tmp_path):
    """Writing the reference file's contents back gives the same tree and the
    same dataset header messages (dataspace, datatype, fill value, layout
    class) as h5py wrote."""
    f = hdf5.read(DETAILS)
    out = str(tmp_path / 'copy.h5')
    hdf5.write(out, f)
    _equal_trees(f, hdf5.read(out))

    def messages(path):
        b = open(path, 'rb').read()
        r = hdf5._Reader(b)
        got = dict()

        def walk(addr, name):
            msgs = r.messages(addr)
            got[name] = [(t, b[p:p + (2 if t == 8 else n)], fl) for t, p, n, fl in msgs
                         if t not in (0, 0x11)]
            for t, p, n, fl in msgs:
                if t == 0x11:
                    for child, h in r._btree_entries(r.u(p, 8), r.heap(r.u(p + 8, 8))):
                        walk(h, name + '/' + child)
        walk(r.root_header, '')
        return got
    assert messages(DETAILS) == messages(out)


def test_hdf5_roundtrip_types(tmp_path):
    rng = np.random.default_rng(3)
    rec = np.zeros(4, dtype=[('a', '<i8'), ('b', '<f4', (3,)), ('s', 'S5'), ('flag', '?'),
                             ('m', '<f8', (2, 2)), ('k', 'i1')])
    rec['a'] = [1, -2, 2 ** 40, 0]
    rec['b'] = rng.normal(size=(4, 3))
    rec['s'] = [b'GAL', b'QSO', b'STAR', b'']
    rec['flag'] = [True, False, True, False]
    rec['m'] = rng.normal(size=(4, 2, 2))
    rec['k'] = [-1, 2, 3, -128]
    plain = np.zeros(3, dtype=[('x', '<f8'), ('y', '<i4')])
    plain['x'] = [0.5, 1.5, 2.5]
    big = {f'n{i:04d}': np.array([i], dtype=np.int32) for i in range(300)}   # 38 symbol nodes
    tree = {
        'f8': rng.normal(size=(3, 4, 5)), 'f4': rng.normal(size=7).astype(np.float32),
        'u2': np.arange(9, dtype=np.uint16), 'i8': np.array([-5, 7]), 'bool': np.array([True, False]),
        'scalar': np.array(2.5), 'empty': np.zeros((0, 3)), 'strings': np.array([b'a', b'bcd']),
        'be': np.arange(4, dtype='>f8'), 'rec': rec, 'plain': plain,
        'g': {'h': {'deep': np.arange(3)}, 'empty_group': dict(), 'big': big},
    }
    path = str(tmp_path / 't.h5')
    hdf5.write(path, tree)
    back = hdf5.read(path)
    want = dict(tree)
    want['be'] = tree['be'].astype('<f8')
    _equal_trees(want, back)
    assert back['g/big/n0299'][0] == 299 and len(back['g']['big']) == 300
    with pytest.raises(TypeError):
        hdf5.write(path, {'u': np.array(['unicode'])})
    with pytest.raises(ValueError):
        hdf5.read(os.path.join(GOLDEN, 'redrock-test.fits'))


def _toy_results(ntarget=4, seed=5):
    """allresults / allzfit shaped like zfind's output for 2 templates."""
    rng = np.random.default_rng(seed)
    tids = np.array([7, 39633393167631849, 3, 11], dtype=np.int64)[:ntarget]
    spec = {'GALAXY': (10, np.linspace(0.0, 1.0, 50)), 'QSO:::HIZ': (4, np.linspace(0.5, 3.0, 30))}
    allresults = {int(t): {ft: {'redshifts': z, 'zchi2': rng.uniform(1e3, 2e3, len(z)),
                                'penalty': rng.uniform(0, 5, len(z)),
                                'zcoeff': rng.normal(size=(len(z), nb))}
                           for ft, (nb, z) in spec.items()} for t in tids}
    n = 3 * ntarget
    zfit = {'targetid': np.repeat(tids, 3), 'z': rng.uniform(0, 2, n), 'zerr': rng.uniform(1e-5, 1e-4, n),
            'zwarn': rng.integers(0, 4, n), 'chi2': rng.uniform(1e3, 2e3, n),
            'zz': rng.uniform(0, 2, (n, 15)), 'zzchi2': rng.uniform(1e3, 2e3, (n, 15)),
            'coeff': rng.normal(size=(n, 10)), 'fitmethod': np.array(['PCA'] * n),
            'npixels': np.full(n, 7900), 'spectype': np.array(['GALAXY', 'QSO', 'GALAXY'] * ntarget),
            'subtype': np.array(['', 'HIZ', ''] * ntarget).astype('<U20'),
            'ncoeff': np.array([10, 4, 10] * ntarget), 'znum': np.tile([0, 1, 2], ntarget),
            'deltachi2': rng.uniform(0, 100, n)}
    return allresults, zfit


def test_write_zscan_roundtrip(tmp_path):
    allresults, zfit = _toy_results()
    path = str(tmp_path / 'sub' / 'rrdetails.h5')
    results.write_zscan(path, allresults, zfit)
    assert not os.path.exists(path + '.tmp')
    f = hdf5.read(path)
    tids = zfit['targetid'][zfit['znum'] == 0]
    assert np.array_equal(f['targetids'], tids)
    zb = f['zbest']
    assert 'znum' not in zb.dtype.names and zb['chi2'].dtype == np.float32
    assert zb['zwarn'].dtype == np.int32 and zb['ncoeff'].dtype == np.int8
    assert np.array_equal(zb['z'], zfit['z'][zfit['znum'] == 0])           # z stays float64
    zscan, back = results.read_zscan(path)
    for t in tids:
        for ft in ('GALAXY', 'QSO:::HIZ'):
            src, got = allresults[int(t)][ft], zscan[t][ft]
            assert np.array_equal(got['redshifts'], src['redshifts'])
            assert got['zchi2'].dtype == np.float32
            assert np.array_equal(got['zchi2'], src['zchi2'].astype(np.float32))
            assert np.array_equal(got['zcoeff'], src['zcoeff'].astype(np.float32))
            assert np.array_equal(got['penalty'], src['penalty'].astype(np.float32))
            assert len(got['zfit']['z']) == (2 if ft == 'GALAXY' else 1)
    # read_zfit stacks the targets in file order; compare row sets per (targetid, znum)
    got = np.lexsort((back['znum'], back['targetid']))
    src = np.lexsort((zfit['znum'], zfit['targetid']))
    for k in ('targetid', 'z', 'spectype', 'subtype', 'fitmethod', 'znum', 'npixels'):
        assert np.array_equal(back[k][got], zfit[k][src]), k
    for k in ('chi2', 'zerr', 'zz', 'coeff', 'deltachi2'):
        assert np.array_equal(back[k][got], zfit[k][src].astype(np.float32)), k
    with pytest.raises(OSError):
        results.write_zscan(str(tmp_path / 'sub' / 'rrdetails.h5' / 'x.h5'), allresults, zfit)
    results.write_zscan(path, allresults, zfit, clobber=True)
    assert results.read_zscan_redrock(path).keys() == zscan.keys()


def test_select_targetids_like_reference():
    """test_io.py:83-140 of the reference on its own details file."""
    zscan, zfit = results.read_zscan(DETAILS)
    for tid in np.unique(zfit['targetid']):
        for ft, zz in zscan[tid].items():
            assert len(zz['redshifts']) == len(zz['zchi2']) and len(zz['zfit']['z']) > 0
    tids = np.unique(zfit['targetid'])
    _, one = results.read_zscan(DETAILS, select_targetids=tids[1])
    assert list(np.unique(one['targetid'])) == [tids[1]]
    _, two = results.read_zscan(DETAILS, select_targetids=tids[0:2])
    assert np.array_equal(np.unique(two['targetid']), tids[0:2])
    _, up = results.read_zscan(DETAILS, upper=True)
    assert 'TARGETID' in up
    only = results.read_zscan(DETAILS, nozfit=True)
    assert isinstance(only, dict) and tids[0] in only
    assert len(np.unique(results.read_zfit(DETAILS)['targetid'])) == 3
    z2 = results.read_zfit(DETAILS, select_targetids=tids[0:2])
    assert np.array_equal(np.unique(z2['targetid']), tids[0:2])
    z1 = results.read_zfit(DETAILS, select_targetids=int(tids[1]))
    assert list(np.unique(z1['targetid'])) == [tids[1]]
    zu = results.read_zfit(DETAILS, upper=True)
    assert 'TARGETID' in zu and 'Z' in zu and 'z' not in zu
