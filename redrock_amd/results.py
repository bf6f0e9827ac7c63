"""redrock's details file: the full χ² scan and all fits of every target, as HDF5.

Same file layout and API as the reference's results module
(results.py:18-241), on this package's own HDF5 reader/writer
(redrock_amd.hdf5; h5py is not a dependency):

    /targetids[nt]                              int64
    /zbest                                      table of the best fits (znum == 0)
    /zscan/{fulltype}/redshifts[nz]             float64
    /zscan/{fulltype}/zchi2[nt, nz]             float32
    /zscan/{fulltype}/penalty[nt, nz]           float32
    /zscan/{fulltype}/zcoeff[nt, nz, nc]        float32
    /zfit/{targetid}/zfit                       table of that target's fits

Tables are dicts of columns here (zfind's allzfit); in the file they are
compound datasets, one row per fit.  A details file written by the reference
(older versions kept float64 columns and no /zbest) reads back through
read_zscan / read_zfit / read_zscan_redrock as well, and read_zscan_redrock's
result is what zfind(chi2_scan=...) resumes from.
"""
from __future__ import annotations

import os

import numpy as np

from . import hdf5
from .templates import parse_fulltype, make_fulltype

_BYTES_COLUMNS = ('spectype', 'subtype', 'fitmethod', 'pca_spectype', 'pca_subtype')
_FLOAT32_COLUMNS = ('zerr', 'chi2', 'zz', 'zzchi2', 'coeff', 'deltachi2')


def _table_to_records(table):
    """Dict of columns (1-D or 2-D) -> structured array, column order kept."""
    names = list(table.keys())
    cols = [np.asarray(table[k]) for k in names]
    n = len(cols[0]) if cols else 0
    fmt = []
    for k, c in zip(names, cols):
        if len(c) != n:
            raise ValueError(f'column {k} has {len(c)} rows, expected {n}')
        fmt.append((k, c.dtype, c.shape[1:]) if c.ndim > 1 else (k, c.dtype))
    rec = np.zeros(n, dtype=fmt)
    for k, c in zip(names, cols):
        rec[k] = c
    return rec


def _records_to_table(rec, upper=False):
    """Structured array -> dict of columns; byte-string columns become str."""
    out = dict()
    for k in rec.dtype.names:
        c = rec[k]
        if c.dtype.kind == 'S':
            c = c.astype((str, c.dtype.itemsize))           # utils.py:42-52
        out[k.upper() if upper else k] = np.array(c)
    return out


def write_zscan(filename, zscan, zfit, clobber=False):
    """Write zfind's results to an HDF5 details file (results.py:18-97).

    zscan: allresults[targetid][fulltype] = dict(redshifts, zchi2, penalty,
    zcoeff); zfit: the table of all fits (dict of columns, with 'targetid'
    and 'znum').  The file is written under a temporary name and renamed.
    """
    filename = os.path.expandvars(filename)
    if clobber and os.path.exists(filename):
        os.remove(filename)
    outdir = os.path.dirname(os.path.abspath(filename))
    if not os.path.exists(outdir):
        os.makedirs(outdir)

    zfit = {k: np.asarray(v) for k, v in zfit.items()}
    for col, v in zfit.items():
        if v.dtype.kind == 'O':       # e.g. pca_subtype of templates made without a subtype (None)
            zfit[col] = np.array(['' if x is None else str(x) for x in v.ravel()]).reshape(v.shape)
    for col in _BYTES_COLUMNS:
        if col in zfit and zfit[col].dtype.kind == 'U':
            zfit[col] = np.char.encode(zfit[col], 'ascii')
    for col in _FLOAT32_COLUMNS:
        zfit[col] = zfit[col].astype(np.float32)
    zfit['zwarn'] = zfit['zwarn'].astype(np.int32)
    zfit['npixels'] = zfit['npixels'].astype(np.int32)
    zfit['znum'] = zfit['znum'].astype(np.int8)
    zfit['ncoeff'] = zfit['ncoeff'].astype(np.int8)

    best = zfit['znum'] == 0
    zbest = {k: v[best] for k, v in zfit.items() if k != 'znum'}
    targetids = np.asarray(zbest['targetid'])
    spectypes = list(zscan[targetids[0]].keys())

    tree = {'targetids': targetids, 'zbest': _table_to_records(zbest), 'zscan': dict(), 'zfit': dict()}
    for ft in spectypes:
        tree['zscan'][ft] = {
            'zchi2': np.vstack([np.asarray(zscan[t][ft]['zchi2'], dtype=np.float32) for t in targetids]),
            'penalty': np.vstack([np.asarray(zscan[t][ft]['penalty'], dtype=np.float32) for t in targetids]),
            'zcoeff': np.stack([np.asarray(zscan[t][ft]['zcoeff'], dtype=np.float32) for t in targetids]),
            'redshifts': np.asarray(zscan[targetids[0]][ft]['redshifts']),
        }
    records = _table_to_records(zfit)
    for tid in targetids:
        tree['zfit'][str(tid)] = {'zfit': records[zfit['targetid'] == tid]}

    tmp = filename + '.tmp'
    hdf5.write(tmp, tree)
    os.rename(tmp, filename)


def _select(targetids, select_targetids):
    if select_targetids is None:
        return np.arange(len(targetids))
    if np.isscalar(select_targetids):
        select_targetids = [select_targetids]
    return np.where(np.isin(targetids, select_targetids))[0]


def read_zfit(filename, select_targetids=None, upper=False):
    """The stacked zfit tables of the selected targets (results.py:183-209)."""
    f = hdf5.read(os.path.expandvars(filename))
    if select_targetids is None:
        select_targetids = f['targetids']
    elif np.isscalar(select_targetids):
        select_targetids = [select_targetids]
    rec = np.concatenate([f[f'zfit/{tid}/zfit'] for tid in select_targetids])
    return _records_to_table(rec, upper=upper)


def read_zscan(filename, select_targetids=None, upper=False, nozfit=False):
    """zscan[targetid][fulltype] = dict(redshifts, zchi2, penalty, zcoeff, zfit)
    and the zfit table, or zscan alone with nozfit (results.py:100-181)."""
    f = hdf5.read(os.path.expandvars(filename))
    targetids = f['targetids']
    fulltypes = list(f['zscan'].keys())
    indx = _select(targetids, select_targetids)
    zscan = {tid: {ft: dict() for ft in fulltypes} for tid in targetids[indx]}
    for ft in fulltypes:
        g = f['zscan'][ft]
        for i, tid in zip(indx, targetids[indx]):
            zscan[tid][ft]['redshifts'] = g['redshifts']
            zscan[tid][ft]['zchi2'] = g['zchi2'][i]
            zscan[tid][ft]['penalty'] = g['penalty'][i]
            zscan[tid][ft]['zcoeff'] = g['zcoeff'][i]
    if nozfit:
        return zscan
    zfit = read_zfit(filename, select_targetids=select_targetids, upper=upper)
    tcol, scol, sub = ('TARGETID', 'SPECTYPE', 'SUBTYPE') if upper else ('targetid', 'spectype', 'subtype')
    keys = list(zip(zfit[tcol], zfit[scol], zfit[sub]))
    for key in dict.fromkeys(keys):
        rows = np.array([k == key for k in keys])
        ft = make_fulltype(key[1], key[2])
        if ft in zscan[key[0]]:       # archetype fits name no scanned template: table only
            zscan[key[0]][ft]['zfit'] = {k: v[rows] for k, v in zfit.items()}
    return zscan, zfit


def read_zscan_redrock(filename):
    """results[targetid][fulltype] = dict(penalty, zcoeff, zchi2, redshifts):
    the scan in the form zfind resumes from (results.py:212-241)."""
    f = hdf5.read(os.path.expandvars(filename))
    targetids = f['targetids']
    out = dict()
    for i, tid in enumerate(targetids):
        out[tid] = dict()
        for ft in f['zscan'].keys():
            g = f['zscan'][ft]
            out[tid][ft] = {'redshifts': g['redshifts'], 'zchi2': g['zchi2'][i],
                            'penalty': g['penalty'][i], 'zcoeff': g['zcoeff'][i]}
    return out


__all__ = ['write_zscan', 'read_zscan', 'read_zfit', 'read_zscan_redrock', 'parse_fulltype']
