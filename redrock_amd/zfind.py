"""Redshift finding: coarse scan, fine fits and the per-target result table.

Mirror of py/redrock/zfind.py (zfind.py:220-714) over the HIP scan
(redrock_amd.zscan) and the batched fitz (redrock_amd.fitz): the chi2 scan of
every target x template runs on the GPU, fitz refines the minima of every
target of a template in batched GPU rebin + scan launches, and the small
per-target assembly (priors, penalty, candidate sorting, nminima trim,
delta chi2 and zwarn bits) runs on the host, as in the reference.

The reference returns an astropy Table; astropy is not a dependency here, so
``allzfit`` is a dict of numpy columns in the reference's column order
(``targetid`` first), one row per (target, minimum).
"""
from __future__ import annotations

import numpy as np

from . import constants
from .fitz import fitz_columns, fitz_targets, get_dv
from .templates import parse_fulltype
from .zwarning import ZWarningMask as ZW
from .zwarning import badfit_mask


def sort_dict_by_col(d, colname):
    """Sort a dict of np.ndarrays by one key (zfind.py:35-48)."""
    if colname not in d:
        raise KeyError('Key ' + str(colname) + ' is not in dictionary')
    for k in d.keys():
        if type(d[k]) is not np.ndarray:
            raise ValueError('Column ' + str(k) + ' is not an np.array')
    idx = d[colname].argsort(0).flatten()
    for k in d.keys():
        d[k] = d[k][idx]


def sort_dict_by_cols(d, colnames, sort_first_column_first=True):
    """Sort a dict of np.ndarrays by several keys (zfind.py:50-81).

    With sort_first_column_first the first name is the primary key (the
    astropy Table.sort order); np.lexsort takes its last key as primary.
    """
    for c in colnames:
        if c not in d:
            raise KeyError('Key ' + str(c) + ' is not in dictionary')
    for k in d.keys():
        if type(d[k]) is not np.ndarray:
            raise ValueError('Column ' + str(k) + ' is not an np.array')
    keys = tuple(d[c] for c in colnames)
    if sort_first_column_first:
        keys = keys[::-1]
    idx = np.lexsort(keys, axis=0).flatten()
    for k in d.keys():
        d[k] = d[k][idx]


def sort_zfit_dict(zfit):
    """Sort candidate rows by (bad-fit zwarn bits, chi2), in place (zfind.py:206-218)."""
    zfit['__badfit__'] = (zfit['zwarn'] & badfit_mask) != 0
    sort_dict_by_cols(zfit, ('__badfit__', 'chi2'), sort_first_column_first=True)
    zfit.pop('__badfit__')


def calc_deltachi2(chi2, z, zwarn, dvlimit=constants.max_velo_diff):
    """chi2 difference of every candidate to the next good fit (zfind.py:103-155).

    deltachi2[i] = min over later candidates k > i that are good fits and lie
    further than min(dvlimit[i], dvlimit[k]) in velocity of chi2[k] - chi2[i]
    (0 when there is none).  setzwarn[i] is True when some other good fit, at
    such a velocity separation, is within constants.min_deltachi2 in chi2
    (either side).  Returns (deltachi2, setzwarn).
    """
    chi2 = np.asarray(chi2, dtype=np.float64)
    z = np.asarray(z, dtype=np.float64)
    n = len(chi2)
    okfit = (np.asarray(zwarn) & badfit_mask) == 0
    dvlimit = np.full(n, dvlimit, dtype=np.float64) if np.isscalar(dvlimit) \
        else np.asarray(dvlimit, dtype=np.float64)
    deltachi2 = np.zeros(n)
    setzwarn = np.zeros(n, dtype=bool)
    if n == 0:
        return deltachi2, setzwarn
    # pairwise velocity separations dv[i, k] = c (z_k - z_i) / (1 + z_i)
    dv = get_dv(z[None, :], z[:, None])
    lim = np.minimum(dvlimit[:, None], dvlimit[None, :])
    far = np.abs(dv) > lim
    dchi2 = chi2[None, :] - chi2[:, None]
    later = np.triu(np.ones((n, n), dtype=bool), k=1)
    cand = later & far & okfit[None, :]
    has = cand.any(axis=1)
    deltachi2[has] = np.where(cand, dchi2, np.inf).min(axis=1)[has]
    # SMALL_DELTA_CHI2: any other good fit at >= the velocity limit within min_deltachi2
    other = ~np.eye(n, dtype=bool)
    close = other & okfit[None, :] & (np.abs(dchi2) < constants.min_deltachi2) & (np.abs(dv) >= lim)
    setzwarn[:] = close.any(axis=1)
    return deltachi2, setzwarn


class Priors(object):
    """Redshift priors added to the coarse chi2 (priors.py:10-126).

    Priors(filename) reads TARGETID, Z, SIGMA and FUNCTION from the PRIORS HDU
    of a FITS file, as the reference does; Priors(targetid, z, sigma,
    function) takes the same columns as arrays.  function is one of
    'gaussian', 'lorentzien', 'tophat'.
    """

    def __init__(self, targetid, z=None, sigma=None, function='gaussian'):
        if isinstance(targetid, str):
            import os
            from . import fits
            t = fits.read(os.path.expandvars(targetid))['PRIORS'].data
            targetid, z, sigma = t['TARGETID'], t['Z'], t['SIGMA']
            function = str(t['FUNCTION'][0]).strip()
        self._param = {int(t): (float(a), float(s)) for t, a, s in zip(targetid, z, sigma)}
        if function not in ('gaussian', 'lorentzien', 'tophat'):
            raise ValueError(f'unknown prior function {function!r}')
        self._type = function
        self._func = getattr(self, function)

    def eval(self, targetid, z):
        """Prior chi2 contribution on the grid z (0. for unknown targets)."""
        if int(targetid) not in self._param:
            return 0.
        z0, s0 = self._param[int(targetid)]
        return self._func(np.asarray(z, dtype=np.float64), z0, s0)

    @staticmethod
    def gaussian(z, z0, s0):
        return ((z - z0) / s0) ** 2

    @staticmethod
    def lorentzien(z, z0, s0):
        return -2 * np.log(1 / (1. + ((z - z0) / s0) ** 2))

    @staticmethod
    def tophat(z, z0, s0):
        """0 inside |z - z0| < s0/2, +inf on the grid points bordering it, NaN outside."""
        prior = np.where(np.abs(z - z0) < s0 / 2, 0., np.nan)
        inside = np.flatnonzero(prior >= 0.0)
        if inside.size == 0:
            return prior
        left, right = inside[0], inside[-1]
        prior[left if left == 0 else left - 1] = np.inf
        prior[right if right == prior.size - 1 else right + 1] = np.inf
        return prior


def _local_targets(targets):
    return targets.local() if hasattr(targets, 'local') else list(targets)


def _all_ids(targets):
    if hasattr(targets, 'all_target_ids'):
        ids = targets.all_target_ids
        return list(ids() if callable(ids) else ids)
    return [t.id for t in targets]


class _Local(object):
    """DistTargets-like view of a plain list of targets."""

    def __init__(self, targets):
        self._t = list(targets)

    def local(self):
        return self._t

    def local_target_ids(self):
        return [t.id for t in self._t]


def zfind(targets, templates, mp_procs=1, nminima=3, archetypes=None, priors=None,
          chi2_scan=None, use_gpu=False, zminfit_npoints=15, per_camera=None,
          deg_legendre=None, n_nearest=None, prior_sigma=None, ncamera=None,
          precision=None, newton=False, polish=None):
    """All redshift fits of the targets (zfind.py:220-714).

    targets: DistTargets (redrock_amd.targets) or a list of Target;
    templates: DistTemplate list.  priors: a Priors object (or None).
    precision: scan precision ('f64', 'f16x3', ...; default
    zscan.default_precision()); the fine fits always run in FP64.
    polish: FP64 re-evaluation of the scan rows near each minimum after a
    mixed-precision scan (zscan.ScanPlan; default on for mixed precisions,
    False turns it off).
    chi2_scan: a details file (results.write_zscan) whose scan is used
    instead of scanning again; the fine fits still run.
    newton=True adds 'znewton'/'chi2newton' columns (Newton refinement on the
    analytic d chi2 / dz beside the parabola fit).
    archetypes: {spectype: Archetype} (All_archetypes(...).archetypes): the
    minima of those types are fit by their archetypes (redrock_amd.archetypes)
    with deg_legendre Legendre polynomials (per camera when per_camera),
    prior_sigma on their coefficients and n_nearest archetypes combined.

    Returns (allresults, allzfit): allresults[targetid][fulltype] =
    dict(redshifts, zchi2, penalty, zcoeff); allzfit the dict of columns.
    """
    import time
    clock = time.perf_counter
    timings = dict(fitz=dict())
    last_timings.clear()
    t0 = clock()
    if isinstance(priors, str):
        priors = Priors(priors)
    local = _local_targets(targets)
    dist = targets if hasattr(targets, 'local_target_ids') else _Local(local)
    from . import zscan
    from .zscan import target_batches
    zscan.unpolished.clear()
    batches = target_batches(local) if local else dict()
    one = next(iter(batches.values())) if len(batches) == 1 else None
    # good pixels per target (the npixels column, fitz.py:388-390), once for every template
    if one is not None:
        npix = [int(n) for n in one.good_pixels]
    else:
        npix = [sum(int(np.count_nonzero(s.ivar > 0.)) for s in tg.spectra) for tg in local]
    # without archetypes or Newton columns the fits go to the table as flat
    # columns (fitz_columns) instead of per-target dicts
    blocks = dict() if not archetypes and not newton else None

    def fit(t, chi2):
        """The fine fits of every target for template t (zfind.py:330-367)."""
        ft = t.template.full_type
        tf = clock()
        archetype = None
        if archetypes and t.template.template_type in archetypes:
            archetype = archetypes[t.template.template_type]
        if blocks is not None:
            blocks[ft] = ([tg.id for tg in local],) + fitz_columns(
                chi2, t.template.redshifts, local, t.template, nminima=nminima,
                zminfit_npoints=zminfit_npoints, batch=one, npixels=npix)
        else:
            fits = fitz_targets(chi2, t.template.redshifts, local, t.template, nminima=nminima,
                                zminfit_npoints=zminfit_npoints, newton=newton, archetype=archetype,
                                deg_legendre=deg_legendre, per_camera=bool(per_camera),
                                n_nearest=n_nearest, prior_sigma=prior_sigma, batch=one,
                                npixels=npix)
            for tg, zfit in zip(local, fits):
                results[tg.id][ft]['zfit'] = zfit
        timings['fitz'][ft] = clock() - tf

    def with_priors(ft, redshifts):
        """zchi2 += the prior (zfind.py:318-322); the (nt, nz) zchi2 + penalty."""
        for tid in results:
            results[tid][ft]['zchi2'] = results[tid][ft]['zchi2'] + priors.eval(tid, redshifts)
        return np.stack([results[tg.id][ft]['zchi2'] + results[tg.id][ft]['penalty']
                         for tg in local])

    if chi2_scan is None and one is not None:                # zfind.py:289-292
        # each template's fine fits start on a second stream as soon as its
        # scan is on the host, while the later templates' scans run
        import torch
        side = zscan.pipeline_streams()[2]
        side.wait_stream(torch.cuda.current_stream())
        results = {tid: dict() for tid in dist.local_target_ids()}
        for t, zchi2, zcoeff, penalty in zscan.scan_templates_iter(local, templates, precision, one,
                                                                   polish):
            ft = t.template.full_type
            zs = t.template.redshifts
            for tg, c2, pn, cf in zip(local, list(zchi2), list(penalty), list(zcoeff)):
                results[tg.id][ft] = dict(redshifts=zs, zchi2=c2, penalty=pn, zcoeff=cf)
            chi2 = zchi2 + penalty if priors is None else with_priors(ft, t.template.redshifts)
            timings['scan'] = clock() - t0
            with torch.cuda.stream(side):
                fit(t, chi2)
        torch.cuda.current_stream().wait_stream(side)
    else:
        full = dict()                 # full_type -> (nt, nz) zchi2 + penalty of `local`
        if chi2_scan is None:
            results, full = zscan._scan_templates(dist, templates, precision=precision,
                                                  batches=batches, polish=polish)
        else:
            from .results import read_zscan_redrock
            saved = read_zscan_redrock(chi2_scan)
            results = dict()
            for tg in local:
                if tg.id not in saved:
                    raise KeyError(f'{chi2_scan}: no scan for target {tg.id}')
                results[tg.id] = {ft: {k: np.asarray(v, dtype=np.float64) for k, v in d.items()}
                                  for ft, d in saved[tg.id].items()}
        timings['scan'] = clock() - t0
        if priors is not None:                               # zfind.py:318-322
            full = dict()
            for tid in results:
                for ft in results[tid]:
                    results[tid][ft]['zchi2'] = (results[tid][ft]['zchi2']
                                                 + priors.eval(tid, results[tid][ft]['redshifts']))
        order = np.array([t.template.full_type for t in templates]).argsort()
        for t in np.array(list(templates), dtype=object)[order]:      # zfind.py:330-367
            ft = t.template.full_type
            if not local:
                continue
            chi2 = full.get(ft)
            if chi2 is None:
                chi2 = np.stack([results[tg.id][ft]['zchi2'] + results[tg.id][ft]['penalty']
                                 for tg in local])
            fit(t, chi2)

    for tg in local:
        results[tg.id]['meta'] = getattr(tg, 'meta', dict())

    arch = None
    if archetypes:
        if ncamera is None:
            ncamera = len({s.wavehash for s in local[0].spectra}) if local else 1
        arch = dict(spectypes=list(archetypes.keys()), per_camera=bool(per_camera),
                    deg_legendre=0 if deg_legendre is None else int(deg_legendre),
                    n_nearest=n_nearest, ncamera=ncamera)
    ta = clock()
    if arch is None:
        allzfit = _assemble_table(results, _all_ids(targets), templates, nminima, blocks)
    else:
        allzfit = _assemble(results, _all_ids(targets), templates, nminima, arch=arch)
    for tid in results:
        results[tid].pop('meta', None)
    timings['assemble'] = clock() - ta
    timings['total'] = clock() - t0
    timings['unpolished'] = dict(zscan.unpolished)
    last_timings.update(timings)
    return results, allzfit


# wall-clock seconds of the phases of the last zfind call (scan, fitz per
# template, assemble, total): host timings for profiling, not part of the API;
# with one wavelength layout a template's fits overlap the later scans, and
# 'scan' is the time until the last scan result reached the host;
# 'unpolished' maps each template whose mixed-precision scan ran without the
# FP64 polish to the reason
last_timings = dict()


def _pad_cols(val, width):
    """Columns of a 2-D numeric array padded with zeros / cut to `width` (zfind.py:537-546)."""
    val = np.asarray(val)
    if val.ndim < 2 or val.shape[1] == width or val.dtype.kind in 'US':
        return val
    if val.shape[1] < width:
        val = np.pad(val, ((0, 0), (0, width - val.shape[1])), constant_values=0)
    return val[:, :width]


def _assemble_table(allresults, target_ids, templates, nminima, blocks=None):
    """_assemble without archetypes, with every target's rows handled at once:
    the same table, row order and dtypes (the per-target sorts, the nminima
    per spectype trim and calc_deltachi2 become sorts keyed by target and
    padded (target, candidate, candidate) arrays).  blocks: {fulltype:
    (target ids, nacc, columns)} from fitz_columns; templates without a block
    come from the per-target 'zfit' dicts."""
    blocks = dict(blocks or {})
    pos = {tid: i for i, tid in enumerate(target_ids)}
    fts = []
    for tid in target_ids:
        for ft in allresults[tid]:
            if ft != 'meta' and ft not in fts:
                fts.append(ft)
    for ft in fts:                        # per-target dicts -> the same flat columns
        if ft in blocks:
            continue
        ids = [tid for tid in target_ids if ft in allresults[tid]]
        per = [allresults[tid][ft].pop('zfit') for tid in ids]
        nacc = np.array([len(z['chi2']) for z in per], dtype=np.int64)
        full = [z for z in per if len(z['chi2'])]
        cols = {k: (lambda v: v.reshape(-1) if v.ndim == 2 and v.shape[1] == 1 and k != 'coeff' else v)(
            np.vstack([z[k] for z in full])) for k in (full[0] if full else [])}
        blocks[ft] = (ids, nacc, cols)
    keys = None
    parts = []                            # (target position, template rank, candidate, ft, cols)
    # a template's rank among each target's templates: one number when every
    # target holds the templates in the same order (zfind fills them so)
    ref = tuple(fts)
    uniform = all(tuple(k for k in allresults[tid] if k != 'meta') == ref for tid in target_ids)
    for ft in fts:
        ids, nacc, cols = blocks[ft]
        for i in np.flatnonzero(np.asarray(nacc) == 0):
            print(f'WARNING: no {ft} chi2 vs. z minima for targetid {ids[i]}')
        if int(np.sum(nacc)) == 0:
            continue
        if keys is None:
            keys = list(cols.keys())
        tpos_b = np.repeat(np.array([pos[tid] for tid in ids], dtype=np.int64), nacc)
        rank = (np.full(len(ids), fts.index(ft), dtype=np.int64) if uniform else
                np.array([list(k for k in allresults[tid] if k != 'meta').index(ft) for tid in ids]))
        cand = np.arange(tpos_b.size) - np.repeat(np.cumsum(nacc) - nacc, nacc)
        parts.append((tpos_b, np.repeat(rank, nacc), cand, ft, cols))
    if keys is None:
        return dict()
    # rows in (target, its template order, candidate) order, as the per-target loop appends them
    tp_all = np.concatenate([p[0] for p in parts])
    orig = np.lexsort((np.concatenate([p[2] for p in parts]), np.concatenate([p[1] for p in parts]),
                       tp_all))
    maxnc = max(p[4]['coeff'].shape[1] for p in parts)
    nrow = tp_all.size
    coeff = np.zeros((nrow, maxnc), dtype=np.result_type(*[p[4]['coeff'] for p in parts]))
    r0 = 0
    for p in parts:
        c = p[4]['coeff']
        coeff[r0:r0 + c.shape[0], :c.shape[1]] = c
        r0 += c.shape[0]
    tab = dict(coeff=coeff[orig])
    for k in keys:
        if k == 'coeff':
            continue
        tab[k] = np.concatenate([p[4][k] for p in parts])[orig]
    spec, subt, ncf, mvd = [], [], [], []
    for p in parts:
        spectype, subtype = parse_fulltype(p[3])
        n = p[0].size
        spec.append(np.full(n, spectype))
        subt.append(np.full(n, subtype))
        ncf.append(np.full(n, p[4]['coeff'].shape[1]))
        mvd.append(np.full(n, constants.max_velo_diff_star if spectype.upper() == 'STAR'
                           else constants.max_velo_diff))
    tab['spectype'] = np.concatenate(spec)[orig]
    tab['subtype'] = np.concatenate(subt)[orig]
    tab['ncoeff'] = np.concatenate(ncf)[orig]
    mv = np.concatenate(mvd)[orig]
    tp = tp_all[orig]
    zw = tab['zwarn']
    zw[tab['npixels'] == 0] |= ZW.NODATA
    zw[tab['npixels'] < 10 * tab['ncoeff']] |= ZW.LITTLE_COVERAGE
    n = len(tp)
    # per target: sort by (bad fit, chi2), stable (sort_zfit_dict)
    bad = (zw & badfit_mask) != 0
    first = np.lexsort((tab['chi2'], bad, tp))
    rank_sorted = np.empty(n, dtype=np.int64)
    rank_sorted[first] = np.arange(n)
    # nminima per spectype: rank of each row among its (target, spectype) in that order
    scode = np.unique(tab['spectype'], return_inverse=True)[1].reshape(-1)
    o = np.lexsort((rank_sorted, scode, tp))
    grp = np.r_[True, (tp[o][1:] != tp[o][:-1]) | (scode[o][1:] != scode[o][:-1])]
    start = np.maximum.accumulate(np.where(grp, np.arange(n), 0))
    within = np.empty(n, dtype=np.int64)
    within[o] = np.arange(n) - start
    keep = within < nminima
    # a target that lost rows is re-sorted from the spectype-grouped order, so
    # (bad fit, chi2) ties then follow spectype (the reference's second sort)
    dropped = np.zeros(len(target_ids), dtype=bool)
    dropped[tp[~keep]] = True
    tie = np.where(dropped[tp], scode, 0)
    idx = np.flatnonzero(keep)
    idx = idx[np.lexsort((rank_sorted[idx], tie[idx], tab['chi2'][idx], bad[idx], tp[idx]))]
    for k in tab:
        tab[k] = tab[k][idx]
    tp, mv = tp[idx], mv[idx]
    n = len(tp)
    # znum and calc_deltachi2 per target over padded (target, row) arrays
    gstart = np.r_[True, tp[1:] != tp[:-1]]
    g = np.cumsum(gstart) - 1
    first_row = np.flatnonzero(gstart)
    znum = np.arange(n) - first_row[g]
    ng, G = len(first_row), int(znum.max()) + 1
    valid = np.zeros((ng, G), dtype=bool)
    valid[g, znum] = True
    C = np.zeros((ng, G))
    Z = np.zeros((ng, G))
    L = np.zeros((ng, G))
    OK = np.zeros((ng, G), dtype=bool)
    C[g, znum], Z[g, znum], L[g, znum] = tab['chi2'], tab['z'], mv
    OK[g, znum] = (tab['zwarn'] & badfit_mask) == 0
    dv = get_dv(Z[:, None, :], Z[:, :, None])            # [t, i, k] = dv of k relative to i
    lim = np.minimum(L[:, :, None], L[:, None, :])
    far = np.abs(dv) > lim
    dchi2 = C[:, None, :] - C[:, :, None]
    okk = (OK & valid)[:, None, :]
    later = np.triu(np.ones((G, G), dtype=bool), k=1)[None]
    cand = later & far & okk
    has = cand.any(axis=2)
    dmin = np.where(cand, dchi2, np.inf).min(axis=2)
    deltachi2 = np.where(has, dmin, 0.0)[g, znum]
    other = ~np.eye(G, dtype=bool)[None]
    close = other & okk & (np.abs(dchi2) < constants.min_deltachi2) & (np.abs(dv) >= lim)
    setzwarn = close.any(axis=2)[g, znum]
    tab['targetid'] = np.asarray(target_ids)[tp]
    tab['znum'] = znum
    tab['deltachi2'] = deltachi2
    tab['zwarn'][setzwarn] |= ZW.SMALL_DELTA_CHI2
    order = keys + ['spectype', 'subtype', 'ncoeff', 'targetid', 'znum', 'deltachi2']
    names = ['targetid'] + [k for k in order if k != 'targetid' and 'pca' not in k.lower()] + \
        [k for k in order if 'pca' in k.lower()]
    allzfit = {k: tab[k] for k in names}
    return _finish_table(allzfit, allresults, templates, None)


def _finish_table(allzfit, allresults, templates, arch):
    """Metadata columns, string widths and the COEFF width (zfind.py:598-693)."""
    meta0 = allresults[allzfit['targetid'][0]].get('meta', dict())
    metakeys = sorted(k for k in meta0 if not k.endswith('_datatype'))
    for mk in metakeys:
        dt = meta0.get(mk + '_datatype')
        vals = [allresults[t]['meta'][mk] for t in allzfit['targetid']]
        allzfit[mk] = np.array(vals, dtype=dt) if dt is not None else np.array(vals)
    allzfit['subtype'] = allzfit['subtype'].astype(
        f"<U{max(20, max((len(s) for s in allzfit['subtype']), default=0))}")
    allzfit['spectype'] = allzfit['spectype'].astype('<U6')
    maxcoeff = max(t.template.nbasis for t in templates)
    if arch is not None and arch['per_camera']:                # zfind.py:677-683
        maxcoeff = max(maxcoeff, arch['ncamera'] * arch['deg_legendre'] + (arch['n_nearest'] or 1))
    coeff = allzfit['coeff']
    if coeff.shape[1] != maxcoeff:
        c = np.zeros((coeff.shape[0], maxcoeff), dtype=coeff.dtype)
        c[:, :coeff.shape[1]] = coeff
        coeff = c
    nc = np.unique(allzfit['ncoeff'])
    if len(nc) == 1:
        coeff = coeff[:, :int(nc[0])]
    allzfit['coeff'] = coeff
    return allzfit


def _assemble(allresults, target_ids, templates, nminima, arch=None):
    """Per-target candidate tables -> one table (zfind.py:467-714).

    arch: None, or the archetype settings (spectypes with archetypes,
    per_camera, deg_legendre, n_nearest, ncamera) for the archetype-mode
    columns: per-row spectype/subtype from the archetype fulltype, pca_*
    columns, legcoeff, NEGATIVE_MODEL and the COEFF width (zfind.py:476-693).
    """
    from .utils import extra_columns_in_archetype_mode
    pca_map, zero_like_keys = extra_columns_in_archetype_mode()
    all_spectype = []
    for t in (templates if arch is not None else []):
        if t.template.template_type not in all_spectype:
            all_spectype.append(t.template.template_type)
    rows = []
    for tid in target_ids:
        tzfit = []
        for fulltype in allresults[tid]:
            if fulltype == 'meta':
                continue
            tmp = allresults[tid][fulltype].pop('zfit')
            if arch is not None and 'fulltype' in tmp:
                parsed = [parse_fulltype(el[0]) for el in tmp.pop('fulltype')]
                spectype = [p[0] for p in parsed]
                subtype = [p[1] for p in parsed]
            else:
                spectype, subtype = parse_fulltype(fulltype)
            nmin = len(tmp['chi2'])
            if nmin == 0:
                print(f'WARNING: no {fulltype} chi2 vs. z minima for targetid {tid}')
                continue
            if isinstance(spectype, str):
                tmp['spectype'] = np.full((nmin, 1), spectype)
                tmp['subtype'] = np.full((nmin, 1), subtype)
                star = spectype.upper() == 'STAR'
            else:
                tmp['spectype'] = np.array([spectype]).reshape((nmin, 1))
                tmp['subtype'] = np.array([subtype]).reshape((nmin, 1))
                star = spectype[0].upper() == 'STAR'
            tmp['ncoeff'] = np.full((nmin, 1), tmp['coeff'].shape[1])
            mv = constants.max_velo_diff_star if star else constants.max_velo_diff
            tmp['max_velo_diff'] = np.full((nmin, 1), mv)
            tzfit.append(tmp)
        if not tzfit:
            continue
        maxnc = max(tmp['coeff'].shape[1] for tmp in tzfit)
        for tmp in tzfit:
            n = maxnc - tmp['coeff'].shape[1]
            if n > 0:
                tmp['coeff'] = np.append(tmp['coeff'], np.zeros((tmp['coeff'].shape[0], n)), axis=1)
        tz = dict()
        for k in tzfit[0].keys():
            ref = tzfit[0][k]
            vals, val = [], None
            for tmp in tzfit:
                if k in tmp:
                    val = tmp[k]
                    if arch is not None and k in pca_map:
                        val = _pad_cols(val, ref.shape[1])
                elif arch is not None and k in pca_map and pca_map[k] in tmp:
                    val = _pad_cols(tmp[pca_map[k]], ref.shape[1])
                elif k in zero_like_keys:
                    val = np.zeros((len(tmp['chi2']),) + np.shape(ref)[1:], dtype=np.asarray(ref).dtype)
                vals.append(val)
            tz[k] = np.vstack(vals)
            if k != 'coeff' and tz[k].shape[1] == 1:
                tz[k] = tz[k].flatten()
        tz['targetid'] = np.full(len(tz['chi2']), tid)
        if arch is not None:                                   # zfind.py:558-579
            nn = arch['n_nearest']
            coeff = tz['coeff'] if tz['coeff'].ndim > 1 else tz['coeff'].reshape(-1, 1)
            if len(arch['spectypes']) == len(all_spectype):
                ibad = coeff[:, 0] <= 0. if nn is None else np.any(coeff[:, :nn] < 0, axis=1)
            else:
                ibad = np.isin(tz['spectype'], np.array(arch['spectypes']))
                for k in np.where(ibad)[0]:
                    if (coeff[k, 0] >= 0.) if nn is None else np.all(coeff[k, :nn] >= 0):
                        ibad[k] = False
            tz['zwarn'][ibad] |= ZW.NEGATIVE_MODEL
        tz['zwarn'][tz['npixels'] == 0] |= ZW.NODATA
        tz['zwarn'][tz['npixels'] < 10 * tz['ncoeff']] |= ZW.LITTLE_COVERAGE
        sort_zfit_dict(tz)
        keep = []
        for spectype in np.unique(tz['spectype']):            # nminima per spectype
            keep.extend(np.where(tz['spectype'] == spectype)[0][:nminima])
        if len(keep) < len(tz['chi2']):
            for k in tz:
                tz[k] = tz[k][keep]
            sort_zfit_dict(tz)
        tz['znum'] = np.arange(len(tz['chi2']))
        dchi2, setzwarn = calc_deltachi2(tz['chi2'], tz['z'], tz['zwarn'], dvlimit=tz['max_velo_diff'])
        tz['deltachi2'] = dchi2
        tz['zwarn'][setzwarn] |= ZW.SMALL_DELTA_CHI2
        del tz['max_velo_diff']
        rows.append(tz)
    if not rows:
        return dict()
    # stack the targets (astropy vstack: coeff padded to the widest)
    maxnc = max(r['coeff'].shape[1] for r in rows)
    for r in rows:
        n = maxnc - r['coeff'].shape[1]
        if n > 0:
            r['coeff'] = np.append(r['coeff'], np.zeros((r['coeff'].shape[0], n)), axis=1)
    # targetid first, pca_* columns last (zfind.py:624-631)
    keys = [k for k in rows[0].keys() if k != 'targetid']
    names = ['targetid'] + [k for k in keys if 'pca' not in k.lower()] + [k for k in keys if 'pca' in k.lower()]
    allzfit = {k: np.concatenate([r[k] for r in rows]) for k in names}
    # target metadata columns (zfind.py:598-636): '<key>_datatype' entries give dtypes
    return _finish_table(allzfit, allresults, templates, arch)


ZBEST_COLUMNS = ('TARGETID', 'Z', 'ZERR', 'ZWARN', 'CHI2', 'COEFF', 'NPIXELS', 'SPECTYPE',
                 'SUBTYPE', 'NCOEFF', 'DELTACHI2')


def zbest_table(allzfit):
    """The best fit of every target (znum == 0) as the REDSHIFTS columns of
    redrock's output file (the rrdesi table; upper-case names)."""
    best = allzfit['znum'] == 0
    return {name: np.asarray(allzfit[name.lower()])[best] for name in ZBEST_COLUMNS}


def write_zbest(path, allzfit, header=None):
    """Write zbest_table(allzfit) as the REDSHIFTS HDU of a FITS file."""
    from . import fits
    fits.write(path, [fits.table_hdu(zbest_table(allzfit), 'REDSHIFTS')], primary_header=header)
