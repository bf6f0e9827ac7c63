"""Refinement of chi2 minima (py/redrock/fitz.py) on the GPU, batched over targets.

fitz(zchi2, redshifts, target, template, nminima=3, ...) keeps the reference
signature (fitz.py:133-395, non-archetype branch).  fitz_targets() refines all
targets of one template at once:

1. every target's candidate minima come from find_minima (fitz.py:46-67);
2. the first few candidates of every target are refined speculatively in one
   pass: one rr_rebin launch per camera builds the template at every
   candidate's 15-point fine grid (IGM fused), one rr_zscan launch computes
   their chi2 and coefficients (FP64), a second rebin + scan gives the
   coefficients at each parabola minimum (fitz.py:265-293);
3. the reference's sequential control flow — skipping minima within
   max_velo_diff of accepted ones, the parabola fit, zwarn bits, dedup and
   sort — is replayed on the host from those results for every target at
   once (select_minima: one numpy step per candidate position over all
   targets still choosing).  A target that reaches a candidate not refined
   yet waits; the next candidates of all waiting targets are refined
   together and they resume where they stopped, so the outcome is the
   reference's.
"""
from __future__ import annotations

import numpy as np

from . import constants
from .zwarning import ZWarningMask as ZW


def get_dv(z, zref):
    """Velocity difference in km/s (fitz.py:28-43)."""
    return constants.c_kms * (z - zref) / (1.0 + zref)


def find_minima(x):
    """Indices of local minima of x including edges, sorted by x (fitz.py:46-67)."""
    x = np.asarray(x)
    m = np.ones(x.shape, dtype=bool)
    m[1:] = x[1:] <= x[:-1]
    m[:-1] &= x[:-1] <= x[1:]
    ii = np.flatnonzero(m)
    jj = np.argsort(x[ii])
    return ii[jj]


def minima_head(x, k):
    """The first k candidates of find_minima for every row of a 2-D x:
    (head, ncand) with head[t, :min(k, ncand[t])] == find_minima(x[t])[:k]
    (-1 after) and ncand[t] = find_minima(x[t]).size.  One host pass per row
    in the library (rr_minima_head); rows where ties or non-finite values
    could leave the order to np.argsort's tie handling take find_minima."""
    from . import _hip
    x = np.ascontiguousarray(x, dtype=np.float64)
    nt, nz = x.shape
    k = int(k)
    head = np.empty((nt, k), dtype=np.int64)
    ncand = np.empty(nt, dtype=np.int64)
    flag = np.empty(nt, dtype=np.uint8)
    if nt:
        rc = _hip.lib().rr_minima_head(x.ctypes.data, nt, nz, nz, k, head.ctypes.data,
                                       ncand.ctypes.data, flag.ctypes.data)
        _hip.check(rc, 'rr_minima_head')
    for t in np.flatnonzero(flag):
        c = find_minima(x[t])[:k]
        head[t] = -1
        head[t, :c.size] = c
    return head, ncand


def _ranges(b, n):
    """(i, k): for every i, k runs over b[i], ..., b[i] + n[i] - 1, concatenated."""
    n = np.asarray(n, dtype=np.int64)
    i = np.repeat(np.arange(n.size), n)
    k = np.arange(i.size, dtype=np.int64) - np.repeat(np.cumsum(n) - n, n) + np.repeat(b, n)
    return i, k


def minfit_batch(x, y):
    """minfit of many 3-point sets at once: x, y (K, 3), x increasing in each
    row.  Returns arrays (x0, xerr, y0, zwarn) of length K.

    A degree-2 fit to 3 distinct points is the interpolating parabola, here
    in closed form about the middle point, p(u) = y1 + b u + a u^2 with
    u = x - x1: vertex x1 - b/(2a), value y1 - b^2/(4a), xerr = 1/sqrt(|a|),
    and the reference's BAD_MINFIT tests (fitz.py:94-111).  np.polyfit on the
    raw Vandermonde matrix (the reference) has condition ~ (x/dx)^2 ~ 1e8 for
    a fine redshift grid, so its vertex carries rounding errors up to ~1e-11
    relative, which this form does not.  Rows that are not 3 increasing
    points with finite y, or whose curvature is exactly 0, go through the
    reference's np.polyfit path (minfit) unchanged."""
    x = np.asarray(x, dtype=np.float64)
    y = np.asarray(y, dtype=np.float64)
    K = x.shape[0]
    x0 = np.full(K, -1.0)
    xerr = np.full(K, -1.0)
    y0 = np.full(K, -1.0)
    zwarn = np.zeros(K, dtype=np.int64)
    if K == 0:
        return x0, xerr, y0, zwarn
    u0 = x[:, 0] - x[:, 1]
    u2 = x[:, 2] - x[:, 1]
    with np.errstate(all='ignore'):
        d1 = (y[:, 0] - y[:, 1]) / u0
        d2 = (y[:, 2] - y[:, 1]) / u2
        a = (d2 - d1) / (u2 - u0)
        b = d1 - a * u0
        ok = (u0 < 0) & (u2 > 0) & np.all(np.isfinite(y), axis=1) & np.isfinite(a) & (a != 0.0)
        x0[ok] = x[ok, 1] - b[ok] / (2 * a[ok])
        y0[ok] = y[ok, 1] - b[ok] ** 2 / (4 * a[ok])
        xerr[ok] = 1 / np.sqrt(np.abs(a[ok]))
    bad = ok & ((x0 <= x[:, 0]) | (x[:, 2] <= x0) | (y0 <= 0.) | (a < 0.0))
    zwarn[bad] |= ZW.BAD_MINFIT
    for k in np.where(~ok)[0]:
        x0[k], xerr[k], y0[k], zwarn[k] = _minfit_polyfit(x[k], y[k])
    return x0, xerr, y0, zwarn


def minfit(x, y):
    """Fit y = y0 + ((x-x0)/xerr)**2 on 3 points (fitz.py:70-111): the closed
    form of minfit_batch for 3 points, np.polyfit as in the reference else."""
    if len(x) == 3:
        r = minfit_batch(np.asarray(x)[None, :], np.asarray(y)[None, :])
        return (float(r[0][0]), float(r[1][0]), float(r[2][0]), int(r[3][0]))
    return _minfit_polyfit(x, y)


def _minfit_polyfit(x, y):
    """fitz.py:70-111 as written: least squares by np.polyfit."""
    if len(x) < 3:
        return (-1, -1, -1, ZW.BAD_MINFIT)
    try:
        a, b, c = np.polyfit(x, y, 2)
    except np.linalg.LinAlgError:
        return (-1, -1, -1, ZW.BAD_MINFIT)
    if a == 0.0:
        return (-1, -1, -1, ZW.BAD_MINFIT)
    x0 = -b / (2 * a)
    y0 = -(b ** 2) / (4 * a) + c
    zwarn = 0
    if (x0 <= np.min(x)) or (np.max(x) <= x0):
        zwarn |= ZW.BAD_MINFIT
    if y0 <= 0.:
        zwarn |= ZW.BAD_MINFIT
    if a > 0.0:
        xerr = 1 / np.sqrt(a)
    else:
        xerr = 1 / np.sqrt(-a)
        zwarn |= ZW.BAD_MINFIT
    return (x0, xerr, y0, zwarn)


class _Refiner(object):
    """GPU refinement of (target, candidate-minimum) pairs for one template."""

    def __init__(self, template, targets, redshifts, nz_fine, precision='f64', batch=None):
        from .zscan import TargetBatch, PixelLayout, _oiisum
        self.template = template
        self.targets = targets
        self.redshifts = np.asarray(redshifts, dtype=np.float64)
        self.nz_fine = nz_fine
        self.precision = precision
        self.dwave = {s.wavehash: s.wave for s in targets[0].spectra}
        from .rebin import centers2edges
        self.edges = [centers2edges(w) for w in self.dwave.values()]
        self.oii = _oiisum(template)
        nt = len(targets)
        if batch is not None and batch.nt == nt:
            # the batch of these targets (its constructor checked that they
            # share one layout): reuse it
            self.batches = {batch.layout.signature: list(range(nt))}
            self._tb = {batch.layout.signature: batch}
        else:
            self.batches = dict()
            for i, tg in enumerate(targets):
                self.batches.setdefault(PixelLayout.signature_of(tg.spectra), []).append(i)
            self._tb = {sig: TargetBatch([targets[i] for i in idx])
                        for sig, idx in self.batches.items()}
        # target -> (layout number, slot in that layout's batch)
        self.sigs = list(self.batches)
        self.tsig = np.zeros(nt, dtype=np.int64)
        self.tslot = np.zeros(nt, dtype=np.int64)
        for g, sig in enumerate(self.sigs):
            idx = np.asarray(self.batches[sig], dtype=np.int64)
            self.tsig[idx] = g
            self.tslot[idx] = np.arange(idx.size)

    def batch_of(self, targets):
        """The TargetBatch of all targets when they share one layout, else None."""
        if len(self._tb) == 1:
            tb = next(iter(self._tb.values()))
            if tb.nt == len(targets):
                return tb
        return None

    @staticmethod
    def _cache_budget():
        """HBM bytes for one fine-grid template cache (nbasis x padded pixels x
        redshifts doubles): an eighth of the free device memory, 1-32 GiB."""
        import torch
        free, _ = torch.cuda.mem_get_info()
        return int(min(max(free // 8, 1 << 30), 32 << 30))

    def _scan(self, zlist, owners, per_item):
        """chi2/coeff/penalty at zlist[j*per_item:(j+1)*per_item] for target owners[j]."""
        nitem = len(owners)
        nb = self.template.nbasis
        if nitem == 0:
            return np.zeros((0, per_item)), np.zeros((0, per_item, nb)), np.zeros((0, per_item))
        npix = sum(len(w) for w in self.dwave.values()) + 64
        chunk = max(1, int(self._cache_budget() // (nb * npix * 8 * per_item)))
        if nitem <= chunk:
            return self._scan_chunk(zlist, owners, per_item)
        parts = [self._scan_chunk(zlist[i * per_item:(i + chunk) * per_item], owners[i:i + chunk], per_item)
                 for i in range(0, nitem, chunk)]
        return tuple(np.concatenate([p[k] for p in parts]) for k in range(3))

    def _scan_blocks(self, zblocks, block, owners):
        """chi2/coeff/penalty of item j: target owners[j] at the redshifts
        zblocks[block[j]] (fine grids: candidates at the same coarse minimum
        share one grid).  The template cache holds each distinct grid once and
        items run ordered by grid, so items of one grid read the same cache
        rows one after another.  Falls back to _scan when the distinct grids
        exceed one cache, or the grids are not the packed 5-16 point kind."""
        nblk, per_item = zblocks.shape
        nitem = len(owners)
        nb = self.template.nbasis
        npix = sum(len(w) for w in self.dwave.values()) + 64
        fits = nblk <= max(1, int(self._cache_budget() // (nb * npix * 8 * per_item)))
        if nitem == 0 or not fits or not (4 < per_item <= 16) or self.precision != 'f64':
            return self._scan(zblocks[block].reshape(-1), owners, per_item)
        from .zscan import TemplateCache, ScanPlan
        chi2 = np.zeros((nitem, per_item))
        coeff = np.zeros((nitem, per_item, nb))
        pen = np.zeros((nitem, per_item))
        dw = dict()
        for tg in self.targets:
            for s in tg.spectra:
                dw.setdefault(s.wavehash, s.wave)
        cache = TemplateCache(self.template, dw, redshifts=zblocks.reshape(-1), check_range=False)
        owners = np.asarray(owners, dtype=np.int64)
        order = np.lexsort((np.arange(nitem), block))            # by grid, stable
        osig = self.tsig[owners[order]]
        for g, sig in enumerate(self.sigs):
            sel = order[osig == g]
            if sel.size == 0:
                continue
            slots = self.tslot[owners[sel]]
            plan = ScanPlan(cache, self._tb[sig], slots, block[sel] * per_item,
                            np.full(sel.size, per_item), self.precision, oiisum=self.oii,
                            solve=self.template.solve_matrices_algorithm)
            plan.run()
            c2, cf, pn, _ = plan.host()
            chi2[sel] = c2.reshape(sel.size, per_item)
            coeff[sel] = cf.reshape(sel.size, per_item, nb)
            pen[sel] = pn.reshape(sel.size, per_item)
        return chi2, coeff, pen

    def _scan_chunk(self, zlist, owners, per_item):
        from .zscan import TemplateCache, ScanPlan, target_runs
        nitem = len(owners)
        nb = self.template.nbasis
        chi2 = np.zeros((nitem, per_item))
        coeff = np.zeros((nitem, per_item, nb))
        pen = np.zeros((nitem, per_item))
        # candidates come from many targets: one cache for all their redshifts
        dw = dict()
        for tg in self.targets:
            for s in tg.spectra:
                dw.setdefault(s.wavehash, s.wave)
        cache = TemplateCache(self.template, dw, redshifts=np.asarray(zlist), check_range=False)
        owners = np.asarray(owners, dtype=np.int64)
        for g, sig in enumerate(self.sigs):
            sel = np.flatnonzero(self.tsig[owners] == g)
            if sel.size == 0:
                continue
            slots = self.tslot[owners[sel]]
            if sel.size == nitem and self.precision == 'f64' and (per_item > 16 or per_item <= 4):
                # consecutive candidates of one target share a scan item (their
                # fine grids are consecutive in zlist): more redshifts per
                # wavefront.  Grids of 5-16 points stay one item each: the FP64
                # kernel packs such items 4 to a wavefront
                tr, zr, nr = target_runs(np.repeat(slots, per_item))
                plan = ScanPlan(cache, self._tb[sig], tr, zr, nr, self.precision, oiisum=self.oii,
                                solve=self.template.solve_matrices_algorithm)
            else:
                plan = ScanPlan(cache, self._tb[sig], slots, sel * per_item,
                                np.full(sel.size, per_item), self.precision, oiisum=self.oii,
                                solve=self.template.solve_matrices_algorithm)
            plan.run()
            c2, cf, pn, _ = plan.host()
            chi2[sel] = c2.reshape(sel.size, per_item)
            coeff[sel] = cf.reshape(sel.size, per_item, nb)
            pen[sel] = pn.reshape(sel.size, per_item)
        return chi2, coeff, pen

    def _edges_ok_many(self, zlo, zhi):
        """rebin.py:332-341 as rebin_template applies it with the xmin / xmax
        fitz passes (the template's wavelength range at zhi / zlo), for
        arrays of (zlo, zhi): True where every camera's edges lie inside."""
        xmin = self.template.minwave * (1 + np.asarray(zhi))
        xmax = self.template.maxwave * (1 + np.asarray(zlo))
        ok = np.ones(xmin.shape, dtype=bool)
        for e in self.edges:
            ok &= ~((e[0] < xmin) | (e[-1] > xmax))
        return ok

    def refine(self, pt, pimin):
        """Refine the candidates (target pt[j], coarse minimum pimin[j]).

        Returns a dict of arrays over the candidates: zz, zzchi2 (n, nz_fine)
        (the fine grid, chi2 + penalty on it), zmin, sigma, chi2min, zwarn
        (the parabola through the 3 points about the fine-grid minimum,
        fitz.py:226-263), coeff (n, nbasis) and zigm (the coefficients at zmin
        and the redshift of their IGM transmission, fitz.py:265-304), and err:
        the candidate's fine grid, or its zmin inside the redshift range,
        leaves the template's wavelength range — the reference raises
        ValueError('edges must be within input x range') there (fitz.py:226-
        231, rebin.py:332-341), if its control flow reaches the candidate.
        """
        nz = self.nz_fine
        zs = self.redshifts
        pt = np.asarray(pt, dtype=np.int64).reshape(-1)
        pimin = np.asarray(pimin, dtype=np.int64).reshape(-1)
        n = pt.size
        nb = self.template.nbasis
        lo = zs[np.maximum(pimin - 1, 0)]
        hi = zs[np.minimum(pimin + 1, len(zs) - 1)]
        # row by row what np.linspace(zs[ilo], zs[ihi], nz) gives (fitz.py:226-231)
        zz_rows = np.linspace(lo, hi, nz, axis=1) if n else np.zeros((0, nz))
        res = dict(zz=zz_rows, zzchi2=np.zeros((n, nz)), zmin=np.zeros(n), sigma=np.zeros(n),
                   chi2min=np.zeros(n), zwarn=np.zeros(n, dtype=np.int64), coeff=np.zeros((n, nb)),
                   zigm=np.full(n, np.nan), err=np.zeros(n, dtype=bool))
        if n == 0:
            return res
        ok = self._edges_ok_many(np.minimum(lo, hi), np.maximum(lo, hi))
        res['err'][~ok] = True
        good = np.flatnonzero(ok)
        if good.size == 0:
            return res
        # candidates at the same coarse minimum have the same fine grid
        _, first, inv = np.unique(pimin[good], return_index=True, return_inverse=True)
        zzchi2, _, zzpen = self._scan_blocks(zz_rows[good][first], inv.reshape(-1), pt[good])
        c2 = zzchi2[:good.size].reshape(good.size, nz).copy()
        if self.oii is not None:
            c2 += zzpen[:good.size].reshape(good.size, nz)
        res['zzchi2'][good] = c2
        # the 3 points about each fine-grid minimum (fitz.py:233-250)
        i = np.clip(np.argmin(c2, axis=1), 1, nz - 2)
        r3 = np.arange(good.size)[:, None]
        c3 = i[:, None] + np.arange(-1, 2)[None, :]
        zmin, sigma, chi2min, zwarn = minfit_batch(zz_rows[good][r3, c3], c2[r3, c3])
        res['zmin'][good], res['sigma'][good] = zmin, sigma
        res['chi2min'][good], res['zwarn'][good] = chi2min, zwarn
        # coefficients at each parabola minimum (fitz.py:265-304)
        eok = self._edges_ok_many(zmin, zmin)
        sel = good[eok]
        _, cf, _ = self._scan(res['zmin'][sel], pt[sel], 1)
        res['coeff'][sel] = cf[:, 0]
        res['zigm'][sel] = res['zmin'][sel]         # transmission at zmin (fitz.py:266-284)
        rest = good[~eok]
        beyond = (res['zmin'][rest] < zs[0]) | (zs[-1] < res['zmin'][rest])
        # beyond the redshift range: zero coefficients (fitz.py:295-301), no transmission
        res['zwarn'][rest[beyond]] |= ZW.Z_FITLIMIT | ZW.BAD_MINFIT
        res['err'][rest[~beyond]] = True
        return res


def select_minima(zchi2, redshifts, nminima, max_velo_diff, refine, speculate=1):
    """The candidate loop of fitz.py:198-327 for every row of zchi2 at once.

    Per row (target) t the reference walks find_minima(zchi2[t]) in order:
    stop at nminima accepted; skip a minimum within max_velo_diff of an
    accepted redshift; else refine it (fine grid, parabola), flag the
    parabola vertex outside the fine grid (BAD_MINFIT, the grid's lowest
    point instead) or near the scan's ends (Z_FITLIMIT), and skip it if the
    refined redshift is within max_velo_diff of an accepted one.  Here one
    numpy step takes the next candidate of every target still choosing.
    refine(pt, pimin) returns the arrays of _Refiner.refine for candidates
    (pt[j], pimin[j]); the first nminima + speculate candidates of every
    target are refined up front, and a target that reaches an unrefined
    candidate waits: the next nminima + speculate candidates of all waiting
    targets are refined together and they resume where they stopped.

    Returns (nacc, acc): nacc[t] accepted minima of target t, and acc a dict
    of (nt, nminima) arrays in acceptance order — z, zerr, zwarn, chi2 and
    row, the row of the candidate in store, the concatenation of every
    refine() result (a dict of arrays, returned as acc['store']).
    """
    zchi2 = np.atleast_2d(np.asarray(zchi2, dtype=np.float64))
    zs = np.asarray(redshifts, dtype=np.float64)
    nt = zchi2.shape[0]
    W = max(1, nminima + speculate)
    head, ncand = minima_head(zchi2, W)
    # candidate positions -> coarse index (listed up to filled[t]), and ->
    # store row once refined
    width = int(min(W, ncand.max())) if nt else 0
    cmat = np.full((nt, max(width, 1)), -1, dtype=np.int64)
    cmat[:, :width] = head[:, :width]
    rowmat = np.full(cmat.shape, -1, dtype=np.int64)
    filled = np.minimum(ncand, width)

    def fill(ts, upto):
        nonlocal cmat, rowmat
        upto = np.minimum(upto, ncand[ts])
        if upto.size and int(upto.max()) > cmat.shape[1]:
            grow = int(upto.max()) - cmat.shape[1] + W
            cmat = np.pad(cmat, ((0, 0), (0, grow)), constant_values=-1)
            rowmat = np.pad(rowmat, ((0, 0), (0, grow)), constant_values=-1)
        more = upto > filled[ts]
        ts, upto = ts[more], upto[more]
        if ts.size == 0:
            return
        h, _ = minima_head(zchi2[ts], int(upto.max()))
        i, k = _ranges(filled[ts], upto - filled[ts])
        cmat[ts[i], k] = h[i, k]
        filled[ts] = upto

    parts, nstore = [], 0
    kpos = np.zeros(nt, dtype=np.int64)
    nacc = np.zeros(nt, dtype=np.int64)
    acc = dict(z=np.full((nt, max(nminima, 1)), np.nan), zerr=np.zeros((nt, max(nminima, 1))),
               zwarn=np.zeros((nt, max(nminima, 1)), dtype=np.int64),
               chi2=np.zeros((nt, max(nminima, 1))), row=np.full((nt, max(nminima, 1)), -1))
    lo_lim, hi_lim = zs[min(1, zs.size - 1)], zs[max(zs.size - 2, 0)]

    # first request: the first W candidates of every target
    tt, kk = _ranges(np.zeros(nt, dtype=np.int64), np.minimum(ncand, W))
    store = None
    while True:
        if tt.size:
            res = refine(tt, cmat[tt, kk])
            rowmat[tt, kk] = nstore + np.arange(tt.size)
            nstore += tt.size
            parts.append(res)
            store = {k: np.concatenate([p[k] for p in parts]) for k in parts[0]}
            parts = [store]
        blocked = np.zeros(nt, dtype=bool)
        while True:
            act = np.flatnonzero((nacc < nminima) & (kpos < ncand) & ~blocked)
            if act.size == 0:
                break
            k = kpos[act]
            short = k >= filled[act]
            if np.any(short):
                fill(act[short], k[short] + W)
            imin = cmat[act, k]
            zprev = acc['z'][act]                 # NaN past the accepted ones: never close
            with np.errstate(invalid='ignore'):
                skip1 = np.any(np.abs(get_dv(zs[imin][:, None], zprev)) < max_velo_diff, axis=1)
            r = rowmat[act, k]
            miss = ~skip1 & (r < 0)
            blocked[act[miss]] = True
            go = ~skip1 & ~miss
            ga, gr = act[go], r[go]
            if gr.size:
                if np.any(store['err'][gr]):
                    raise ValueError('edges must be within input x range')
                zmin = store['zmin'][gr]
                zwarn = store['zwarn'][gr].copy()
                chi2min = store['chi2min'][gr].copy()
                zz = store['zz'][gr]
                zbest = zmin.copy()
                out = (zbest < zz[:, 0]) | (zbest > zz[:, -1])
                if np.any(out):
                    zwarn[out] |= ZW.BAD_MINFIT
                    zzc = store['zzchi2'][gr[out]]
                    i = np.argmin(zzc, axis=1)
                    zbest[out] = zz[out, i]
                    chi2min[out] = zzc[np.arange(i.size), i]
                lim = (zbest < lo_lim) | (zbest > hi_lim) | (zmin < lo_lim) | (zmin > hi_lim)
                zwarn[lim] |= ZW.Z_FITLIMIT
                with np.errstate(invalid='ignore'):
                    skip2 = np.any(np.abs(get_dv(zbest[:, None], acc['z'][ga])) < max_velo_diff,
                                   axis=1)
                a = ~skip2
                at, slot = ga[a], nacc[ga[a]]
                acc['z'][at, slot] = zbest[a]
                acc['zerr'][at, slot] = store['sigma'][gr[a]]
                acc['zwarn'][at, slot] = zwarn[a]
                acc['chi2'][at, slot] = chi2min[a]
                acc['row'][at, slot] = gr[a]
                nacc[at] += 1
            kpos[act[~miss]] += 1
        wait = np.flatnonzero(blocked)
        if wait.size == 0:
            break
        # the next W candidates of every waiting target
        need = np.minimum(kpos[wait] + W, ncand[wait])
        fill(wait, need)
        i, kk = _ranges(kpos[wait], need - kpos[wait])
        tt = wait[i]
        unref = rowmat[tt, kk] < 0
        tt, kk = tt[unref], kk[unref]
    acc['store'] = store
    return nacc, acc


def _stack(results, target, npix=None):
    """Sort candidates by chi2 and stack them into the reference's dict of arrays
    (fitz.py:359-395); npix: the target's good-pixel count if known."""
    ii = np.argsort([tmp['chi2'] for tmp in results])
    results = [results[i] for i in ii]
    for r in results:
        r.pop('_zigm', None)
    if len(results) == 0:
        xfloat = np.zeros((0, 1), dtype=float)
        xint = np.zeros((0, 1), dtype=int)
        xstr = np.zeros((0, 1), dtype=str)
        return dict(z=xfloat, zerr=xfloat, zwarn=xint, chi2=xfloat, zz=xfloat, zzchi2=xfloat,
                    coeff=xfloat, fitmethod=xstr, npixels=xint)
    if npix is None:
        npix = sum(int((s.ivar > 0.).sum()) for s in target.spectra)
    for r in results:
        r['npixels'] = npix
    return {k: _vstack([r[k] for r in results]) for k in results[0].keys()}


_SCALARS = (float, int, str, np.floating, np.integer, np.str_)


def _vstack(vals):
    """np.vstack of one result field over the candidates: scalars -> (n, 1),
    equal-length 1-D arrays -> (n, m), with np.vstack's dtype; the common
    cases skip its per-element atleast_2d."""
    v0 = vals[0]
    if isinstance(v0, _SCALARS) and all(isinstance(v, _SCALARS) for v in vals):
        return np.array(vals).reshape(len(vals), 1)
    if isinstance(v0, np.ndarray) and v0.ndim == 1 and all(
            isinstance(v, np.ndarray) and v.shape == v0.shape for v in vals):
        return np.array(vals)
    return np.vstack(vals)


def fitz_targets(zchi2, redshifts, targets, template, nminima=3, zminfit_npoints=15,
                 precision='f64', speculate=1, newton=False, archetype=None, deg_legendre=None,
                 per_camera=False, n_nearest=None, prior_sigma=None, batch=None, npixels=None):
    """fitz for every target of one template (zchi2: (nt, nz) incl. penalty).

    Returns a list with the reference's per-target result dict.  With
    newton=True every minimum is also refined by Newton iterations on the
    analytic d chi2/dz inside its fine grid (redrock_amd.grad.newton_refine)
    and the dicts gain 'znewton' and 'chi2newton' beside the parabola's 'z'
    and 'chi2'.  With an archetype (redrock_amd.archetypes.Archetype) every
    minimum's chi2 and coefficients come from the archetype fit at its
    redshift (fitz.py:329-357): deg_legendre Legendre polynomials (per
    camera with per_camera), prior_sigma on their coefficients, n_nearest
    archetypes combined.  batch: the TargetBatch of `targets` (one layout),
    reused instead of uploading them again.  npixels: each target's number of
    pixels with ivar > 0, when the caller has it already.
    """
    out = _fitz_targets(zchi2, redshifts, targets, template, nminima, zminfit_npoints,
                        precision, speculate, archetype, deg_legendre, per_camera, n_nearest,
                        prior_sigma, batch, npixels)
    if newton:
        _add_newton(out, targets, template)
    return out


def _add_newton(out, targets, template):
    from .grad import GradProblem, newton_refine
    owner, z0, lo, hi = [], [], [], []
    for t, r in enumerate(out):
        for k in range(r['z'].shape[0]):
            owner.append((t, k))
            z0.append(r['z'][k, 0])
            lo.append(r['zz'][k, 0])
            hi.append(r['zz'][k, -1])
    for r in out:
        r['znewton'] = np.zeros_like(r['z'])
        r['chi2newton'] = np.zeros_like(r['chi2'])
    if not owner:
        return
    prob = GradProblem(template, targets, [t for t, _ in owner])
    zn, c2, _ = newton_refine(prob, np.array(z0), np.array(lo), np.array(hi))
    for (t, k), zv, cv in zip(owner, zn, c2):
        out[t]['znewton'][k, 0] = zv
        out[t]['chi2newton'][k, 0] = cv


def _fitz_targets(zchi2, redshifts, targets, template, nminima, zminfit_npoints, precision,
                  speculate, archetype=None, deg_legendre=None, per_camera=False, n_nearest=None,
                  prior_sigma=None, batch=None, npixels=None):
    if archetype is None:
        return _as_dicts(*fitz_columns(zchi2, redshifts, targets, template, nminima, zminfit_npoints,
                                       precision, speculate, batch, npixels))
    zchi2 = np.atleast_2d(np.asarray(zchi2, dtype=np.float64))
    nt = len(targets)
    assert zchi2.shape == (nt, len(redshifts))
    nz = 15 if zminfit_npoints is None else zminfit_npoints
    ref = _Refiner(template, targets, redshifts, nz, precision, batch=batch)
    mvd = (constants.max_velo_diff_star if template.template_type == 'STAR'
           else constants.max_velo_diff)
    nacc, acc = select_minima(zchi2, ref.redshifts, nminima, mvd, ref.refine, speculate)
    st = acc['store']
    if npixels is None:
        npixels = [sum(int((s.ivar > 0.).sum()) for s in tg.spectra) for tg in targets]
    # per-candidate dicts in acceptance order, refit by the archetypes, sorted
    out = [[dict(z=acc['z'][t, k], zerr=acc['zerr'][t, k], zwarn=int(acc['zwarn'][t, k]),
                 chi2=acc['chi2'][t, k], zz=st['zz'][acc['row'][t, k]],
                 zzchi2=st['zzchi2'][acc['row'][t, k]], coeff=st['coeff'][acc['row'][t, k]],
                 fitmethod=template.method, _zigm=st['zigm'][acc['row'][t, k]])
            for k in range(nacc[t])] for t in range(nt)]
    _archetype_pass(out, targets, template, archetype, ref, deg_legendre, per_camera, n_nearest,
                    prior_sigma)
    return [_stack(r, tg, n) for r, tg, n in zip(out, targets, npixels)]


def _columns(acc, nacc, st, template, npixels, nz):
    """Every target's accepted minima sorted by chi2 (fitz.py:359-362) as flat
    columns, target by target: z, zerr, zwarn, chi2, fitmethod, npixels (n,),
    zz, zzchi2 (n, nz), coeff (n, nbasis)."""
    nm = acc['z'].shape[1]
    valid = np.arange(nm)[None, :] < nacc[:, None]
    o = np.argsort(np.where(valid, acc['chi2'], np.inf), axis=1, kind='stable')
    sel = np.take_along_axis(valid, o, axis=1)
    col = {k: np.take_along_axis(acc[k], o, axis=1)[sel] for k in ('z', 'zerr', 'zwarn', 'chi2', 'row')}
    rows = col.pop('row')
    ntot = rows.size
    nbasis = template.nbasis
    cols = dict(z=col['z'], zerr=col['zerr'], zwarn=col['zwarn'], chi2=col['chi2'],
                zz=st['zz'][rows] if ntot else np.zeros((0, nz)),
                zzchi2=st['zzchi2'][rows] if ntot else np.zeros((0, nz)),
                coeff=st['coeff'][rows] if ntot else np.zeros((0, nbasis)),
                fitmethod=np.full(ntot, template.method),
                npixels=np.repeat(np.asarray(npixels, dtype=np.int64), nacc))
    return nacc, cols


def _as_dicts(nacc, cols):
    """Per-target result dicts (fitz.py:359-395) viewing the flat columns."""
    ntot = cols['z'].size
    c2 = {k: (v.reshape(ntot, 1) if v.ndim == 1 else v) for k, v in cols.items()}
    bounds = np.concatenate([[0], np.cumsum(nacc)])
    xfloat = np.zeros((0, 1), dtype=float)
    xint = np.zeros((0, 1), dtype=int)
    xstr = np.zeros((0, 1), dtype=str)
    empty = dict(z=xfloat, zerr=xfloat, zwarn=xint, chi2=xfloat, zz=xfloat, zzchi2=xfloat,
                 coeff=xfloat, fitmethod=xstr, npixels=xint)
    keys = ('z', 'zerr', 'zwarn', 'chi2', 'zz', 'zzchi2', 'coeff', 'fitmethod', 'npixels')
    res = []
    for t in range(len(nacc)):
        a, b = bounds[t], bounds[t + 1]
        res.append(dict(empty) if a == b else {k: c2[k][a:b] for k in keys})
    return res


def fitz_columns(zchi2, redshifts, targets, template, nminima=3, zminfit_npoints=15,
                 precision='f64', speculate=1, batch=None, npixels=None):
    """fitz_targets without archetypes, as flat columns: (nacc, cols) with
    nacc[t] the number of minima of target t and cols the dict of their
    z, zerr, zwarn, chi2, fitmethod, npixels (n,), zz, zzchi2 (n, nz_fine)
    and coeff (n, nbasis), target by target, each target's rows sorted by
    chi2 — the per-target dicts of fitz_targets stacked (zfind builds its
    table from them)."""
    zchi2 = np.atleast_2d(np.asarray(zchi2, dtype=np.float64))
    nt = len(targets)
    assert zchi2.shape == (nt, len(redshifts))
    nz = 15 if zminfit_npoints is None else zminfit_npoints
    ref = _Refiner(template, targets, redshifts, nz, precision, batch=batch)
    mvd = (constants.max_velo_diff_star if template.template_type == 'STAR'
           else constants.max_velo_diff)
    nacc, acc = select_minima(zchi2, ref.redshifts, nminima, mvd, ref.refine, speculate)
    if npixels is None:
        npixels = [sum(int((s.ivar > 0.).sum()) for s in tg.spectra) for tg in targets]
    return _columns(acc, nacc, acc['store'], template, npixels, nz)


def _archetype_pass(rows, targets, template, archetype, ref, deg_legendre, per_camera, n_nearest,
                    prior_sigma):
    """Replace every candidate's fit by the archetype fit at its redshift
    (fitz.py:333-357), all candidates of all targets in one ArchetypeFit."""
    from .archetypes import ArchetypeFit, prior_on_coeffs, split_archetype_coeff
    from .igm import lyman_table
    from .templates import parse_fulltype
    nleg = 0 if deg_legendre is None else int(deg_legendre)
    ncamera = len({s.wavehash for s in targets[0].spectra}) if per_camera else 1
    prior = None
    if prior_sigma is not None:
        prior = prior_on_coeffs(1 if n_nearest is None else n_nearest, nleg, prior_sigma, ncamera)
    owner = [(t, k) for t, r in enumerate(rows) for k in range(len(r))]
    if not owner:
        return
    for tg in targets:
        tg.legendre(nleg)
    fit = ArchetypeFit(archetype, targets, per_camera=per_camera, nleg=nleg, prior=prior,
                       n_nearest=n_nearest, ncamera=ncamera, batch=ref.batch_of(targets))
    tidx = np.array([t for t, _ in owner])
    zb = np.array([rows[t][k]['z'] for t, k in owner], dtype=np.float64)
    zigm = np.array([rows[t][k]['_zigm'] for t, k in owner], dtype=np.float64)
    lines = lyman_table(template.igm_model) or []
    chi2, coeff, ftype = fit.run(tidx, zb, zigm=zigm, lines=lines)
    for (t, k), c2, cf, ft in zip(owner, chi2, coeff, ftype):
        r = rows[t][k]
        archcoeff, legcoeff = split_archetype_coeff(parse_fulltype(ft)[1], cf, ncamera, nleg)
        if per_camera:
            legcoeff = [np.vstack(legcoeff)]
        r.update(chi2=c2, pca_coeff=r['coeff'], coeff=archcoeff, legcoeff=legcoeff, fulltype=ft,
                 fitmethod=archetype.method, pca_spectype=template._rrtype,
                 pca_subtype=template.sub_type)


def fitz(zchi2, redshifts, target, template, nminima=3, archetype=None, use_gpu=False,
         deg_legendre=None, zminfit_npoints=15, per_camera=False, n_nearest=None,
         prior_sigma=None):
    """Refine the redshift around up to nminima minima (fitz.py:133-395)."""
    assert len(zchi2) == len(redshifts)
    return fitz_targets(np.asarray(zchi2)[None, :], redshifts, [target], template,
                        nminima=nminima, zminfit_npoints=zminfit_npoints, archetype=archetype,
                        deg_legendre=deg_legendre, per_camera=per_camera, n_nearest=n_nearest,
                        prior_sigma=prior_sigma)[0]
