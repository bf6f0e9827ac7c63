"""FP64 polish of a mixed-precision scan (rr_zscan_polish).

After a mixed-precision scan the rows near each item's minimum are
re-evaluated in FP64.  Checked here against the plain FP64 GPU scan (itself
held to the oracle at 1e-9 by test_gpu_zscan.py):

* the polished scan's best redshift equals the FP64 scan's for every target,
  also in bf16, whose unpolished scan misses some (the fine STAR grid and
  bright galaxies are the hard cases);
* polished rows agree with the FP64 rows to 1e-12 relative (same arithmetic,
  pixel sums in another order); the other rows are the unpolished scan's, bit
  for bit;
* the windows chosen equal a CPU restatement of the selection rule, written
  below from the rule's statement (window minima of chi2 and chi2 + penalty,
  threshold minimum + tol max(f'Wf, |minimum|), halved until at most
  max_windows windows qualify, no window for an item with a single row under
  the full threshold);
* a target with no weight (HUGE_CHI2 everywhere) is left alone.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HUGE = 9999999.0
WZ = 4


def select_windows(chi2, pen, status, fwf, sumw, tol, maxw):
    nz = chi2.size
    nw = (nz + WZ - 1) // WZ
    ok = status == 0                 # singular rows (HUGE_CHI2 by convention) take no part
    pad = np.full(nw * WZ, np.inf)
    c = pad.copy()
    c[:nz] = np.where(ok, chi2, np.inf)
    s = pad.copy()
    s[:nz] = np.where(ok, chi2 + pen, np.inf)
    w1 = c.reshape(nw, WZ).min(axis=1)
    w2 = s.reshape(nw, WZ).min(axis=1)
    m1, m2 = w1.min(), w2.min()
    if not (sumw > 0 and m1 < np.inf):  # no weight: HUGE_CHI2 everywhere
        return np.zeros(0, dtype=np.int64)
    d1, d2 = tol * max(fwf, abs(m1)), tol * max(fwf, abs(m2))
    it = 0
    while True:
        t1, t2 = m1 + d1, m2 + d2
        sel = (w1 <= t1) | (w2 <= t2)
        if sel.sum() <= maxw or (d1 == 0.0 and d2 == 0.0):
            break
        if it >= 60:
            d1 = d2 = 0.0
        else:
            d1 *= 0.5
            d2 *= 0.5
        it += 1
    if it == 0 and np.sum(ok & ((chi2 <= t1) | (chi2 + pen <= t2))) <= 1:
        return np.zeros(0, dtype=np.int64)      # isolated minimum
    return np.flatnonzero(sel)[:maxw]


def fwf_of(plan, k=0):
    """f'Wf (k = 0) or sum(w) (k = 1) of every item, as the scan's statistics
    stage left them in the workspace."""
    import torch
    return plan.work[:16 * plan.nitems].view(torch.float64).cpu().numpy()[k::2]


def _plans(tm, targets, precision, polish, windows=None):
    import torch
    from redrock_amd import zscan
    from redrock_amd.templates import DistTemplate
    windows = windows or zscan.POLISH_WINDOWS
    dwave = {s.wavehash: s.wave for s in targets[0].spectra}
    dt = DistTemplate(tm, dwave)
    batch = zscan.TargetBatch(targets)
    nt, nz = len(targets), dt.cache.nz
    oii = zscan._oiisum(tm)
    items = (np.arange(nt), np.zeros(nt), np.full(nt, nz))
    f64 = zscan.ScanPlan(dt.cache, batch, *items, precision='f64', oiisum=oii).run()
    raw = zscan.ScanPlan(dt.cache, batch, *items, precision=precision, oiisum=oii,
                         polish=False).run()
    pol = zscan.ScanPlan(dt.cache, batch, *items, precision=precision, oiisum=oii,
                         polish=polish, polish_windows=windows).run()
    torch.cuda.synchronize()
    return f64, raw, pol, nt, nz


def _check(f64, raw, pol, nt, nz, tol, windows, dead=()):
    c64, f64c, p64, _ = f64.host()
    craw, rawc, praw, sraw = raw.host()
    cpol, polc, ppol, spol = pol.host()
    c64, craw, cpol, sraw = (a.reshape(nt, nz) for a in (c64, craw, cpol, sraw))
    praw, ppol, p64 = (a.reshape(nt, nz) for a in (praw, ppol, p64))
    nwin = pol.nwindows.cpu().numpy()
    win = pol.windows.cpu().numpy()
    fwf, sumw = fwf_of(pol), fwf_of(pol, 1)
    nrows = 0
    for i in range(nt):
        want = select_windows(craw[i], praw[i], sraw[i], fwf[i], sumw[i], tol, windows)
        assert nwin[i] == want.size, i
        assert np.array_equal(win[i, :nwin[i]], want), i
        mask = np.zeros(nz, dtype=bool)
        for k in want:
            mask[k * WZ:(k + 1) * WZ] = True
        nrows += mask.sum()
        assert np.array_equal(cpol[i, ~mask], craw[i, ~mask]), i
        assert np.array_equal(ppol[i, ~mask], praw[i, ~mask]), i
        if mask.any():
            rel = np.abs(cpol[i, mask] - c64[i, mask]) / np.abs(c64[i, mask])
            assert rel.max() <= 1e-12, (i, rel.max())
            assert np.all(spol.reshape(nt, nz)[i, mask] == 0)
        if i in dead:
            assert nwin[i] == 0
            assert np.all(cpol[i] == HUGE)
    return c64, craw, cpol, nrows / (nt * nz)


@pytest.mark.parametrize('precision,snr', [('f16x3', 1.0), ('bf16', 1.0), ('f16x3', 10.0),
                                           ('f16', 30.0)])
def test_polish_galaxy_best_z(gpu, precision, snr):
    """Bright targets (per-pixel S/N 10 and 30) are where the mixed chi2 error,
    relative to chi2, is largest (it scales with f'Wf)."""
    from redrock_amd import synth, zscan
    tm = synth.galaxy_template()
    targets, _ = synth.make_targets(tm, 48, seed=21, snr=snr)
    for s in targets[5].spectra:            # a target with no weight at all
        s.ivar[:] = 0.0
    tol = zscan.POLISH_TOL[precision]
    print(f'case: polish galaxy {precision}, S/N {snr}, 48 targets x {tm.redshifts.size} z, '
          f'tol {tol}', flush=True)
    f64, raw, pol, nt, nz = _plans(tm, targets, precision, None)
    c64, craw, cpol, frac = _check(f64, raw, pol, nt, nz, tol, zscan.POLISH_WINDOWS, dead=(5,))
    live = np.array([i for i in range(nt) if i != 5])
    agree_raw = np.mean(craw[live].argmin(axis=1) == c64[live].argmin(axis=1))
    print(f'{precision}: polished fraction of rows {frac:.4f}; best-z agreement '
          f'unpolished {agree_raw:.4f}')
    assert np.array_equal(cpol[live].argmin(axis=1), c64[live].argmin(axis=1))
    best = c64[live].argmin(axis=1)
    rel = np.abs(cpol[live, best] - c64[live, best]) / c64[live, best]
    did = pol.nwindows.cpu().numpy()[live] > 0
    print(f'{precision}: {did.sum()} of {live.size} targets polished, the rest isolated')
    assert rel[did].max(initial=0.0) <= 1e-12
    err_f = np.abs(cpol[live, best] - c64[live, best]) / fwf_of(pol)[live]
    assert err_f.max() <= tol
    if snr >= 30:
        # bright targets: real chi2 minima above HUGE_CHI2 are polished like any
        # other (a wide tolerance puts several rows of every target in range)
        big = c64[live].min(axis=1) > HUGE
        assert big.any()
        _, _, wide, _, _ = _plans(tm, targets, precision, 0.05)
        assert np.all(wide.nwindows.cpu().numpy()[live][big] > 0)
        assert np.array_equal(wide.zchi2.cpu().numpy().reshape(nt, nz)[live].argmin(axis=1),
                              c64[live].argmin(axis=1))


def test_polish_star_fine_grid(gpu):
    """STAR: 501 redshifts 2e-5 apart, where neighbouring chi2 are closest."""
    from redrock_amd import synth, zscan
    tm = synth.star_template()
    targets, _ = synth.make_targets(tm, 64, seed=4)
    print(f'case: polish star f16x3 and f16, 64 targets x {tm.redshifts.size} z', flush=True)
    for precision in ('f16x3', 'f16'):
        f64, raw, pol, nt, nz = _plans(tm, targets, precision, None)
        tol = zscan.POLISH_TOL[precision]
        c64, craw, cpol, frac = _check(f64, raw, pol, nt, nz, tol, zscan.POLISH_WINDOWS)
        print(f'{precision}: polished fraction {frac:.4f}; unpolished agreement '
              f'{np.mean(craw.argmin(axis=1) == c64.argmin(axis=1)):.4f}')
        assert np.array_equal(cpol.argmin(axis=1), c64.argmin(axis=1))


def test_polish_window_cap(gpu):
    """A tolerance so wide that every window qualifies: the cap keeps the
    windows nearest the minimum (the rule above), here at most 2."""
    from redrock_amd import synth
    tm = synth.qso_template()
    targets, _ = synth.make_targets(tm, 8, seed=9)
    print('case: polish qso f16, tol 10, 2 windows', flush=True)
    f64, raw, pol, nt, nz = _plans(tm, targets, 'f16', 10.0, windows=2)
    _check(f64, raw, pol, nt, nz, 10.0, 2)
    n = pol.nwindows.cpu().numpy()
    assert np.all((n >= 1) & (n <= 2))

