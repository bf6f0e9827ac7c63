"""Pins the CPU oracle (oracle/rr_oracle.py) before it is used to check the GPU.

Known answers come from the reference's own tests and docstrings (cited per
test) and from mathematics that fixes the result independently of the
restatement (closed-form integrals, an independent least-squares route,
exact parabolas).  Two tests plant a mistake in the oracle and show the checks
catch it.
"""
import numpy as np
import pytest
import scipy.sparse

from oracle import rr_oracle as orc


# ---------------------------------------------------------------- rebin
def test_centers2edges_reference_vectors():
    # test_rebin.py:28-32
    assert np.allclose(orc.centers2edges([1, 2, 3]), [0.5, 1.5, 2.5, 3.5])
    assert np.allclose(orc.centers2edges([1, 3, 5]), [0, 2, 4, 6])
    assert np.allclose(orc.centers2edges([1, 3, 4]), [0, 2, 3.5, 4.5])


def test_trapz_constant_density_exact():
    # test_rebin.py:34-51: constant density rebins to exactly 1.0
    nx = 10
    x = np.arange(nx) * 1.1
    y = np.ones(nx)
    for nedge in range(3, 10):
        edges = np.linspace(min(x), max(x), nedge)
        yy = orc.trapz_rebin(x, y, edges=edges)
        assert np.all(yy == 1.0), yy
    for nedge in range(3, 3 * nx):
        edges = np.linspace(0.5, 8.3, nedge)
        assert np.allclose(orc.trapz_rebin(x, y, edges=edges), 1.0)


def test_trapz_centers_input():
    # test_rebin.py:53-60
    x = np.arange(10)
    yy = orc.trapz_rebin(x, np.ones(10), np.linspace(0.5, 10 - 1.5))
    assert np.allclose(yy, 1.0)


def test_trapz_range_errors():
    # test_rebin.py:62-70
    x = np.arange(10)
    y = np.ones(10)
    with pytest.raises(ValueError):
        orc.trapz_rebin(x, y, edges=np.arange(-1, 9))
    with pytest.raises(ValueError):
        orc.trapz_rebin(x, y, edges=np.arange(1, 11))


def test_trapz_sin_integrals():
    # test_rebin.py:72-86
    for nx in range(5, 12):
        x = np.linspace(0, 2 * np.pi, nx)
        assert np.allclose(orc.trapz_rebin(x, np.sin(x), edges=[0, 2 * np.pi]), 0.0)
    x = np.linspace(0, 2 * np.pi, 100)
    yy = orc.trapz_rebin(x, np.sin(x), edges=[0, 0.5 * np.pi, np.pi, 1.5 * np.pi, 2 * np.pi])
    assert np.allclose(yy[0:2], 2 / np.pi, atol=5e-4)
    assert np.allclose(yy[2:4], -2 / np.pi, atol=5e-4)


def test_trapz_docstring_vectors():
    # rebin.py:222-250
    x = np.arange(10)
    assert np.allclose(orc.trapz_rebin(x, np.ones(10), edges=[0, 2, 4, 6, 8]), 1.0)
    y = np.ones((2, 10))
    y[1, :] = np.arange(10)
    r = orc.trapz_rebin(x, y, edges=[0, 2, 4, 6, 8])
    assert r.shape == (4, 2)
    assert np.allclose(r, [[1, 1], [1, 3], [1, 5], [1, 7]])
    r = orc.trapz_rebin(x, y, edges=[0, 2, 4, 6, 8], myz=[0, 0.5])
    assert r.shape == (2, 4, 2)
    assert np.allclose(r[1], [[1, 0.66666667], [1, 2], [1, 3.33333333], [1, 4.66666667]])


def _exact_bin_means(x, y, edges):
    """Independent exact integral of the piecewise-linear interpolant.

    Splits every bin at the knots inside it and integrates each linear piece
    with the midpoint rule (exact for linear functions).
    """
    out = []
    for a, b in zip(edges[:-1], edges[1:]):
        pts = np.concatenate([[a], x[(x > a) & (x < b)], [b]])
        mids = 0.5 * (pts[1:] + pts[:-1])
        out.append(np.sum(np.interp(mids, x, y) * np.diff(pts)) / (b - a))
    return np.array(out)


@pytest.mark.parametrize('seed', range(5))
def test_trapz_matches_exact_integral(seed):
    rng = np.random.default_rng(seed)
    x = np.cumsum(rng.uniform(0.2, 2.0, 60))
    y = rng.normal(size=(3, 60))
    for z in (0.0, 0.37, 1.9):
        xs = x * (1 + z)
        edges = np.sort(rng.uniform(xs[2], xs[-3], 25))
        r = orc.trapz_rebin(x, y, edges=edges, myz=[z])[0]
        for b in range(3):
            assert np.allclose(r[:, b], _exact_bin_means(xs, y[b], edges), rtol=1e-11, atol=1e-12)


def test_trapz_strided_redshifts():
    """A strided z view gives the same result as its contiguous copy (regression)."""
    x = np.linspace(1000, 2000, 501)
    y = np.sin(x / 37.0)
    z = np.linspace(0.2, 0.3, 40)
    e = np.linspace(1400, 2300, 30)
    assert np.array_equal(orc.trapz_rebin(x, y, edges=e, myz=z[::7]),
                          orc.trapz_rebin(x, y, edges=e, myz=z[::7].copy()))


def _constant_density_ok(rebin):
    x = np.arange(10) * 1.1
    return all(np.all(rebin(x, np.ones(10), edges=np.linspace(0.5, 8.3, n)) == 1.0)
               for n in range(3, 10))


def test_planted_mistake_in_sweep_is_caught():
    """A sweep that forgets the bin-width normalisation fails the exact check."""
    assert _constant_density_ok(orc.trapz_rebin)

    def no_normalisation(x, y, edges=None, **kw):
        return orc.trapz_rebin(x, y, edges=edges, **kw) * np.diff(edges)
    assert not _constant_density_ok(no_normalisation)


# ---------------------------------------------------------------- IGM
def test_igm_semantics():
    # test_igm.py:103-135 (CPU): shapes, T == 1 redward of Lya, T < 1 somewhere
    obswave = np.arange(3600, 9801, 50).astype(float)
    zz = np.arange(0, 5.1, 0.5)
    restwave = np.outer(1 / (1 + zz), obswave)
    for model in ('Calura12', 'Kamble20'):
        t = orc.transmission_Lyman(zz, obswave, model=model, always_return_array=False)
        assert t.shape == (len(zz), len(obswave))
        assert np.any(t < 1.0)
        assert np.all(t[restwave > orc.LYA_WAVELENGTH] == 1.0)
    assert orc.transmission_Lyman(zz, obswave, model='None', always_return_array=False) is None
    assert np.all(orc.transmission_Lyman(zz, obswave, model='None') == 1.0)
    # scalar / None rules (test_igm.py:32-51)
    assert orc.transmission_Lyman(0.0, 3600.0, model='Calura12', always_return_array=False) is None
    assert orc.transmission_Lyman(0.0, 3600.0, model='Calura12') == 1.0
    assert np.isscalar(orc.transmission_Lyman(4.0, 3600.0, model='Calura12'))


def test_igm_closed_form():
    z, lobs = 3.5, 4000.0   # rest 888.9 A: all five Calura12 lines absorb
    t = orc.transmission_Lyman(np.array([z]), np.array([lobs]), model='Calura12')[0, 0]
    expect = 1.0
    for (line, A, B) in orc.LYMAN_SERIES['Calura12']:
        expect *= np.exp(-A * (lobs / line) ** B)
    assert t == pytest.approx(expect, rel=1e-14)
    # custom Lya-only table: exp(-0.0036 (1+z_abs)^3.46) with 1+z_abs = lobs/line
    t = orc.transmission_Lyman(np.array([z]), np.array([lobs]), model='x',
                               lines=[(1215.67, 0.0036, 3.46)])[0, 0]
    assert t == pytest.approx(np.exp(-0.0036 * (lobs / 1215.67) ** 3.46), rel=1e-14)


# ---------------------------------------------------------------- chi2
class _Spec:
    def __init__(self, wave, flux, ivar, R, key):
        self.wave, self.flux, self.ivar, self.R, self.wavehash = wave, flux, ivar, R, key
        self.nwave = len(wave)


def _toy(seed=0, nb=3, nz=4, n=(40, 30)):
    rng = np.random.default_rng(seed)
    spectra, tdata = [], {}
    for i, m in enumerate(n):
        band = rng.uniform(0, 1, (m, 11))
        band /= band.sum(axis=1, keepdims=True)
        dense = np.zeros((m, m))
        for p in range(m):
            for k in range(11):
                q = p + k - 5
                if 0 <= q < m:
                    dense[p, q] = band[p, k]
        R = scipy.sparse.csr_matrix(dense)
        spectra.append(_Spec(np.arange(m, dtype=float), rng.normal(size=m),
                             rng.uniform(0.5, 2, m), R, f'c{i}'))
        tdata[f'c{i}'] = rng.normal(size=(nz, m, nb))
    return spectra, tdata


def _lstsq_ok(batch_fn):
    spectra, tdata = _toy()
    w, f, wf = orc.spectral_data(spectra)
    chi2, coeff = batch_fn(spectra, tdata, w, f, wf, 4, 3)
    ok = True
    for i in range(4):
        A = np.vstack([s.R.toarray() @ tdata[s.wavehash][i] for s in spectra])
        sw = np.sqrt(w)
        c, *_ = np.linalg.lstsq(sw[:, None] * A, sw * f, rcond=None)
        r = f - A @ c
        ok &= np.allclose(coeff[i], c, rtol=1e-10, atol=1e-12)
        ok &= bool(np.isclose(chi2[i], np.sum(w * r * r), rtol=1e-10))
        # quadratic identity chi2 = f'Wf - y'c at the solution
        y = A.T @ (w * f)
        ok &= bool(np.isclose(chi2[i], np.sum(w * f * f) - y @ coeff[i], rtol=1e-9))
    return ok


def test_calc_zchi2_batch_matches_lstsq():
    assert _lstsq_ok(orc.calc_zchi2_batch)


def test_planted_mistake_in_normal_equations_is_caught():
    """Normal equations that drop the weights fail the least-squares check."""
    def unweighted(spectra, tdata, w, f, wf, nz, nb):
        return orc.calc_zchi2_batch(spectra, tdata, np.ones_like(w), f, f, nz, nb)
    assert not _lstsq_ok(unweighted)


def test_calc_zchi2_batch_huge_conventions():
    spectra, tdata = _toy(1)
    w, f, wf = orc.spectral_data(spectra)
    chi2, coeff = orc.calc_zchi2_batch(spectra, tdata, 0 * w, f, 0 * wf, 4, 3)
    assert np.all(chi2 == orc.HUGE_CHI2) and np.all(coeff == 0)
    for k in tdata:
        tdata[k][2, :, 1] = 0.0      # basis 1 vanishes at z index 2 -> singular M
    chi2, coeff = orc.calc_zchi2_batch(spectra, tdata, w, f, wf, 4, 3)
    assert chi2[2] == orc.HUGE_CHI2 and np.all(coeff[2] == 0)
    assert np.all(chi2[[0, 1, 3]] < orc.HUGE_CHI2)


def test_negOII_penalty_closed_form():
    rng = np.random.default_rng(3)
    O = rng.normal(size=(3, 19))
    c = rng.normal(size=(5, 3))
    pen = orc.calc_negOII_penalty(O, c)
    flux = c @ O.sum(axis=1)
    assert np.allclose(pen, np.where(flux < 0, -flux, 0.0))


# ---------------------------------------------------------------- minima
def test_find_minima_reference_vector():
    # fitz.py:52-53
    assert list(orc.find_minima([1, 1, 1, 2, 2, 2])) == [0, 1, 2, 4, 5]
    assert list(orc.find_minima([3, 1, 2, 0, 5])) == [3, 1]


def test_minfit_exact_parabola():
    x = np.array([0.1, 0.2, 0.3])
    x0, s, y0 = 0.17, 0.05, 12.0
    y = y0 + ((x - x0) / s) ** 2
    rx0, rs, ry0, zw = orc.minfit(x, y)
    assert rx0 == pytest.approx(x0, rel=1e-10)
    assert rs == pytest.approx(s, rel=1e-10)
    assert ry0 == pytest.approx(y0, rel=1e-10)
    assert zw == 0
    assert orc.minfit(x[:2], y[:2])[3] == orc.ZW_BAD_MINFIT
    assert orc.minfit(x, -y)[3] & orc.ZW_BAD_MINFIT


# ---------------------------------------------------------------- zfind helpers
def test_calc_deltachi2_reference_vectors():
    # test_zfind.py:141-170
    chi2 = np.array([1.0, 2.0, 20.0, 40.0])
    d, s = orc.calc_deltachi2(chi2, np.array([3.0, 3.1, 3.2, 3.3]), np.array([0, 0, 0, 0]))
    assert np.all(d == [1, 18, 20, 0.0]) and np.all(s == [True, True, False, False])
    d, s = orc.calc_deltachi2(chi2, np.array([3.0, 3.0001, 4.0, 4.0001]), np.array([0, 0, 0, 0]))
    assert np.all(d == [19, 18, 0.0, 0.0]) and not np.any(s)
    d, s = orc.calc_deltachi2(chi2, np.array([3.0, 3.1, 3.2, 3.3]), np.array([0, 32, 0, 0]))
    assert np.all(d == [19, 18, 20, 0.0]) and np.all(s == [False, True, False, False])
    # test_zscan.py:101-111
    chi2 = np.array([1.0, 2.0, 4.0, 8.0])
    d, _ = orc.calc_deltachi2(chi2, np.array([3.0, 3.1, 3.2, 3.3]), np.zeros(4, int))
    assert np.all(d == [1, 2, 4, 0.0])
    d, _ = orc.calc_deltachi2(chi2, np.array([3.0, 3.0, 4.0, 4.0]), np.zeros(4, int))
    assert np.all(d == [3, 2, 0.0, 0.0])


def test_sort_zfit_dict_reference_vectors():
    # test_zfind.py:86-139
    cases = [([0, 0, 0], [2, 1, 0]), ([0, 0, 32], [1, 0, 2]), ([0, 0, 1024], [1, 0, 2]),
             ([0, 0, 256], [2, 1, 0]), ([0, 256, 288], [1, 0, 2])]
    for zwarn, rank in cases:
        z = dict(id=np.array([1, 2, 3]), zwarn=np.array(zwarn), chi2=np.array([30, 20, 10]),
                 rank=np.array(rank))
        orc.sort_zfit_dict(z)
        assert list(z['rank']) == [0, 1, 2]


def test_resolution_band_dia_fast_path():
    """The dia_matrix slicing path of resolution_band equals the generic COO path
    (non-square shapes, data longer than the matrix, zero wide diagonals)."""
    import scipy.sparse
    from redrock_amd.targets import resolution_band
    rng = np.random.default_rng(4)
    data = rng.normal(size=(5, 60))
    offs = np.array([0, 2, -3, 5, 7])
    data[4] = 0.0
    for shape in ((50, 55), (55, 50), (60, 60)):
        R = scipy.sparse.dia_matrix((data, offs), shape=shape)
        assert np.array_equal(resolution_band(R), resolution_band(R.tocsr()))
    data[4, 10] = 1.0
    with pytest.raises(NotImplementedError):
        resolution_band(scipy.sparse.dia_matrix((data, offs), shape=(50, 55)))


def _nnls_bruteforce(M, y):
    """min |M x - y|^2 over x >= 0 by enumerating every free set: the feasible
    least-squares solution whose gradient M^T (M x - y) is >= 0 on the fixed
    (zero) coefficients -- the Kuhn-Tucker point, unique for full-rank M."""
    import itertools
    n = M.shape[1]
    best = None
    for k in range(n + 1):
        for F in itertools.combinations(range(n), k):
            x = np.zeros(n)
            if F:
                x[list(F)] = np.linalg.lstsq(M[:, F], y, rcond=None)[0]
            if np.any(x[list(F)] <= 0):
                continue
            g = M.T @ (M @ x - y)
            fixed = [j for j in range(n) if j not in F]
            scale = np.abs(M).sum() * np.abs(y).max()
            if np.all(g[fixed] >= -1e-10 * scale):
                r = np.sum((M @ x - y) ** 2)
                if best is None or r < best[0]:
                    best = (r, x)
    return best[1]


@pytest.mark.parametrize('seed', range(8))
def test_oracle_nnls_is_the_kkt_point(seed):
    """scipy.optimize.nnls(M, y) -- the reference's NMF / NNLS solve
    (zscan.py:1153-1172) -- equals the brute-force Kuhn-Tucker point of
    min |M x - y|^2, x >= 0; with no active bound it is np.linalg.solve."""
    rng = np.random.default_rng(seed)
    nb = 2 + seed % 5
    A = np.abs(rng.normal(size=(300, nb))) + 0.2 * rng.normal(size=(300, nb))
    w = rng.uniform(0.5, 2.0, 300)
    M = A.T @ (w[:, None] * A)
    truth = rng.normal(size=nb) * (1.0 if seed % 2 else 0.3)
    truth[0] = abs(truth[0])
    y = M @ truth + rng.normal(size=nb)
    got = orc.solve_matrices_nnls(M, y)
    want = _nnls_bruteforce(M, y)
    assert np.all(got >= 0)
    assert np.allclose(got, want, rtol=1e-8, atol=1e-10 * max(np.abs(want).max(), 1.0))
    free = np.linalg.solve(M, y)
    if np.all(free > 0):
        assert np.allclose(got, free, rtol=1e-8)


def test_oracle_nnls_in_the_scan_clamps_coefficients():
    """calc_zchi2_batch with algo 'NMF' gives non-negative coefficients and a
    chi2 no smaller than the unconstrained (PCA) one at every redshift."""
    from redrock_amd import synth
    tm = synth.galaxy_template()
    targets, _ = synth.make_targets(tm, 1, seed=3)
    dwave = {s.wavehash: s.wave for s in targets[0].spectra}
    zsel = tm.redshifts[400:412]
    tdata = orc.template_cache(tm, dwave, myz=zsel)
    w, f, wf = orc.spectral_data(targets[0].spectra)
    c_pca, _ = orc.calc_zchi2_batch(targets[0].spectra, tdata, w, f, wf, len(zsel), tm.nbasis)
    c_nn, cf = orc.calc_zchi2_batch(targets[0].spectra, tdata, w, f, wf, len(zsel), tm.nbasis,
                                    algo='NMF')
    assert np.all(cf >= 0)
    assert np.all(c_nn >= c_pca * (1 - 1e-12))


def test_masked_ivar_known_row_sums():
    """targets.py:33-35 on resolution matrices whose row sums are known by
    construction: rows below 0.99 (not at it) lose their ivar, in the oracle
    and in the product's Spectrum alike."""
    import scipy.sparse
    from redrock_amd.targets import Spectrum
    diag = np.array([1.0, 0.995, 0.99, 0.9899, 0.5, 1.2])
    R = scipy.sparse.dia_matrix((diag[None, :], [0]), shape=(6, 6))
    ivar = np.arange(1.0, 7.0)
    want = ivar.copy()
    want[[3, 4]] = 0.0
    assert np.array_equal(orc.masked_ivar(R, ivar), want)
    s = Spectrum(np.linspace(4000, 4005, 6), np.zeros(6), ivar.copy(), R)
    assert np.array_equal(s.ivar, want)


def test_masked_ivar_truncated_edges():
    """The test recipe's Gaussian (sigma 2 px unsquared) loses 3.5 % of its
    mass in the third row from each edge and 0.6 % in the fourth: rows 0-2 and
    n-3..n-1 are masked, none else."""
    from tests import refsim
    from redrock_amd.targets import Spectrum
    n = 100
    R = refsim.resolution(n)
    ivar = np.ones(n)
    want = np.ones(n)
    want[[0, 1, 2, n - 3, n - 2, n - 1]] = 0.0
    assert np.array_equal(orc.masked_ivar(R, ivar), want)
    s = Spectrum(np.arange(4000.0, 4000.0 + 5 * n, 5.0), np.zeros(n), ivar.copy(), R)
    assert np.array_equal(s.ivar, want)
