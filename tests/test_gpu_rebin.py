"""GPU rebin (rr_rebin) against the oracle's sweep and the reference's known answers."""
import numpy as np
import pytest

from oracle import rr_oracle as orc

pytestmark = pytest.mark.gpu


def test_docstring_vectors(gpu):
    # rebin.py:222-250, both return types
    from redrock_amd import rebin
    x = np.arange(10)
    y = np.ones((2, 10))
    y[1, :] = np.arange(10)
    r = rebin.trapz_rebin(x, np.ones(10), edges=[0, 2, 4, 6, 8])
    assert isinstance(r, np.ndarray) and r.shape == (4,) and np.all(r == 1.0)
    g = rebin.trapz_rebin(x, y, edges=[0, 2, 4, 6, 8], use_gpu=True)
    assert g.is_cuda and tuple(g.shape) == (4, 2)
    assert np.allclose(g.cpu().numpy(), [[1, 1], [1, 3], [1, 5], [1, 7]])
    r = rebin.trapz_rebin(x, y, edges=[0, 2, 4, 6, 8], myz=[0.0, 0.5])
    assert r.shape == (2, 4, 2)
    assert np.allclose(r[1], [[1, 0.66666667], [1, 2], [1, 3.33333333], [1, 4.66666667]])
    assert rebin.trapz_rebin(x, y, edges=[0, 2, 4, 6, 8], myz=np.zeros(0)).shape == (0, 4, 1)


def test_constant_density_exact(gpu):
    # test_rebin.py:34-51 / :88-109 (exact equality on the GPU too)
    from redrock_amd import rebin
    x = np.arange(10) * 1.1
    y = np.ones(10)
    for nedge in range(3, 10):
        yy = rebin.trapz_rebin(x, y, edges=np.linspace(min(x), max(x), nedge))
        assert np.all(yy == 1.0), yy
    for nedge in range(3, 30):
        assert np.allclose(rebin.trapz_rebin(x, y, edges=np.linspace(0.5, 8.3, nedge)), 1.0)


def test_range_errors(gpu):
    from redrock_amd import rebin
    x = np.arange(10)
    with pytest.raises(ValueError):
        rebin.trapz_rebin(x, np.ones(10), edges=np.arange(-1, 9))
    with pytest.raises(ValueError):
        rebin.trapz_rebin(x, np.ones(10), edges=np.arange(1, 11))


def test_uneven_and_multidim_match_oracle(gpu):
    # test_rebin.py:111-154 shapes, compared with the oracle bit for bit
    from redrock_amd import rebin
    x = np.linspace(0, 10, 100) ** 2
    y = np.sqrt(x)
    edges = np.linspace(0, 100, 21)
    assert np.array_equal(rebin.trapz_rebin(x, y, edges=edges), orc.trapz_rebin(x, y, edges=edges))
    x = np.arange(10.)
    y = np.ones((2, 10))
    y[1] = np.arange(10)
    myz = np.linspace(0, 1, 11)
    g = rebin.trapz_rebin(x, y, edges=[0, 2, 4, 6, 8], myz=myz)
    assert g.shape == (11, 4, 2)
    assert np.array_equal(g, orc.trapz_rebin(x, y, edges=[0, 2, 4, 6, 8], myz=myz))


def test_galaxy_cache_bit_exact_vs_sweep(gpu):
    """Every (z, pixel, basis) of the GALAXY cache equals the oracle sweep's double."""
    from redrock_amd import synth
    from redrock_amd.zscan import TemplateCache
    tm = synth.galaxy_template()
    grids = synth.camera_grids()
    zsel = tm.redshifts[::97]
    cache = TemplateCache(tm, grids, redshifts=zsel)
    got = {k: v.cpu().numpy() for k, v in cache.reference_layout().items()}
    for k, w in grids.items():
        ref = orc.trapz_rebin(tm.wave, tm.flux, xnew=w, myz=zsel)
        assert got[k].shape == ref.shape
        assert np.array_equal(got[k], ref), f'camera {k}: max diff {np.max(np.abs(got[k]-ref))}'


def test_qso_cache_with_fused_igm(gpu):
    """IGM fused in the rebin equals rebin-then-multiply (templates.py:586-597)."""
    from redrock_amd import synth
    from redrock_amd.zscan import TemplateCache
    tm = synth.qso_template()
    grids = synth.camera_grids()
    zsel = tm.redshifts[::151]
    cache = TemplateCache(tm, grids, redshifts=zsel)
    got = {k: v.cpu().numpy() for k, v in cache.reference_layout().items()}
    ref = orc.template_cache(tm, grids, myz=zsel, lines=[(1215.67, 0.0036, 3.46)])
    for k in grids:
        scale = np.max(np.abs(ref[k]))
        assert np.max(np.abs(got[k] - ref[k])) <= 1e-14 * scale
    # redward of Lya in every pixel at the lowest z: no attenuation at all
    plain = orc.trapz_rebin(tm.wave, tm.flux, xnew=grids['z'], myz=zsel[:1])
    assert np.array_equal(got['z'][:1], plain)


def test_cache_layouts_agree(gpu):
    """The f64 [b][row][z] and 16-bit [z][b][row] caches hold the same values."""
    import torch
    from redrock_amd import synth
    from redrock_amd.zscan import TemplateCache, PixelLayout
    tm = synth.galaxy_template()
    grids = synth.camera_grids()
    zsel = tm.redshifts[100:140]
    cache = TemplateCache(tm, grids, redshifts=zsel)
    lay = cache.layout(list(grids))
    # the f64 cache zeroes only what the rebin does not write: hand it NaN memory
    ldz = 64
    junk = torch.full((tm.nbasis * lay.cache_rows * ldz,), float('nan'), dtype=torch.float64, device='cuda')
    del junk
    t64, ldz = cache.get(lay, 'f64')
    t16, _ = cache.get(lay, 'f16')
    full = t64.reshape(tm.nbasis, lay.cache_rows, ldz)
    assert not torch.any(torch.isnan(full))
    assert torch.all(full[:, lay.halo + lay.npix:] == 0) and torch.all(full[:, :, len(zsel):] == 0)
    a = full[:, :, :len(zsel)].permute(2, 0, 1)
    n16 = len(zsel) * tm.nbasis * lay.cache_rows
    assert torch.all(t16[n16:] == 0) and t16.numel() >= n16 + 64   # zero block after the cache
    b = t16[:n16].view(torch.float16).reshape(len(zsel), tm.nbasis, lay.cache_rows).double() / cache.tscale
    assert torch.all(a[:, :, :lay.halo] == 0) and torch.all(b[:, :, :lay.halo] == 0)
    assert torch.allclose(a, b, rtol=2 ** -10, atol=1e-6 * float(a.abs().max()))


@pytest.mark.parametrize('step', [0.5, 0.02])
def test_lds_window_equals_global_path(gpu, step):
    """Bins-fast outputs stage each workgroup's template window in LDS; the
    z-fastest output reads the samples from global memory: the same doubles.
    A 0.02 A template sampling gives windows of more than 2048 samples at low z,
    which take the global path inside the bins-fast launch; both equal the
    oracle's sweep, IGM factor included."""
    import torch
    from redrock_amd import rebin, synth
    rng = np.random.default_rng(5)
    wave = np.arange(3000.0, 9000.0 + step / 2, step)
    flux = 1.0 + 0.3 * np.sin(wave[None, :] / np.array([[37.0], [91.0], [13.0]])) \
        + 0.05 * rng.standard_normal((3, wave.size))
    obs = synth.camera_grids()['b']
    edges = rebin.centers2edges(obs)
    z = np.linspace(0.0, 0.19, 37)                # rest frame inside [3000, 9000] A
    lyman = [(4500.0, 0.002, 3.0)]              # a line inside the window, so the factor varies
    d = lambda a: torch.as_tensor(np.ascontiguousarray(a), dtype=torch.float64, device='cuda')
    x, y, e, zz, w = d(wave), d(flux), d(edges), d(z), d(obs)
    nz, nbin, nb = len(z), len(obs), flux.shape[0]
    print(f'case: rebin LDS window vs global, step {step}, {wave.size} samples', flush=True)
    a = torch.empty((nz, nbin, nb), dtype=torch.float64, device='cuda')     # bins-fast: LDS
    rebin.rebin_into(a, x, y, e, zz, nbin * nb, nb, 1, lyman=lyman, wave=w)
    b = torch.empty((nb, nbin, nz), dtype=torch.float64, device='cuda')     # z-fastest: global
    rebin.rebin_into(b, x, y, e, zz, 1, nz, nbin * nz, lyman=lyman, wave=w)
    a = a.cpu().numpy()
    b = b.permute(2, 1, 0).cpu().numpy()
    assert np.array_equal(a, b)
    ref = orc.trapz_rebin(wave, flux, xnew=obs, myz=z)
    T = np.ones((nz, nbin))
    for i, zi in enumerate(z):
        lrest = obs / (1 + zi)
        m = lrest < 4500.0
        T[i, m] = np.exp(-0.002 * (obs[m] / 4500.0) ** 3.0)
    assert np.allclose(a, ref * T[:, :, None], rtol=1e-14, atol=0)
