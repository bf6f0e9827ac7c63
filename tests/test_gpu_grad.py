"""Gradient of chi2(z) (rr_zscan_grad) against central differences and a float64
torch reference, the autograd Function, and the Newton refinement."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope='module')
def problem(gpu):
    from redrock_amd import synth
    from redrock_amd.grad import GradProblem
    tm = synth.galaxy_template()
    targets, zt = synth.make_targets(tm, 3, seed=17, snr=0.5, zrange=(0.2, 1.2))
    return tm, targets, zt, GradProblem(tm, targets, [0, 1, 2, 0, 1, 2])


def test_grad_chi2_matches_forward(problem):
    """The adjoint kernel's explicit residual equals the forward scan's chi2."""
    tm, targets, zt, P = problem
    z = np.array([0.3, 0.71, 0.95, 0.301, zt[1], zt[2] + 0.01])
    chi2, coeff = P.forward(z)
    _, echi2, _ = P.grad(z, coeff)
    rel = np.abs(echi2.cpu().numpy() - chi2.cpu().numpy()) / chi2.cpu().numpy()
    assert rel.max() < 1e-10


def test_dz_matches_central_difference(problem):
    """Central differences with h = 1e-6 in z."""
    tm, targets, zt, P = problem
    z = np.array([0.3, 0.71, 0.95, 0.45, zt[1] + 3e-4, zt[2] - 5e-4])
    h = 1e-6
    chi2, coeff = P.forward(z)
    g, _, _ = P.grad(z, coeff)
    cp, _ = P.forward(z + h)
    cm, _ = P.forward(z - h)
    fd = (cp - cm).cpu().numpy() / (2 * h)
    g = g.cpu().numpy()
    print('analytic', g, 'fd', fd)
    scale = np.maximum(np.abs(fd), 1e-3 * np.abs(chi2.cpu().numpy()))
    assert np.all(np.abs(g - fd) <= 2e-4 * scale)


def _torch_chi2(tm, target, z, flux):
    """Float64 torch reference of chi2(z): antiderivative rebin, band R, normal equations."""
    import torch
    from redrock_amd.rebin import centers2edges
    dev = flux.device
    x = torch.as_tensor(tm.wave, device=dev)
    w = torch.as_tensor(np.concatenate([s.ivar for s in target.spectra]), device=dev)
    f = torch.as_tensor(np.concatenate([s.flux for s in target.spectra]), device=dev)
    Ts, Rs = [], []
    for s in target.spectra:
        e = torch.as_tensor(centers2edges(s.wave), device=dev)
        xs = x * (1 + z)
        seg = 0.5 * (flux[:, 1:] + flux[:, :-1]) * (xs[1:] - xs[:-1])
        F = torch.cat([torch.zeros_like(flux[:, :1]), torch.cumsum(seg, dim=1)], dim=1)
        j = torch.clamp(torch.searchsorted(xs.detach(), e.detach(), right=True) - 1, 0, len(x) - 2)
        t = e - xs[j]
        slope = (flux[:, j + 1] - flux[:, j]) / (xs[j + 1] - xs[j])
        Fe = F[:, j] + flux[:, j] * t + 0.5 * slope * t * t
        T = ((Fe[:, 1:] - Fe[:, :-1]) / (e[1:] - e[:-1])).T          # (n, nb)
        band = torch.as_tensor(s.rband(), device=dev)
        Tp = torch.nn.functional.pad(T.T, (5, 5)).T                    # (n+10, nb)
        A = sum(band[:, k:k + 1] * Tp[k:k + len(s.wave)] for k in range(11))
        Ts.append(A)
    A = torch.cat(Ts)
    M = A.T @ (w[:, None] * A)
    y = A.T @ (w * f)
    c = torch.linalg.solve(M, y)
    r = f - A @ c
    return torch.sum(w * r * r)


def test_autograd_matches_torch_reference(problem):
    """d chi2/dz and d chi2/d(template flux) vs autograd through a float64 torch reference."""
    import torch
    from redrock_amd.grad import GradProblem, chi2_at_redshift
    tm, targets, zt, _ = problem
    P1 = GradProblem(tm, targets[:1], [0])
    z0 = float(zt[0]) + 2e-4
    flux = torch.as_tensor(np.ascontiguousarray(tm.flux), device='cuda').requires_grad_(True)
    z = torch.tensor([z0], dtype=torch.float64, device='cuda', requires_grad=True)
    chi2 = chi2_at_redshift(z, flux, P1)
    chi2.sum().backward()
    flux_r = torch.as_tensor(np.ascontiguousarray(tm.flux), device='cuda').requires_grad_(True)
    z_r = torch.tensor(z0, dtype=torch.float64, device='cuda', requires_grad=True)
    chi2_r = _torch_chi2(tm, targets[0], z_r, flux_r)
    chi2_r.backward()
    assert abs(float(chi2) - float(chi2_r)) <= 1e-9 * float(chi2_r)
    assert abs(float(z.grad) - float(z_r.grad)) <= 1e-6 * abs(float(z_r.grad)) + 1e-9 * float(chi2_r)
    gy, gy_r = flux.grad.cpu().numpy(), flux_r.grad.cpu().numpy()
    assert np.allclose(gy, gy_r, rtol=1e-6, atol=1e-8 * np.abs(gy_r).max())


def test_newton_refinement(problem):
    """Newton on the analytic gradient reaches a stationary point next to the parabola fit."""
    from redrock_amd import zscan, fitz
    from redrock_amd.templates import DistTemplate
    from redrock_amd.grad import newton_refine
    tm, targets, zt, P = problem
    dwave = {s.wavehash: s.wave for s in targets[0].spectra}
    dt = DistTemplate(tm, dwave)
    chi2, _, pen = zscan.calc_zchi2([0, 1, 2], targets, dt)
    res = fitz.fitz_targets(chi2 + pen, tm.redshifts, targets, tm, nminima=1)
    zpar = np.array([r['z'][0, 0] for r in res])
    zz = [r['zz'][0] for r in res]
    from redrock_amd.grad import GradProblem
    P3 = GradProblem(tm, targets, [0, 1, 2])
    info = {}
    zn, c2n, gn = newton_refine(P3, zpar, [q[0] for q in zz], [q[-1] for q in zz], info=info)
    step = np.array([q[1] - q[0] for q in zz])
    print('parabola', zpar, 'newton', zn, 'grad', gn, info)
    assert np.all(info['converged'])
    assert np.all(np.abs(zn - zpar) < 2 * step)
    c2p, _ = P3.forward(zpar)
    assert np.all(c2n <= c2p.cpu().numpy())
    assert np.all(np.abs(gn) * step < 1e-6 * np.abs(c2n))


def test_newton_backtracks_and_keeps_reference(problem):
    """Started far from the minimum (a bracket edge) Newton only accepts steps
    that lower chi2, ends at a stationary point, and a reference point with a
    lower chi2 than its result is returned instead."""
    from redrock_amd import zscan, fitz
    from redrock_amd.templates import DistTemplate
    from redrock_amd.grad import GradProblem, newton_refine
    tm, targets, zt, P = problem
    dwave = {s.wavehash: s.wave for s in targets[0].spectra}
    dt = DistTemplate(tm, dwave)
    chi2, _, pen = zscan.calc_zchi2([0, 1, 2], targets, dt)
    zg = tm.redshifts
    imin = (chi2 + pen).argmin(axis=1)
    lo, hi = zg[imin - 5], zg[imin + 5]
    P3 = GradProblem(tm, targets, [0, 1, 2])
    c_lo = P3.forward(lo)[0].cpu().numpy()
    info = {}
    zn, c2n, gn = newton_refine(P3, lo, lo, hi, info=info)
    print('start', lo, 'newton', zn, info)
    assert np.all(c2n <= c_lo)
    assert np.all(info['converged'])
    step = zg[imin + 1] - zg[imin]
    assert np.all(np.abs(gn) * step < 1e-6 * np.abs(c2n))
    # a reference with a lower chi2 than Newton's result wins
    fake = c2n - 1.0
    info2 = {}
    zk, ck, gk = newton_refine(P3, lo, lo, hi, zref=lo, chi2ref=fake, info=info2)
    assert np.all(info2['kept_reference']) and np.allclose(zk, lo)


def test_flux_gradient_run_to_run_identical(problem):
    """d chi2 / d(template flux) is summed in a fixed order: two runs give the
    same bits (no floating-point atomics)."""
    tm, targets, zt, P = problem
    z = np.array([0.3, 0.71, 0.95, 0.301, zt[1], zt[2] + 0.01])
    chi2, coeff = P.forward(z)
    out = [P.grad(z, coeff, want_dy=True, scale=np.linspace(0.5, 2.0, 6))[2].cpu().numpy()
           for _ in range(3)]
    assert np.any(out[0] != 0)
    assert np.array_equal(out[0], out[1]) and np.array_equal(out[0], out[2])
