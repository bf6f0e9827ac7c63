#!/usr/bin/env python
"""Benchmark of the chi2 scan: zchi2 evaluations per second on one MI355X.

Workload (the galaxy petal): 500 GALAXY targets x 1977 redshifts x 10 basis
vectors over 3 cameras (7958 pixels), synthetic data (redrock_amd.synth).
One eval = one (target, template, z) chi2 including its coefficient solve.
A step is one full scan of all targets over all redshifts with every input
already resident in HBM (template cache built once, as DistTemplate does).

The headline value is the 'f16r' scan: fp16 normal equations on MFMA (fp32
accumulation) for the coefficients, FP64 Cholesky, then chi2 as the
reference's explicit residual sum w (f - R T c)^2 at those coefficients
(u = T c and the R-apply as split-fp16 MFMAs, fp32 accumulation, FP64 sums),
plus the FP64 polish of the rows near each minimum.  The plain FP64 scan (one target per lane) runs on the same inputs
for the speedup and the best-z agreement; the split-fp16 (f16x3) and the
single-pass fp16 / bf16 / fp32-MFMA modes are reported beside it as
variants.  The CPU baseline is the float64 oracle (a numpy restatement of
the reference's CPU loop) on a bounded sample of the same targets, one
process per core and on one core, run before the GPU is initialised.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--ntargets 500]
"""
import os

for _v in ('OMP_NUM_THREADS', 'OPENBLAS_NUM_THREADS', 'MKL_NUM_THREADS'):
    os.environ[_v] = '1'      # the CPU baseline runs one single-threaded process per core

import argparse
import json
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = 'zchi2 evals/sec on 1 GPU: mixed-precision vs FP64 plain at same best-z'
# reference-form FLOPs per eval (R-apply, weights, Gram, y, solve, model, chi2)
FLOP_PER_EVAL = {'GALAXY': 3763200, 'QSO': 1137348, 'STAR': 1495337}
PEAK_FLOPS = {'f16': 2.5e15, 'bf16': 2.5e15, 'f16x3': 2.5e15, 'f16r': 2.5e15, 'f64': 78.6e12,
              'f32': 157.3e12}
# fields of rr_zscan_resid_tiles entries by kind (1: FP32 lane kernel, 2: matrix cores)
RESID_FIELDS = {1: ('kind', 'waves', 'items_per_wave', 'z_per_wave'),
                2: ('kind', 'waves', 'items_per_workgroup', 'z_per_workgroup')}
# the GPU box allots this many host cores to a one-GPU job
CPU_SHARE = 16


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def kernel_signature(precision, nb, cfg):
    """Demangled name of the mixed scan kernel instance for `cfg` = (wa, wb, tw, zw)."""
    wa, wb, tw, zw = cfg
    if precision == 'f16x3':
        return f'void zscan_mfma3_kernel<{nb}, {wa}, {wb}, {tw}, {zw}>'
    ty = {'f16': 1, 'bf16': 2, 'f16r': 1}[precision]
    return f'void zscan_mfma_kernel<{nb}, {ty}, {wa}, {wb}, {tw}, {zw}>'


def measured_counter(name, signature):
    """A per-launch figure of `signature` on the galaxy petal from the committed
    rocprofv3 counter summary: 'traffic_bytes_per_launch' (FETCH_SIZE +
    WRITE_SIZE, calibrated against a plain copy) or 'mfma_util_percent'
    (MfmaUtil); None when the summary does not cover this kernel instance."""
    path = os.path.join(ROOT, 'profiles', 'zscan_galaxy500_counters.md')
    if not os.path.exists(path):
        return None
    key = f'{name} {signature} = '
    with open(path) as fh:
        for line in fh:
            if line.startswith(key):
                return float(line[len(key):])
    return None


_CPU = {}


def _cpu_one(i):
    from oracle import rr_oracle as orc
    tm, targets, tdata = _CPU['tm'], _CPU['targets'], _CPU['tdata']
    orc.calc_zchi2([targets[i]], tdata, tm.nbasis, getattr(tm, 'OIItemplate', None))
    return i


def _cpu_pool_run(ncores, nsample):
    import multiprocessing as mp
    ctx = mp.get_context('fork')
    with ctx.Pool(ncores) as pool:
        t0 = time.perf_counter()
        list(pool.imap_unordered(_cpu_one, range(nsample)))
        return time.perf_counter() - t0


def cpu_baseline(tm, targets, dwave, seconds):
    """Oracle scan on host cores: evals/s over a bounded sample of targets, one
    single-threaded process per core (targets split between them, as
    rrdesi --mp N, zscan.py:1037-1094), and on one core."""
    from oracle import rr_oracle as orc
    visible = len(os.sched_getaffinity(0))
    ncores = min(CPU_SHARE, visible)
    # the QSO workload's IGM: Lyman-alpha only, tau = 0.0036 (1+z_abs)^3.46
    lines = [(1215.67, 0.0036, 3.46)] if tm.template_type == 'QSO' else None
    tdata = orc.template_cache(tm, dwave, lines=lines)
    _CPU.update(tm=tm, targets=targets, tdata=tdata)
    t0 = time.perf_counter()
    _cpu_one(0)
    one = time.perf_counter() - t0
    nz = len(tm.redshifts)
    # one core: a few targets in this process
    n1 = max(1, min(len(targets), int(0.4 * seconds / max(one, 1e-3))))
    t0 = time.perf_counter()
    for i in range(n1):
        _cpu_one(i)
    wall1 = time.perf_counter() - t0
    nsample = max(ncores, min(len(targets), int(seconds * ncores / max(one, 1e-3))))
    nsample = min(nsample, len(targets))
    wall = _cpu_pool_run(ncores, nsample)
    return dict(value=nsample * nz / wall, unit='evals/s', cores=ncores, kind='port',
                value_1core=n1 * nz / wall1, host_cores_visible=visible,
                sample=f'{nsample} of the workload targets x {nz} z, float64 oracle '
                       f'(numpy CSR R-apply, normal equations, LAPACK gesv, explicit residual), '
                       f'one process per core on {ncores} cores (the host share of a one-GPU job; '
                       f'{visible} cores visible), {wall:.1f} s wall; 1 core: {n1} targets, '
                       f'{wall1:.1f} s')


def calibrate_copy(mib):
    """A plain device-to-device copy of known size, for scaling HBM byte counters."""
    import torch
    n = mib << 20
    x = torch.ones(n, dtype=torch.uint8, device='cuda')
    y = torch.empty_like(x)
    for _ in range(3):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({'copy_bytes': n, 'read_bytes': n, 'write_bytes': n}), flush=True)


def bench_mixed(args):
    """All targets x all template classes (nbasis 10 / 4 / 5, 1977 / 3795 / 501
    redshifts) with 0-30 % of every spectrum masked in runs of 50-500 pixels,
    each precision's whole scan replayed as one captured HIP graph."""
    import numpy as np
    import torch
    from redrock_amd import synth, zscan, _hip
    from redrock_amd.templates import DistTemplate
    t0 = time.perf_counter()
    tms = (synth.galaxy_template(), synth.qso_template(), synth.star_template())
    targets, types, ztrue = synth.mixed_targets(args.mixed_targets, seed=args.seed, snr=args.snr,
                                                mask_frac=(0.0, 0.3), templates=tms)
    dwave = {s.wavehash: s.wave for s in targets[0].spectra}
    log(f'mixed workload: {len(targets)} targets generated in {time.perf_counter() - t0:.1f} s')
    assert torch.cuda.is_available(), 'bench.py needs a GPU'
    dts = [DistTemplate(t, dwave) for t in tms]
    batch = zscan.TargetBatch(targets)
    caches = [(dt.cache, zscan._oiisum(dt.template)) for dt in dts]
    nt = len(targets)
    evals = nt * sum(len(t.redshifts) for t in tms)
    out = {}
    res = {}
    runs = [(args.precision, not args.no_polish)] + ([('raw', False)] if not args.no_polish else [])
    runs.append(('f64', False))
    for key, pol in runs:
        p = args.precision if key == 'raw' else key
        sg = zscan.ScanGraph(caches, batch, p, tune=(p != 'f64'), polish=pol)
        sg.capture()
        for _ in range(args.warmup):
            sg.replay()
        torch.cuda.synchronize()
        k = max(1, args.steps if p != 'f64' else args.steps // 2)
        t = time.perf_counter()
        for _ in range(k):
            sg.replay()
        torch.cuda.synchronize()
        dt_s = (time.perf_counter() - t) / k
        out[key] = dt_s
        res[key] = sg.results()
        log(f'{key}{" + FP64 polish" if pol else ""}: {dt_s * 1e3:.2f} ms per graph replay '
            f'({evals / dt_s / 1e6:.1f} M evals/s)')
        del sg
    rel, agree, agree_type = [], [], []
    fwf = (batch.wf * batch.f).sum(dim=1).cpu().numpy()
    best = {p: np.stack([r[0].min(axis=1) for r in res[p]], axis=1) for p in res}
    per_template = {}
    for tm, (c2m, _, pm), (c2d, _, pd) in zip(tms, res[args.precision], res['f64']):
        r = np.abs(c2m - c2d) / np.abs(c2d)
        rel.append(r)
        agree.append(np.mean(c2m.argmin(axis=1) == c2d.argmin(axis=1)))
        rb = r[np.arange(nt), c2d.argmin(axis=1)]
        per_template[tm.template_type] = {
            'max_rel_chi2_err': float(r.max()), 'median_rel_chi2_err': float(np.median(r)),
            'max_chi2_err_over_fwf': float((np.abs(c2m - c2d) / fwf[:, None]).max()),
            'max_rel_chi2_err_at_best_z': float(rb.max()), 'best_z_agreement': float(agree[-1])}
        log(f'{tm.template_type}: max rel err {r.max():.2e} (at best z {rb.max():.2e}), '
            f'best-z agreement {agree[-1]:.4f}')
    agree_type = float(np.mean(best[args.precision].argmin(axis=1) == best['f64'].argmin(axis=1)))
    rel = np.concatenate([r.ravel() for r in rel])
    unpolished = None
    if 'raw' in res:
        unpolished = {'value': evals / out['raw'], 'ms_per_step': out['raw'] * 1e3, 'per_template': {}}
        for tm, (c2m, _, _), (c2d, _, _) in zip(tms, res['raw'], res['f64']):
            r = np.abs(c2m - c2d) / np.abs(c2d)
            a = float(np.mean(c2m.argmin(axis=1) == c2d.argmin(axis=1)))
            unpolished['per_template'][tm.template_type] = {
                'best_z_agreement': a, 'max_rel_chi2_err': float(r.max()),
                'max_rel_chi2_err_at_best_z': float(r[np.arange(nt), c2d.argmin(axis=1)].max())}
            log(f'{tm.template_type} without the polish: best-z agreement {a:.4f}, '
                f'max rel err {r.max():.2e}')
    print(json.dumps({
        'metric': 'zchi2 evals/sec on 1 GPU, mixed GALAXY/QSO/STAR batch in one HIP graph',
        'value': evals / out[args.precision], 'unit': 'evals/s', 'n_gpus': 1,
        'steps': args.steps, 'warmup': args.warmup, 'ms_per_step': out[args.precision] * 1e3,
        'higher_is_better': True, 'scaling': 'weak', 'vs_baseline': None,
        'dtype': args.precision, 'data': 'synthetic',
        'config': {'workload': f'{nt} mixed GALAXY/QSO/STAR targets, 0-30% masked in 50-500 px runs, '
                               f'x GALAXY (nb 10, 1977 z) + QSO (nb 4, 3795 z) + STAR (nb 5, 501 z)',
                   'targets': nt, 'evals_per_step': evals},
        'fp64': {'value': evals / out['f64'], 'ms_per_step': out['f64'] * 1e3},
        'speedup_vs_fp64': out['f64'] / out[args.precision],
        'best_z_agreement_per_template': [float(a) for a in agree],
        'best_template_agreement': agree_type,
        'max_rel_chi2_err': float(rel.max()), 'median_rel_chi2_err': float(np.median(rel)),
        'per_template': per_template,
        'polish': (None if args.no_polish else
                   {'tol': zscan.POLISH_TOL[args.precision], 'max_windows': zscan.POLISH_WINDOWS,
                    'window': _hip.RR_POLISH_WINDOW}),
        'unpolished': unpolished,
    }), flush=True)


def bench_grad(args):
    """GALAXY targets: the 3 lowest minima of each coarse scan refined by Newton
    iterations on the analytic d chi2/dz (adjoint kernel, rebin recomputed)
    inside +-5 grid steps, beside the reference's 15-point parabola fit."""
    import numpy as np
    import torch
    from redrock_amd import synth, zscan
    from redrock_amd import fitz as fz
    from redrock_amd.templates import DistTemplate
    from redrock_amd.grad import GradProblem, newton_refine
    tm = synth.galaxy_template()
    targets, ztrue = synth.make_targets(tm, args.grad_targets, seed=args.seed, snr=args.snr)
    dwave = {s.wavehash: s.wave for s in targets[0].spectra}
    assert torch.cuda.is_available(), 'bench.py needs a GPU'
    nt, zg = len(targets), tm.redshifts
    nz = len(zg)
    dt = DistTemplate(tm, dwave)
    batch = zscan.TargetBatch(targets)
    oii = zscan._oiisum(tm)
    plan = zscan.ScanPlan(dt.cache, batch, np.arange(nt), np.zeros(nt), np.full(nt, nz),
                          precision=args.precision, oiisum=oii)
    plan.run()
    c2 = (plan.zchi2 + plan.penalty).reshape(nt, nz).cpu().numpy()
    del plan
    owner, imin = [], []
    for t in range(nt):
        for i in fz.find_minima(c2[t])[:3]:
            if 5 <= i < nz - 5:
                owner.append(t)
                imin.append(int(i))
    owner, imin = np.array(owner), np.array(imin)
    n = len(owner)
    log(f'grad workload: {nt} targets, {n} minima (3 lowest per target, +-5 grid steps)')
    z0, zlo, zhi = zg[imin], zg[imin - 5], zg[imin + 5]
    prob = GradProblem(tm, targets, owner, batch=batch)

    def timed(fn, k):
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(k):
            r = fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t) / k, r

    chi2, coeff = prob.forward(z0)
    prob.grad(z0, coeff, want_dy=True)             # warm-up
    k = max(2, args.steps // 2)
    t_fwd, _ = timed(lambda: prob.forward(z0), k)
    t_grad, _ = timed(lambda: prob.grad(z0, coeff), k)
    t_grad_dy, _ = timed(lambda: prob.grad(z0, coeff, want_dy=True), k)
    # reference method: 15-point fine grid over z[i-1..i+1] and a parabola through the minimum
    ref = fz._Refiner(tm, targets, zg, 15, 'f64', batch=batch)
    t_par, rr = timed(lambda: ref.refine(owner, imin), 1)
    zpar = np.where(rr['err'], np.nan, rr['zmin'])
    # Newton starts at the parabola vertex where it lies inside the +-5 step bracket
    ok = np.isfinite(zpar) & (zpar >= zlo) & (zpar <= zhi)
    zstart = np.where(ok, zpar, z0)
    info = {}
    t_newton, (zn, c2n, gn) = timed(lambda: newton_refine(prob, zstart, zlo, zhi, info=info), 1)
    c2par = prob.forward(zstart)[0].cpu().numpy()
    dv = 299792.458 * (zn - zpar) / (1 + zpar)
    better = np.mean(c2n[ok] <= c2par[ok])
    conv = info['converged']
    dzgrid = zg[imin + 1] - zg[imin]
    interior = (zn > zlo) & (zn < zhi)
    gscaled = np.abs(gn) * dzgrid / np.abs(c2n)
    gmax = float(gscaled[conv & interior].max()) if np.any(conv & interior) else None
    gp99 = float(np.quantile(gscaled[conv & interior], 0.99)) if np.any(conv & interior) else None
    log(f'forward (rebin + FP64 solve at {n} z) {t_fwd * 1e3:.1f} ms; adjoint dchi2/dz '
        f'{t_grad * 1e3:.1f} ms, with dchi2/dflux {t_grad_dy * 1e3:.1f} ms; Newton '
        f'{t_newton * 1e3:.0f} ms ({info["evaluations"]} evaluations, converged {conv.mean():.4f})')
    log(f'parabola fit {t_par * 1e3:.0f} ms; |Newton - parabola| median {np.median(np.abs(dv[ok])):.3f} '
        f'km/s, p99 {np.quantile(np.abs(dv[ok]), 0.99):.3f} km/s; chi2(Newton) <= chi2(parabola) '
        f'for {better:.4f} of minima; |dchi2/dz| dz_grid / chi2 at converged interior points: max {gmax}, '
        f'p99 {gp99}')
    print(json.dumps({
        'metric': 'adjoint d chi2/dz evaluations/sec on 1 GPU (GALAXY, rebin recomputed)',
        'value': n / t_grad, 'unit': 'gradients/s', 'n_gpus': 1, 'steps': k,
        'warmup': 1, 'ms_per_step': t_grad * 1e3, 'higher_is_better': True, 'scaling': 'weak',
        'vs_baseline': None, 'dtype': 'f64', 'data': 'synthetic',
        'config': {'workload': f'{nt} GALAXY targets, 3 lowest minima each, Newton inside +-5 '
                               f'grid steps (coarse scan in {args.precision})',
                   'targets': nt, 'minima': n},
        'forward_ms': t_fwd * 1e3, 'grad_ms': t_grad * 1e3, 'grad_with_dflux_ms': t_grad_dy * 1e3,
        'newton_ms': t_newton * 1e3, 'parabola_fit_ms': t_par * 1e3,
        'newton_minus_parabola_kms': {'median_abs': float(np.median(np.abs(dv[ok]))),
                                      'p99_abs': float(np.quantile(np.abs(dv[ok]), 0.99))},
        'newton_chi2_not_above_parabola': float(better),
        'minima_compared': int(ok.sum()),
        'newton_converged_fraction': float(conv.mean()),
        'newton_evaluations': int(info['evaluations']),
        'newton_at_bracket_edge_fraction': float(np.mean(~interior)),
        'max_dchi2dz_x_dzgrid_over_chi2_converged': gmax,
        'p99_dchi2dz_x_dzgrid_over_chi2_converged': gp99,
        'max_abs_dchi2dz_at_newton': float(np.max(np.abs(gn))),
    }), flush=True)


def _cpu_arch_one(k):
    from oracle import rr_oracle as orc
    a, targets, z, prior = _CPU['arch'], _CPU['targets'], _CPU['z'], _CPU['prior']
    t = targets[k]
    orc.get_best_archetype(t.spectra, t.bands, a.wave, a.flux, list(a.full_type), z[k], 2, True,
                           prior=prior)
    return k


def cpu_baseline_archetype(arch, targets, z, prior, seconds):
    """Oracle archetype fit (scipy lsq_linear BVLS per archetype, as the
    reference) on host cores: archetype fits/s over a bounded sample of pairs."""
    import multiprocessing as mp
    ncores = min(CPU_SHARE, len(os.sched_getaffinity(0)))
    _CPU.update(arch=arch, targets=targets, z=z, prior=prior)
    t0 = time.perf_counter()
    _cpu_arch_one(0)
    one = time.perf_counter() - t0
    nsample = min(len(targets), max(ncores, int(seconds * ncores / max(one, 1e-3))))
    ctx = mp.get_context('fork')
    with ctx.Pool(ncores) as pool:
        t0 = time.perf_counter()
        list(pool.imap_unordered(_cpu_arch_one, range(nsample)))
        wall = time.perf_counter() - t0
    return dict(value=nsample * arch.narch / wall, unit='archetype fits/s', cores=ncores, kind='port',
                sample=f'{nsample} (target, z) pairs x {arch.narch} archetypes, float64 oracle '
                       f'(CSR R-apply, normal equations, scipy lsq_linear BVLS, explicit residual), '
                       f'one process per core, {wall:.1f} s wall')


def bench_archetype(args):
    """Archetype final fit: every GALAXY archetype at the 3 lowest minima of
    each target's coarse scan, per-camera Legendre polynomials of degree < 2
    and a 0.1 prior on them (rrdesi's defaults), bounded archetype
    coefficient; one step = ArchetypeFit.run over all pairs (archetype rebin
    into HBM, rr_arch_fit, best archetype per pair)."""
    import numpy as np
    from redrock_amd import synth
    from redrock_amd.archetypes import prior_on_coeffs
    tm = synth.galaxy_template()
    targets, ztrue = synth.make_targets(tm, args.ntargets, seed=args.seed, snr=args.snr)
    arch = synth.archetypes_from(tm, args.narch, seed=args.seed + 1)
    prior = prior_on_coeffs(1, 2, 0.1, 3)
    cpu = None
    if not args.no_cpu:
        log('CPU baseline (oracle archetype fit) ...')
        cpu = cpu_baseline_archetype(arch, targets, ztrue, prior, args.cpu_seconds)
        log(f"CPU baseline: {cpu['value']:.4g} archetype fits/s on {cpu['cores']} cores")
    import torch
    from redrock_amd import zscan
    from redrock_amd import fitz as fz
    from redrock_amd.templates import DistTemplate
    from redrock_amd.archetypes import ArchetypeFit
    from oracle import rr_oracle as orc
    assert torch.cuda.is_available(), 'bench.py needs a GPU'
    dwave = {s.wavehash: s.wave for s in targets[0].spectra}
    nt, zg = len(targets), tm.redshifts
    nz = len(zg)
    dt = DistTemplate(tm, dwave)
    batch = zscan.TargetBatch(targets)
    plan = zscan.ScanPlan(dt.cache, batch, np.arange(nt), np.zeros(nt), np.full(nt, nz),
                          precision='f16x3', oiisum=zscan._oiisum(tm))
    plan.run()
    c2 = (plan.zchi2 + plan.penalty).reshape(nt, nz).cpu().numpy()
    del plan
    tidx, zp = [], []
    for t in range(nt):
        for i in fz.find_minima(c2[t])[:3]:
            tidx.append(t)
            zp.append(zg[i])
    tidx, zp = np.array(tidx), np.array(zp)
    npair = len(tidx)
    fit = ArchetypeFit(arch, targets, per_camera=True, nleg=2, prior=prior, batch=batch)
    events = []
    inner = fit._fit

    def timed_fit(*a, **k):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        r = inner(*a, **k)
        e1.record()
        events.append((e0, e1))
        return r
    fit._fit = timed_fit
    for _ in range(max(1, args.warmup)):
        fit.run(tidx, zp)
    torch.cuda.synchronize()
    events.clear()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        chi2, coeff, ftype = fit.run(tidx, zp)
    torch.cuda.synchronize()
    step = (time.perf_counter() - t0) / args.steps
    fit_ms = sum(a.elapsed_time(b) for a, b in events) / args.steps
    nfit = npair * arch.narch
    log(f'archetype workload: {nt} targets, {npair} pairs x {arch.narch} archetypes, '
        f'{step * 1e3:.1f} ms/step (rr_arch_fit {fit_ms:.1f} ms)')
    # parity on a sample of pairs
    rng = np.random.default_rng(0)
    sample = rng.choice(npair, size=min(8, npair), replace=False)
    errs, same = [], 0
    for k in sample:
        t = targets[tidx[k]]
        wc2, wcf, wty = orc.get_best_archetype(t.spectra, t.bands, arch.wave, arch.flux,
                                               list(arch.full_type), zp[k], 2, True, prior=prior)
        errs.append(abs(chi2[k] - wc2) / wc2)
        same += ftype[k] == wty
    log(f'oracle sample: max rel chi2 err {max(errs):.2e}, same archetype {same}/{len(sample)}')
    # reference-form FLOPs of one archetype fit (the eval formula of the scan
    # with nb = 1 + 6 columns): R-apply, weights, Gram, y, solve, model, chi2
    nb, npix = 7, batch.layout.npix
    nnz = 11 * npix
    flop_fit = (2 * nnz * nb + npix * nb + 2 * npix * nb * nb + 2 * npix * nb
                + (2 * nb ** 3) // 3 + 2 * nb * nb + 2 * npix * nb + 3 * npix)
    print(json.dumps({
        'metric': 'archetype fits/sec on 1 GPU (target x minimum x archetype, BVLS + Legendre)',
        'value': nfit / step, 'unit': 'archetype fits/s', 'n_gpus': 1, 'steps': args.steps,
        'warmup': args.warmup, 'ms_per_step': step * 1e3, 'higher_is_better': True, 'scaling': 'weak',
        'vs_baseline': None, 'dtype': 'f64', 'data': 'synthetic',
        'config': {'workload': f'{nt} GALAXY targets, 3 lowest minima each, {arch.narch} archetypes, '
                               f'per-camera Legendre degree < 2, prior sigma 0.1',
                   'targets': nt, 'pairs': int(npair), 'archetypes': arch.narch},
        'roofline': {'bound': 'fp64', 'achieved': nfit * flop_fit / (fit_ms * 1e-3) / 1e12, 'peak': 78.6,
                     'unit': 'TFLOP/s', 'frac': nfit * flop_fit / (fit_ms * 1e-3) / 78.6e12, 'traffic': None,
                     'note': 'FP64 vector ALU; reference-form FLOPs per fit x fits per launch',
                     'kernel': 'rr_arch_fit (arch_geom_kernel + arch_accum_kernel<1, 6, true> + '
                               'arch_solve_kernel<1, 6>)',
                     'kernel_ms': fit_ms, 'flop_per_fit': int(flop_fit), 'fits_per_launch': int(nfit)},
        'oracle_sample_max_rel_chi2_err': float(max(errs)),
        'oracle_sample_same_archetype': f'{same}/{len(sample)}',
        'cpu_baseline': cpu,
    }), flush=True)


def bench_zfind(args):
    """zfind end to end (scan of every template, FP64 fine fits of the minima,
    result table) on mixed GALAXY/QSO/STAR targets: targets per second."""
    import numpy as np
    import torch
    from redrock_amd import synth, zfind as zf
    from redrock_amd.templates import DistTemplate
    from redrock_amd.targets import DistTargetsCopy
    tms = (synth.galaxy_template(), synth.qso_template(), synth.star_template())
    targets, types, ztrue = synth.mixed_targets(args.zfind_targets, seed=args.seed, snr=args.snr,
                                                mask_frac=(0.0, 0.3), templates=tms)
    assert torch.cuda.is_available(), 'bench.py needs a GPU'
    dwave = {s.wavehash: s.wave for s in targets[0].spectra}
    dts = [DistTemplate(t, dwave) for t in tms]
    dist = DistTargetsCopy(targets)
    zf.zfind(DistTargetsCopy(targets[:64]), dts, nminima=3, precision=args.precision)      # warm-up
    out = {}
    for p in (args.precision, 'f64'):
        torch.cuda.synchronize()
        t = time.perf_counter()
        _, zt = zf.zfind(dist, dts, nminima=3, precision=p)
        torch.cuda.synchronize()
        out[p] = (time.perf_counter() - t, dict(zf.last_timings), zt)
        tm = zf.last_timings
        log(f"{p}: zfind {out[p][0]:.2f} s (scan {tm['scan']:.2f}, fitz "
            + ', '.join(f'{k} {v:.2f}' for k, v in tm['fitz'].items()) + f", assemble {tm['assemble']:.2f})")
    a, b = out[args.precision][2], out['f64'][2]
    best = lambda zt: {t: (zt['z'][i], zt['spectype'][i]) for i, t in enumerate(zt['targetid']) if zt['znum'][i] == 0}
    ba, bb = best(a), best(b)
    same_z = np.mean([abs(ba[t][0] - bb[t][0]) <= 1e-9 * (1 + bb[t][0]) for t in bb])
    same_type = np.mean([ba[t][1] == bb[t][1] for t in bb])
    truth = {i: (z, ty) for i, (z, ty) in enumerate(zip(ztrue, types))}
    right = np.mean([abs(bb[t][0] - truth[t][0]) <= 3e-3 * (1 + truth[t][0]) and bb[t][1] == truth[t][1]
                     for t in bb])
    nt = len(targets)
    wall = out[args.precision][0]
    print(json.dumps({
        'metric': 'targets/sec on 1 GPU, zfind end to end (scan + fine fits + table)',
        'value': nt / wall, 'unit': 'targets/s', 'n_gpus': 1, 'steps': 1, 'warmup': 1,
        'ms_per_step': wall * 1e3, 'higher_is_better': True, 'scaling': 'weak', 'vs_baseline': None,
        'dtype': args.precision, 'data': 'synthetic',
        'config': {'workload': f'{nt} mixed GALAXY/QSO/STAR targets, 0-30% masked, templates GALAXY '
                               f'(nb 10, 1977 z) + QSO (nb 4, 3795 z) + STAR (nb 5, 501 z), nminima 3',
                   'targets': nt},
        'phases_s': out[args.precision][1],
        'fp64': {'value': nt / out['f64'][0], 'seconds': out['f64'][0], 'phases_s': out['f64'][1]},
        'best_z_same_as_fp64': float(same_z), 'best_type_same_as_fp64': float(same_type),
        'fp64_right_z_and_type': float(right),
    }), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=10)
    ap.add_argument('--warmup', type=int, default=2)
    ap.add_argument('--ntargets', type=int, default=500)
    ap.add_argument('--precision', default='f16r', choices=('f16r', 'f16x3', 'f16', 'bf16'))
    ap.add_argument('--snr', type=float, default=1.0)
    ap.add_argument('--seed', type=int, default=0)
    ap.add_argument('--cpu-seconds', type=float, default=12.0)
    ap.add_argument('--no-cpu', action='store_true')
    ap.add_argument('--snr-sweep', default='3,10',
                    help='per-pixel S/N of extra target sets whose chi2 error against FP64 is reported')
    ap.add_argument('--variants', default='f16x3,f16,bf16,f32',
                    help='other mixed precisions to time and compare (comma separated)')
    ap.add_argument('--tile', type=int, default=-1,
                    help='tile configuration of the mixed kernel (-1: autotune)')
    ap.add_argument('--calibrate-copy-mb', type=int, default=0,
                    help='only run a plain device copy of this many MiB (HBM counter calibration)')
    ap.add_argument('--workload', default='galaxy', choices=('galaxy', 'qso', 'mixed', 'grad', 'archetype', 'zfind'),
                    help="galaxy: the 500-target galaxy petal (headline); qso: 500 QSO targets "
                         "with the fused Lyman-alpha IGM; mixed: GALAXY/QSO/STAR targets with "
                         "masked runs over all three templates in one HIP graph; grad: adjoint "
                         "gradient and Newton refinement of the 3 lowest minima per target; "
                         "archetype: the archetype final fit at the 3 lowest minima per target; "
                         "zfind: zfind end to end on mixed targets")
    ap.add_argument('--mixed-targets', type=int, default=5000)
    ap.add_argument('--grad-targets', type=int, default=5000)
    ap.add_argument('--narch', type=int, default=256, help='archetypes of the archetype workload')
    ap.add_argument('--zfind-targets', type=int, default=2000)
    ap.add_argument('--no-polish', action='store_true',
                    help='leave out the FP64 polish of the rows near each minimum after the '
                         'mixed-precision scan')
    args = ap.parse_args()
    if args.calibrate_copy_mb:
        calibrate_copy(args.calibrate_copy_mb)
        return
    if args.gpus != 1:
        print(f'bench.py measures one GPU; --gpus {args.gpus} is not supported '
              f'(targets shard across GPUs with no communication; see the design notes)',
              file=sys.stderr)
        sys.exit(2)

    if args.workload == 'mixed':
        return bench_mixed(args)
    if args.workload == 'grad':
        return bench_grad(args)
    if args.workload == 'archetype':
        return bench_archetype(args)
    if args.workload == 'zfind':
        return bench_zfind(args)

    return bench_petal(args)


def _petal_targets(tm, n, seed, snr):
    from redrock_amd import synth
    return synth.make_targets(tm, n, seed=seed, snr=snr)


def bench_petal(args):
    """The galaxy petal (headline) or the QSO petal: every target x every
    redshift of one template, inputs resident in HBM."""
    import numpy as np
    from redrock_amd import synth

    tm = synth.galaxy_template() if args.workload == 'galaxy' else synth.qso_template()
    kind = tm.template_type
    targets, ztrue = _petal_targets(tm, args.ntargets, args.seed, args.snr)
    dwave = {s.wavehash: s.wave for s in targets[0].spectra}
    nt, nz, nb = len(targets), len(tm.redshifts), tm.nbasis
    evals = nt * nz
    log(f'workload: {nt} targets x {nz} z x {nb} basis, npix={sum(len(w) for w in dwave.values())}')

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(tm, targets, dwave, args.cpu_seconds)
        log(f'cpu baseline: {cpu["value"]:.0f} evals/s on {cpu["cores"]} cores, '
            f'{cpu["value_1core"]:.0f} on 1 core')

    import torch
    from redrock_amd import zscan, _hip, autotune
    from redrock_amd.templates import DistTemplate
    assert torch.cuda.is_available(), 'bench.py needs a GPU'

    torch.cuda.synchronize()
    t0 = time.perf_counter()
    dt = DistTemplate(tm, dwave)
    batch = zscan.TargetBatch(targets)
    oii = tm.OIItemplate.sum(axis=1) if hasattr(tm, 'OIItemplate') else None
    items = (np.arange(nt), np.zeros(nt), np.full(nt, nz))
    plans = {p: zscan.ScanPlan(dt.cache, batch, *items, precision=p, oiisum=oii,
                               polish=(p != 'f64' and not args.no_polish))
             for p in ('f64', args.precision)}
    torch.cuda.synchronize()
    setup_s = time.perf_counter() - t0
    log(f'setup (template caches, target upload, packing): {setup_s:.2f} s')

    def timed(plan, k, stages=0):
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(k):
            plan.run(stages=stages)
        torch.cuda.synchronize()
        return (time.perf_counter() - t) / k

    def kernel_ms(plan, k, stage=_hip.RR_STAGE_MAIN):
        """Average duration of one stage's kernel (HIP events on its stream)."""
        evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
               for _ in range(k)]
        for e0, e1 in evs:
            e0.record()
            plan.run(stages=stage)
            e1.record()
        torch.cuda.synchronize()
        return sum(e0.elapsed_time(e1) for e0, e1 in evs) / k

    mix = plans[args.precision]
    table = autotune.tile_table(args.precision)
    if args.tile < 0:
        tile, tile_ms = autotune.choose(mix, use_cache=False)
        log(f'autotune: tile {tile} {table[tile]} ; ' +
            ', '.join(f'{table[t]}: {ms:.2f} ms' for t, ms in sorted(tile_ms.items())))
    else:
        tile, tile_ms = args.tile, {}
        mix.desc.tile = tile
    rtable = autotune.resid_table() if args.precision == 'f16r' else None
    if rtable is not None:
        # resid_tile 0 with the matrix-core inputs is that kernel's first entry
        rentry = (rtable[mix.desc.resid_tile] if mix.desc.resid_tile or not mix.desc.d_tcache_u
                  else next(r for r in rtable if r[0] == 2))
        rdesc = (RESID_FIELDS[rentry[0]], rentry)
        log(f'residual kernel configuration {mix.desc.resid_tile}: {dict(zip(*rdesc))}')
    for p, plan in plans.items():
        for _ in range(args.warmup):
            plan.run()
    torch.cuda.synchronize()
    t_mix = timed(mix, args.steps)
    k_mix = kernel_ms(mix, args.steps)
    k_resid = kernel_ms(mix, args.steps, _hip.RR_STAGE_RESID) if args.precision == 'f16r' else None
    f64 = plans['f64']
    t_f64 = timed(f64, max(1, args.steps // 2))
    k_f64 = kernel_ms(f64, max(1, args.steps // 2))
    # the redshift-per-lane FP64 kernel (the FP64 scan before the lane-per-target one)
    old = zscan.ScanPlan(dt.cache, batch, *items, precision='f64', oiisum=oii, lanes=False)
    old.run()
    k_old = kernel_ms(old, 2)
    torch.cuda.synchronize()
    del old
    polish = None
    c_raw = None
    if mix.polish is not None:
        # the same plan without its polish: time and accuracy of the bare scan,
        # then the polish stage alone on the bare scan's rows (HIP events)
        bare = _hip.RR_STAGE_STATS | _hip.RR_STAGE_MAIN | _hip.RR_STAGE_FINISH | _hip.RR_STAGE_RESID
        t_raw = timed(mix, args.steps, bare)
        c_raw = mix.zchi2.reshape(nt, nz).cpu().numpy()
        evs = []
        for _ in range(args.steps):
            mix.run(stages=bare)
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            mix.run(stages=zscan.STAGE_POLISH)
            e1.record()
            evs.append((e0, e1))
        torch.cuda.synchronize()
        polish = {'tol': zscan.polish_tol(args.precision, True), 'max_windows': zscan.POLISH_WINDOWS,
                  'window': _hip.RR_POLISH_WINDOW,
                  'ms': sum(a.elapsed_time(b) for a, b in evs) / len(evs),
                  'windows_per_target': float(mix.nwindows.float().mean()),
                  'targets_polished': int((mix.nwindows > 0).sum()),
                  'unpolished_ms_per_step': t_raw * 1e3, 'unpolished_value': evals / t_raw}
    # the per-batch packing (R tiles, lane pack) the step leaves out, and the
    # copy of the results to host numpy arrays (the reference returns numpy)
    torch.cuda.synchronize()
    t = time.perf_counter()
    for _ in range(args.steps):
        batch._mixed.clear()
        batch.mixed(args.precision)        # f16r: the split-fp16 tiles of both its stages
        if args.precision == 'f16r' and mix.resid in ('lanes', 'both'):
            batch.lanes('f16r')
    torch.cuda.synchronize()
    t_pack = (time.perf_counter() - t) / args.steps
    # final full runs so the outputs hold complete results
    mix.run()
    f64.run()
    torch.cuda.synchronize()
    host = mix.host()                      # new host arrays on every call, as for every petal
    t = time.perf_counter()
    for _ in range(args.steps):
        host = mix.host()
    t_d2h = (time.perf_counter() - t) / args.steps

    c_mix = host[0].reshape(nt, nz)
    c_f64 = f64.zchi2.reshape(nt, nz).cpu().numpy()
    # f'Wf per target: the scale of a mixed chi2's error (the sum chi2 cancels against)
    fwf = (batch.wf * batch.f).sum(dim=1).cpu().numpy()
    if polish is not None:
        rr_ = np.abs(c_raw - c_f64) / np.abs(c_f64)
        polish.update({
            'unpolished_best_z_agreement': float(np.mean(c_raw.argmin(axis=1) == c_f64.argmin(axis=1))),
            'unpolished_max_rel_chi2_err': float(rr_.max()),
            'unpolished_max_rel_chi2_err_at_best_z': float(rr_[np.arange(nt), c_f64.argmin(axis=1)].max())})
        log(f'polish: {polish["ms"]:.3f} ms, {polish["targets_polished"]} targets with '
            f'{polish["windows_per_target"] * nt:.0f} windows; '
            f'without it {t_raw * 1e3:.2f} ms/step, best-z agreement '
            f'{polish["unpolished_best_z_agreement"]:.4f}')
    bz_mix, bz_f64 = c_mix.argmin(axis=1), c_f64.argmin(axis=1)
    agree = float(np.mean(bz_mix == bz_f64))
    rel = np.abs(c_mix - c_f64) / np.abs(c_f64)
    rel_best = rel[np.arange(nt), bz_f64]
    zgrid = tm.redshifts
    zerr_ok = float(np.mean(np.abs(zgrid[bz_f64] - ztrue) < 5e-3 * (1 + ztrue)))
    log(f'{args.precision}: {t_mix*1e3:.2f} ms/step (normal equations {k_mix:.2f} ms'
        + (f', residual {k_resid:.2f} ms' if k_resid is not None else '') + f'); '
        f'f64: {t_f64*1e3:.2f} ms/step (kernel {k_f64:.2f} ms; redshift-per-lane kernel {k_old:.2f} ms); '
        f'packing {t_pack*1e3:.2f} ms, results to host {t_d2h*1e3:.1f} ms')
    log(f'best-z agreement {agree:.4f}; max rel chi2 err {rel.max():.2e}, at best z {rel_best.max():.2e}; '
        f'f64 best z within 5e-3(1+z) of truth: {zerr_ok:.3f}')

    # other precisions on the same inputs: speed, chi2 error vs FP64, best-z agreement
    variants = {}
    for p in [v for v in args.variants.split(',') if v and v != args.precision]:
        plan = zscan.ScanPlan(dt.cache, batch, *items, precision=p, oiisum=oii,
                              polish=not args.no_polish)
        plan.run()
        tv = timed(plan, max(2, args.steps // 2))
        plan.run()
        torch.cuda.synchronize()
        cv = plan.zchi2.reshape(nt, nz).cpu().numpy()
        rv = np.abs(cv - c_f64) / np.abs(c_f64)
        variants[p] = {'value': evals / tv, 'ms_per_step': tv * 1e3,
                       'speedup_vs_fp64': t_f64 / tv,
                       'best_z_agreement': float(np.mean(cv.argmin(axis=1) == bz_f64)),
                       'max_rel_chi2_err': float(rv.max()),
                       'median_rel_chi2_err': float(np.median(rv)),
                       'max_chi2_err_over_fwf': float((np.abs(cv - c_f64) / fwf[:, None]).max()),
                       'max_rel_chi2_err_at_best_z': float(rv[np.arange(nt), bz_f64].max())}
        log(f'{p}: {tv*1e3:.2f} ms/step, max rel err {rv.max():.2e}, median {np.median(rv):.2e}, '
            f'best-z agreement {variants[p]["best_z_agreement"]:.4f}')
        del plan

    # error against FP64 at higher per-pixel S/N (same template caches)
    sweep = []
    for snr in [float(v) for v in args.snr_sweep.split(',') if v]:
        tg2, _ = _petal_targets(tm, nt, args.seed + 7, snr)
        b2 = zscan.TargetBatch(tg2)
        res2 = {}
        for p in (args.precision, 'f16x3', 'f64'):
            pl = zscan.ScanPlan(dt.cache, b2, *items, precision=p, oiisum=oii,
                                polish=(p != 'f64' and not args.no_polish))
            pl.run()
            res2[p] = pl.zchi2.reshape(nt, nz).cpu().numpy()
            del pl
        fwf2 = (b2.wf * b2.f).sum(dim=1).cpu().numpy()
        row = {'snr_per_pixel': snr}
        for p in (args.precision, 'f16x3'):
            r2 = np.abs(res2[p] - res2['f64']) / np.abs(res2['f64'])
            row[p] = {'max_rel_chi2_err': float(r2.max()),
                      'max_chi2_err_over_fwf': float((np.abs(res2[p] - res2['f64']) / fwf2[:, None]).max()),
                      'best_z_agreement': float(np.mean(res2[p].argmin(axis=1) == res2['f64'].argmin(axis=1)))}
        sweep.append(row)
        log(f'S/N {snr:g}: ' + '; '.join(f'{p} max rel {row[p]["max_rel_chi2_err"]:.2e}, best-z '
                                        f'{row[p]["best_z_agreement"]:.4f}' for p in (args.precision, 'f16x3')))
        del b2

    flops = FLOP_PER_EVAL[kind] * evals
    sig = kernel_signature(args.precision, nb, table[tile])
    ach = flops / (k_mix * 1e-3)
    npix = batch.layout.npix
    resid_flop = 2 * npix * (nb + 11 + 1)          # u = T c, R-apply, square-accumulate per eval
    resid_mfma = bool(args.precision == 'f16r' and mix.desc.d_tcache_u
                      and (mix.desc.resid_tile == 0 or rtable[mix.desc.resid_tile][0] == 2))
    # matrix-core residual, executed MFMA work per eval: per 16-pixel tile 3
    # 16x16x32 MFMAs over 16 redshifts (R-apply) and nmf over 16 targets (u = T c)
    nmf = (_hip.resid_u_slots(nb)[0] // 32) if resid_mfma else 0
    resid_exec = (batch.layout.npix_pad // 16) * (3 + nmf) * 16384 // 16 if resid_mfma else None
    out = {
        'metric': METRIC,
        'value': evals / t_mix,
        'unit': 'evals/s',
        'n_gpus': 1,
        'steps': args.steps,
        'warmup': args.warmup,
        'ms_per_step': t_mix * 1e3,
        'higher_is_better': True,
        'scaling': 'weak',
        'vs_baseline': None,
        'dtype': 'f16+f32' if args.precision == 'f16r' else args.precision,
        'data': 'synthetic',
        'config': {'workload': (f'galaxy petal: {nt} GALAXY targets x {nz} z x {nb} basis, '
                                if kind == 'GALAXY' else
                                f'{nt} QSO targets x {nz} z x {nb} basis, Lyman-alpha IGM fused, ')
                               + '3 cameras 7958 px, 11-diagonal R',
                   'targets': nt, 'redshifts': nz, 'nbasis': nb, 'snr_per_pixel': args.snr,
                   'precision': args.precision,
                   'solve': 'FP64 Cholesky',
                   'chi2': ('explicit residual sum w (f - R T c)^2 at the fp16 coefficients: u = T c and '
                            'the R-apply as split-fp16 MFMAs (hi + lo operands, fp32 accumulation), '
                            'squares in fp32 per 32 pixels, FP64 sums' if args.precision == 'f16r' else
                            "compensated FP64 f'Wf - 2c'y + c'Mc"),
                   'polish': (None if args.no_polish else
                              f'FP64 re-evaluation of the {_hip.RR_POLISH_WINDOW}-redshift windows within '
                              f'{zscan.polish_tol(args.precision, True):g} of each target minimum '
                              f'(relative to max(f\'Wf, |min|), in the timed step)'),
                   'accumulate': ('fp32 MFMA partial sums over 64 pixels, added in FP64'
                                  if args.precision == 'f16x3' else 'fp32 (MFMA), FP64 after the scan'),
                   'tile': dict(zip(('waves_items', 'waves_z', 'items_per_wave', 'z_per_wave'),
                                    table[tile])),
                   'residual_tile': None if rtable is None else dict(zip(*rdesc))},
        'autotune_ms': {'x'.join(map(str, table[t])): ms for t, ms in tile_ms.items()},
        'fp64': {'value': evals / t_f64, 'ms_per_step': t_f64 * 1e3, 'kernel_ms': k_f64,
                 'kernel': f'void zscan_f64r_kernel<{nb}, {12 if nb <= 10 else 8}>',
                 'roofline_frac': flops / (k_f64 * 1e-3) / PEAK_FLOPS['f64'],
                 'redshift_per_lane_kernel_ms': k_old,
                 'redshift_per_lane_roofline_frac': flops / (k_old * 1e-3) / PEAK_FLOPS['f64']},
        'speedup_vs_fp64': t_f64 / t_mix,
        'best_z_agreement': agree,
        'max_rel_chi2_err': float(rel.max()),
        'max_rel_chi2_err_at_best_z': float(rel_best.max()),
        'max_chi2_err_over_fwf': float((np.abs(c_mix - c_f64) / fwf[:, None]).max()),
        'median_rel_chi2_err': float(np.median(rel)),
        'fp64_best_z_near_truth': zerr_ok,
        'setup_s': setup_s,
        'step_with_packing': {'pack_ms': t_pack * 1e3, 'ms_per_step': (t_mix + t_pack) * 1e3,
                              'value': evals / (t_mix + t_pack)},
        'end_to_end': {'results_to_host_ms': t_d2h * 1e3,
                       'ms_per_petal': (t_mix + t_pack + t_d2h) * 1e3,
                       'value': evals / (t_mix + t_pack + t_d2h),
                       'note': 'packing + scan + copy of zchi2, zcoeff, penalty, status to new host numpy '
                               'arrays (mean of several petals); '
                               'template caches and target upload (setup_s) not included'},
        'roofline': {'bound': 'mfma', 'achieved': ach / 1e12, 'peak': PEAK_FLOPS[args.precision] / 1e12,
                     'unit': 'TFLOP/s', 'frac': ach / PEAK_FLOPS[args.precision],
                     'traffic': (measured_counter('traffic_bytes_per_launch', sig)
                                 if kind == 'GALAXY' and nt == 500 else None),
                     'mfma_busy': (measured_counter('mfma_util_percent', sig) / 100
                                   if kind == 'GALAXY' and nt == 500
                                   and measured_counter('mfma_util_percent', sig) is not None else None),
                     'kernel': sig,
                     'kernel_ms': k_mix,
                     'flop_per_eval': FLOP_PER_EVAL[kind], 'evals_per_launch': evals},
        'roofline_residual': (None if k_resid is None else {
            'bound': 'mfma', 'kernel': f'zscan_resid_mfma_kernel<{nb}, ...>', 'kernel_ms': k_resid,
            'executed_flop_per_eval': resid_exec,
            'achieved': resid_exec * evals / (k_resid * 1e-3) / 1e12, 'peak': PEAK_FLOPS['f16'] / 1e12,
            'unit': 'TFLOP/s', 'frac': resid_exec * evals / (k_resid * 1e-3) / PEAK_FLOPS['f16'],
            'flop_per_eval': resid_flop,
            'algorithmic_tflops': resid_flop * evals / (k_resid * 1e-3) / 1e12,
            'note': ('split-fp16 MFMA work executed (u = T c over 16 targets, R-apply over 16 '
                     'redshifts, hi/lo products); flop_per_eval is the FP32 form (nb + 12 FMAs '
                     'per pixel)')} if resid_mfma else {
            'bound': 'fp32 vector ALU', 'kernel': f'zscan_resid_kernel<{nb}, ...>', 'kernel_ms': k_resid,
            'flop_per_eval': resid_flop, 'achieved': resid_flop * evals / (k_resid * 1e-3) / 1e12,
            'peak': PEAK_FLOPS['f32'] / 1e12, 'unit': 'TFLOP/s',
            'frac': resid_flop * evals / (k_resid * 1e-3) / PEAK_FLOPS['f32'],
            'note': 'peak counts packed FP32; the kernel issues unpacked FMAs (half that rate)'}),
        'polish': polish,
        'variants': variants,
        'snr_sweep': sweep,
        'cpu_baseline': cpu,
    }
    print(json.dumps(out), flush=True)

if __name__ == '__main__':
    main()
