"""Tile-shape autotuner for the scan kernels.

The kernels have several tile configurations (items x redshifts per
wavefront and workgroup, rr_zscan_tile_table; the explicit-residual kernel of
'f16r', rr_zscan_resid_tiles).  tune() times every configuration on a
prepared ScanPlan — the kernel's stage only, HIP events on the launch stream —
and returns the fastest; choose() caches the winners per (nbasis, padded
pixels, precision, item count bucket, uniform z) in a JSON file so later
plans of the same shape reuse them without timing.

Cache file: $RR_TUNE_CACHE, default ~/.cache/redrock_amd/tiles.json.
"""
from __future__ import annotations

import ctypes
import json
import os

from . import _hip


def tile_table(precision=_hip.RR_PREC_F16):
    """[(wa, wb, tw, zw), ...] indexed by rr_zscan_desc.tile (entry 0 = default)
    for the kernel of `precision` (an RR_PREC_* code or a PRECISIONS name);
    empty for modes with a single configuration.

    A workgroup covers wa*tw items x wb*zw redshifts.
    """
    if isinstance(precision, str):
        precision = _hip.PRECISIONS[precision]
    n = 32
    cfg = (ctypes.c_int * (4 * n))()
    k = _hip.lib().rr_zscan_tile_table(int(precision), cfg, n)
    if k < 0:
        _hip.check(k, 'rr_zscan_tile_table')
    return [tuple(cfg[4 * i:4 * i + 4]) for i in range(min(k, n))]


def resid_table():
    """Configurations of the explicit-residual stage of precision 'f16r',
    indexed by rr_zscan_desc.resid_tile (entry 0 = default): the FP32 lane
    kernel's (1, waves, 64 items per wave, redshifts per wave), then the
    matrix-core kernel's (2, waves, 16 items, redshifts per workgroup)."""
    n = 32
    cfg = (ctypes.c_int * (4 * n))()
    k = _hip.lib().rr_zscan_resid_tiles(cfg, n)
    if k < 0:
        _hip.check(k, 'rr_zscan_resid_tiles')
    return [tuple(cfg[4 * i:4 * i + 4]) for i in range(min(k, n))]


def _time_stage(plan, stage, reps):
    import torch
    plan.run(stages=stage)   # warm-up (code load, caches)
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
           for _ in range(reps)]
    for e0, e1 in evs:
        e0.record()
        plan.run(stages=stage)
        e1.record()
    torch.cuda.synchronize()
    return min(e0.elapsed_time(e1) for e0, e1 in evs)


def tune_resid(plan, reps=3):
    """Time every residual-kernel configuration of an 'f16r' plan (after a full
    run, so the coefficients exist); returns (best, {resid_tile: ms})."""
    table = resid_table()
    saved = plan.desc.resid_tile
    plan.run()
    # configurations whose inputs the plan holds (ScanPlan resid=)
    have = {1: bool(plan.desc.d_rlane), 2: bool(plan.desc.d_tcache_u)}
    timings = {t: _time_stage(_with(plan, 'resid_tile', t), _hip.RR_STAGE_RESID, reps)
               for t in range(1, len(table)) if have.get(table[t][0], False)}
    plan.desc.resid_tile = saved
    return min(timings, key=timings.get), timings


def _with(plan, field, value):
    setattr(plan.desc, field, value)
    return plan


def _cache_path():
    return os.environ.get('RR_TUNE_CACHE',
                          os.path.join(os.path.expanduser('~'), '.cache', 'redrock_amd', 'tiles.json'))


def _key(plan):
    d = plan.desc
    bucket = 1
    while bucket < d.nitems:
        bucket *= 2
    return f'nb{d.nb}-p{d.npix_pad}-prec{d.precision}-items{bucket}-u{d.uniform_z}'


def tune(plan, reps=3):
    """Time every tile configuration on `plan`; returns (best_tile, {tile: ms})."""
    import torch
    table = tile_table(plan.desc.precision)
    if len(table) < 2:
        return 0, dict()
    saved_tile = plan.desc.tile
    timings = dict()
    plan.run()   # stats + a full pass so the workspace holds valid statistics
    for t in range(1, len(table)):
        wa, wb, tw, zw = table[t]
        if wa * tw > 1 and not plan.desc.uniform_z:
            continue
        plan.desc.tile = t
        plan.run(stages=_hip.RR_STAGE_MAIN)   # warm-up (code load, caches)
        evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
               for _ in range(reps)]
        for e0, e1 in evs:
            e0.record()
            plan.run(stages=_hip.RR_STAGE_MAIN)
            e1.record()
        torch.cuda.synchronize()
        timings[t] = min(e0.elapsed_time(e1) for e0, e1 in evs)
    plan.desc.tile = saved_tile
    best = min(timings, key=timings.get)
    return best, timings


def choose(plan, reps=3, use_cache=True):
    """Best tile for plan's shape, from the cache or by timing; sets plan.desc.tile."""
    path = _cache_path()
    key = _key(plan)
    cache = dict()
    if use_cache and os.path.exists(path):
        try:
            with open(path) as fh:
                cache = json.load(fh)
        except (OSError, ValueError):
            cache = dict()
    resid = plan.desc.precision == _hip.RR_PREC_F16R
    if key in cache and (not resid or 'resid_tile' in cache[key]):
        plan.desc.tile = int(cache[key]['tile'])
        if resid:
            plan.desc.resid_tile = int(cache[key]['resid_tile'])
        return plan.desc.tile, cache[key].get('ms', {})
    best, timings = tune(plan, reps)
    plan.desc.tile = best
    cache[key] = dict(tile=best, ms={str(k): v for k, v in timings.items()})
    if resid:
        rbest, rtimes = tune_resid(plan, reps)
        plan.desc.resid_tile = rbest
        cache[key].update(resid_tile=rbest, resid_ms={str(k): v for k, v in rtimes.items()})
    try:
        os.makedirs(os.path.dirname(path), exist_ok=True)
        with open(path + '.tmp', 'w') as fh:
            json.dump(cache, fh, indent=1)
        os.replace(path + '.tmp', path)
    except OSError:
        pass
    return best, timings
