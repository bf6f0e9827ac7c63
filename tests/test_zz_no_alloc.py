"""rr_zscan_noalloc (workspace in the library's static device array) against
rr_zscan with a caller workspace: bit-identical outputs on two calls in a row,
the second smaller (the array keeps the first call's contents), in FP64 and
in the fp16 + explicit-residual mode; a scan too large for the array is
refused before any launch.
"""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _outputs(plan):
    return (plan.zchi2.cpu().numpy(), plan.zcoeff.cpu().numpy(), plan.penalty.cpu().numpy(),
            plan.status.cpu().numpy())


@pytest.mark.parametrize('precision', ['f64', 'f16r'])
def test_noalloc_equals_zscan(gpu, precision):
    import torch
    from redrock_amd import synth, zscan, _hip
    from redrock_amd.templates import DistTemplate
    tm = synth.galaxy_template()
    tm._redshifts = tm.redshifts[400:720]
    targets, _ = synth.make_targets(tm, 40, seed=17)
    for s in targets[3].spectra:       # a target without weight: HUGE_CHI2 rows
        s.ivar[:] = 0.0
    dwave = {s.wavehash: s.wave for s in targets[0].spectra}
    dt = DistTemplate(tm, dwave)
    oii = tm.OIItemplate.sum(axis=1)
    for nt, nz in ((40, 320), (12, 200)):     # the second call is smaller
        batch = zscan.TargetBatch(targets[:nt])
        items = (np.arange(nt), np.zeros(nt), np.full(nt, nz))
        print(f'case: rr_zscan_noalloc {precision}, {nt} targets x {nz} z', flush=True)
        ref = zscan.ScanPlan(dt.cache, batch, *items, precision=precision, oiisum=oii, polish=False)
        nof = zscan.ScanPlan(dt.cache, batch, *items, precision=precision, oiisum=oii, polish=False)
        need = _hip.lib().rr_zscan_workspace_bytes(ctypes.byref(nof.desc))
        assert 0 < need <= _hip.RR_NOALLOC_SCRATCH_BYTES
        ref.run()
        rc = _hip.lib().rr_zscan_noalloc(ctypes.byref(nof.desc), _hip.stream_handle())
        _hip.check(rc, 'rr_zscan_noalloc')
        torch.cuda.synchronize()
        for a, b in zip(_outputs(ref), _outputs(nof)):
            assert np.array_equal(a, b)
        assert np.all(ref.zchi2.cpu().numpy().reshape(nt, nz)[3] == zscan.HUGE_CHI2)


def test_noalloc_rejects_large_scan(gpu):
    from redrock_amd import zscan, _hip
    d = _hip.ZscanDesc()
    d.nitems, d.npix, d.npix_pad, d.nb, d.halo = 500, 7958, 7968, 10, 8
    d.precision = _hip.RR_PREC_F16
    d.total_rows = 500 * 1977
    rc = _hip.lib().rr_zscan_noalloc(ctypes.byref(d), None)
    assert rc == -1
    assert 'static scratch' in _hip.lib().rr_last_error().decode()
