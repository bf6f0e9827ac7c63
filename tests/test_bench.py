"""bench.py's contract that can be checked without a GPU.

* more than one GPU is refused with a message and exit status 2, before any
  device work;
* the committed counter summary covers every tile configuration the
  autotuner can choose for the headline's fp16 normal-equation kernel on the
  galaxy petal, so the line's roofline.traffic and roofline.mfma_busy are
  filled whichever tile wins;
* the algorithmic FLOPs per evaluation follow the reference-form count.
"""
import os
import subprocess
import sys

from tests.conftest import ROOT


def test_several_gpus_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, 'bench.py'), '--gpus', '2'],
                       capture_output=True, text=True, timeout=120, cwd=ROOT)
    assert r.returncode == 2
    assert 'one GPU' in r.stderr
    assert r.stdout == ''


def test_counters_cover_every_autotuned_tile():
    sys.path.insert(0, ROOT)
    import bench
    from redrock_amd import autotune
    table = autotune.tile_table('f16r')
    assert table
    for cfg in table:
        sig = bench.kernel_signature('f16r', 10, cfg)
        assert bench.measured_counter('traffic_bytes_per_launch', sig) is not None, sig
        assert bench.measured_counter('mfma_util_percent', sig) is not None, sig


def test_flop_per_eval_formula():
    """F = 2 nnzR nb + npix nb + 2 npix nb^2 + 2 npix nb + (2/3) nb^3 + 2 nb^2
    + 2 npix nb + 3 npix, with nnzR the band entries of the three cameras."""
    sys.path.insert(0, ROOT)
    import bench
    cams = (2751, 2326, 2881)
    npix = sum(cams)
    nnz = sum(11 * n - 30 for n in cams)          # 11 diagonals less the corners
    for kind, nb in (('GALAXY', 10), ('QSO', 4), ('STAR', 5)):
        f = (2 * nnz * nb + npix * nb + 2 * npix * nb * nb + 2 * npix * nb
             + round(2 * nb ** 3 / 3) + 2 * nb * nb + 2 * npix * nb + 3 * npix)
        assert abs(bench.FLOP_PER_EVAL[kind] - f) <= 1, (kind, f)
