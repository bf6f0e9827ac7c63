"""XCD tile map (redrock_amd/csrc/xcd_map.hpp) checked on the host: every
(item tile, z tile) is launched exactly once and each super-tile stays on
one XCD, over a sweep of grid shapes and residencies."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.mark.skipif(shutil.which('hipcc') is None and not os.path.exists('/opt/rocm/bin/hipcc'),
                    reason='hipcc not available')
def test_tile_map_is_a_bijection(tmp_path):
    hipcc = shutil.which('hipcc') or '/opt/rocm/bin/hipcc'
    exe = str(tmp_path / 'tilemap_check')
    r = subprocess.run([hipcc, '-std=c++17', '-O1', os.path.join(HERE, 'tilemap_check.cpp'), '-o', exe,
                        '-Wno-unused-command-line-argument'], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-2000:]
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert 'tile map ok' in r.stdout
