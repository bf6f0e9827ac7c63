"""Build libredrock_hip.so for gfx950 (hipcc, in-tree).

    python -m redrock_amd.build [--debug] [--incremental]

Compiles every redrock_amd/csrc/*.hip to an object under redrock_amd/_build/
(in parallel) and links redrock_amd/libredrock_hip.so.
"""
from __future__ import annotations

import concurrent.futures
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, 'csrc')
BUILD = os.path.join(HERE, '_build')
LIB = os.path.join(HERE, 'libredrock_hip.so')
HIPCC = os.environ.get('HIPCC', '/opt/rocm/bin/hipcc')

# Packed FP32 instructions are disabled for all device code.  MFMA results are
# kept in VGPRs (the scan kernel converts every R-apply result with VALU ops,
# and the AGPR form would copy each one out and back).
COMMON = ['--offload-arch=gfx950', '-std=c++17', '-fPIC',
          '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops',
          '-mllvm', '-amdgpu-mfma-vgpr-form',
          '-I', os.path.join(os.path.dirname(HERE), 'include'),
          '-Wno-unused-command-line-argument']


def _stale(src, obj):
    """True when obj is missing or older than src or any shared header."""
    if not os.path.exists(obj):
        return True
    heads = [os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith('.hpp')]
    heads.append(os.path.join(os.path.dirname(HERE), 'include', 'redrock_hip.h'))
    t = os.path.getmtime(obj)
    return any(os.path.getmtime(f) > t for f in [src] + heads)


def _compile(src, opt, incremental=False):
    obj = os.path.join(BUILD, os.path.basename(src) + '.o')
    if incremental and not _stale(src, obj):
        return obj
    cmd = [HIPCC] + COMMON + [opt, '-c', src, '-o', obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f'hipcc failed for {src}:\n{r.stderr}')
    return obj


def build(debug=False, verbose=True, incremental=False):
    """Compile every source (incremental: only those older than their object
    or a shared header) and link the library."""
    os.makedirs(BUILD, exist_ok=True)
    srcs = sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith('.hip'))
    opt = '-O1' if debug else '-O3'
    with concurrent.futures.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(lambda s: _compile(s, opt, incremental), srcs))
    tmp = LIB + '.tmp'
    cmd = [HIPCC, '--offload-arch=gfx950', '-shared', '-fPIC', '-o', tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f'link failed:\n{r.stderr}')
    os.replace(tmp, LIB)
    if verbose:
        print(f'built {LIB}')
    return LIB


if __name__ == '__main__':
    build(debug='--debug' in sys.argv, incremental='--incremental' in sys.argv)
