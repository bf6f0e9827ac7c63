"""The C-ABI library loads and exports every entry point include/*.h declares.

No compute calls: these run without a GPU.
"""
import ctypes
import glob
import os
import re
import subprocess

import pytest

from tests.conftest import ROOT

HEADERS = glob.glob(os.path.join(ROOT, 'include', '*.h'))


def declared_functions():
    names = []
    for h in HEADERS:
        text = open(h).read()
        text = re.sub(r'/\*.*?\*/', '', text, flags=re.S)
        for m in re.finditer(r'^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b(rr_\w+)\s*\(', text, re.M):
            names.append(m.group(1))
    return sorted(set(names))


def test_header_declares_entry_points():
    names = declared_functions()
    for n in ('rr_abi_version', 'rr_last_error', 'rr_rebin', 'rr_zscan',
              'rr_zscan_workspace_bytes', 'rr_pack_mixed'):
        assert n in names


def test_library_exports_every_declared_symbol():
    from redrock_amd import _hip
    assert os.path.exists(_hip.LIB_PATH), 'build the library first (python -m redrock_amd.build)'
    out = subprocess.run(['nm', '-D', '--defined-only', _hip.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if line.strip()}
    missing = [n for n in declared_functions() if n not in exported]
    assert not missing, missing
    assert set(_hip.EXPORTS) <= set(declared_functions())


def test_abi_version_and_error_text():
    from redrock_amd import _hip
    lib = _hip.lib()
    assert lib.rr_abi_version() == _hip.RR_ABI_VERSION
    # argument validation happens before any device call
    d = _hip.ZscanDesc()
    d.nitems = 1
    d.npix = 0
    rc = lib.rr_zscan(ctypes.byref(d), None)
    assert rc == -1
    assert b'bad pixel sizes' in lib.rr_last_error()
    assert lib.rr_zscan_workspace_bytes(None) < 0


def test_descriptor_matches_header_field_order():
    """ctypes struct mirrors rr_zscan_desc field for field."""
    from redrock_amd import _hip
    text = open(os.path.join(ROOT, 'include', 'redrock_hip.h')).read()
    body = text[text.index('typedef struct {\n    int nitems;'):text.index('} rr_zscan_desc;')]
    body = re.sub(r'/\*.*?\*/', '', body, flags=re.S)
    fields = re.findall(r'(\w+)\s*;', body)
    assert fields == [f for f, _ in _hip.ZscanDesc._fields_]


def test_workspace_size_formula():
    from redrock_amd import _hip
    lib = _hip.lib()
    d = _hip.ZscanDesc()
    d.nitems, d.nb, d.total_rows = 3, 10, 1000
    d.precision = _hip.RR_PREC_F64
    ne = 10 * 11 - 45
    assert lib.rr_zscan_workspace_bytes(ctypes.byref(d)) == 256 + ((ne * 1000 * 8 + 255) // 256) * 256
    d.precision = _hip.RR_PREC_F16
    ne = 10 * 12 - 45
    assert lib.rr_zscan_workspace_bytes(ctypes.byref(d)) == 256 + ((ne * 1000 * 8 + 255) // 256) * 256
