"""The C-ABI library loads and exports every entry point include/*.h declares.

No compute calls: these run without a GPU.
"""
import ctypes
import glob
import os
import re
import subprocess

import numpy as np
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


def test_polish_argument_errors():
    """rr_zscan_polish validates before any device call; an empty scan is a no-op."""
    from redrock_amd import _hip
    lib = _hip.lib()
    assert lib.rr_zscan_polish(None, None, None) == -1
    assert b'null descriptor' in lib.rr_last_error()
    d = _hip.ZscanDesc()
    d.nitems, d.npix, d.npix_pad, d.nb, d.halo = 0, 10, 32, 3, 8
    d.cache_rows = 64
    d.precision = _hip.RR_PREC_F16X3
    assert lib.rr_zscan_polish(ctypes.byref(d), None, None) == 0
    d.nitems, d.max_item_nz, d.total_rows = 1, 5, 5
    d.d_tcache = d.d_rband = d.d_w = d.d_wf = d.d_f = 8
    d.d_item_target = d.d_item_z0 = d.d_item_nz = d.d_item_out = d.d_zchi2 = d.d_zcoeff = 8
    d.d_rtile = d.d_scales = 8
    d.d_work, d.work_bytes = 8, 1 << 20
    assert lib.rr_zscan_polish(ctypes.byref(d), None, None) == -1
    assert b'null polish descriptor' in lib.rr_last_error()
    d.solve = _hip.RR_SOLVE_NNLS                   # NNLS scans are not polished
    assert lib.rr_zscan_polish(ctypes.byref(d), None, None) == -3
    assert b'NNLS' in lib.rr_last_error()
    d.solve = _hip.RR_SOLVE_PCA
    p = _hip.PolishDesc()
    p.d_tcache64 = p.d_windows = p.d_nwindows = 8
    assert lib.rr_zscan_polish(ctypes.byref(d), ctypes.byref(p), None) == -1
    assert b'null required pointer' in lib.rr_last_error()
    p.d_queue = 8
    p.cache_ldz, p.tol, p.max_windows = 0, 1e-4, 8
    d.cache_nz = 5
    assert lib.rr_zscan_polish(ctypes.byref(d), ctypes.byref(p), None) == -1
    assert b'cache_ldz' in lib.rr_last_error()
    p.cache_ldz = 64
    for tol in (-1.0, float('nan'), float('inf')):
        p.tol = tol
        assert lib.rr_zscan_polish(ctypes.byref(d), ctypes.byref(p), None) == -1
        assert b'tol' in lib.rr_last_error()
    p.tol = 1e-4
    for mw in (0, _hip.RR_POLISH_MAX_WINDOWS + 1):
        p.max_windows = mw
        assert lib.rr_zscan_polish(ctypes.byref(d), ctypes.byref(p), None) == -1
        assert b'max_windows' in lib.rr_last_error()
    # an FP64 scan has nothing to polish
    d.precision = _hip.RR_PREC_F64
    d.cache_ldz = 64
    assert lib.rr_zscan_polish(ctypes.byref(d), ctypes.byref(p), None) == 0


def test_solve_argument():
    """Unknown solve codes are refused before any device call; the method names
    map as the reference's solve_matrices does (zscan.py:1107-1205)."""
    from redrock_amd import _hip, zscan
    assert zscan.solve_code(None) == _hip.RR_SOLVE_PCA
    assert zscan.solve_code('pca') == _hip.RR_SOLVE_PCA
    assert zscan.solve_code('NMF') == zscan.solve_code('NNLS') == _hip.RR_SOLVE_NNLS
    assert zscan.solve_spec('NMF', 4) == (_hip.RR_SOLVE_NNLS, 0b1111, None)
    # BVLS with the archetype bounds (zscan.py:289-294): archetype >= 0,
    # Legendre terms free, the masked NNLS solve; no bounds at all = every
    # coefficient free
    inf = np.inf
    assert zscan.solve_spec('BVLS', 4, [[0, 0, -inf, -inf], [inf] * 4]) == (_hip.RR_SOLVE_NNLS, 0b0011, None)
    assert zscan.solve_spec('bvls', 3) == (_hip.RR_SOLVE_NNLS, 0, None)
    assert zscan.solve_spec('BVLS', 3, (0, inf)) == (_hip.RR_SOLVE_NNLS, 0b111, None)
    assert zscan.solve_spec('BVLS', 3, (-inf, inf)) == (_hip.RR_SOLVE_NNLS, 0, None)
    # any other bounds: the two-sided BVLS solve with the bounds array
    code, mask, b = zscan.solve_spec('BVLS', 2, [[0, 0], [1, inf]])
    assert code == _hip.RR_SOLVE_BVLS and b.shape == (2, 2) and b.dtype == np.float64
    assert np.array_equal(b, [[0, 0], [1, inf]])
    code, _, b = zscan.solve_spec('BVLS', 3, (-1.0, 2.0))        # broadcast as lsq_linear
    assert code == _hip.RR_SOLVE_BVLS and np.array_equal(b, [[-1.0] * 3, [2.0] * 3])
    for bad in ([[0, 1], [0, 2]], [[0, np.nan], [1, 2]], [[2, 0], [1, 1]]):
        with pytest.raises(ValueError, match='strictly less'):   # lsq_linear's rule
            zscan.solve_spec('BVLS', 2, bad)
    with pytest.raises(ValueError):
        zscan.solve_spec('BVLS', 3, [[0, 0], [inf, inf]])
    with pytest.raises(NotImplementedError):
        zscan.solve_code('LSQ')
    lib = _hip.lib()
    d = _hip.ZscanDesc()
    d.nitems, d.npix, d.npix_pad, d.nb, d.halo, d.cache_rows = 1, 10, 32, 3, 8, 64
    d.max_item_nz, d.total_rows, d.cache_nz, d.cache_ldz = 5, 5, 5, 64
    d.d_tcache = d.d_rband = d.d_w = d.d_wf = d.d_f = 8
    d.d_item_target = d.d_item_z0 = d.d_item_nz = d.d_item_out = d.d_zchi2 = d.d_zcoeff = 8
    d.d_work, d.work_bytes = 8, 1 << 20
    d.solve = 7
    assert lib.rr_zscan(ctypes.byref(d), None) == -1
    assert b'unknown solve' in lib.rr_last_error()
    d.solve = _hip.RR_SOLVE_BVLS                       # BVLS without its bounds
    assert lib.rr_zscan(ctypes.byref(d), None) == -1
    assert b'd_bounds' in lib.rr_last_error()


def test_polish_tolerance_argument():
    import math
    from redrock_amd import zscan
    assert zscan.polish_tol('f64', None) is None
    assert zscan.polish_tol('f16x3', False) is None
    assert zscan.polish_tol('bf16', True) == zscan.POLISH_TOL['bf16']
    assert zscan.polish_tol('f16', 0.25) == 0.25
    for bad in (-1.0, math.inf, math.nan):
        with pytest.raises(ValueError):
            zscan.polish_tol('f16', bad)


def test_descriptor_matches_header_field_order():
    """ctypes struct mirrors rr_zscan_desc field for field."""
    from redrock_amd import _hip
    text = open(os.path.join(ROOT, 'include', 'redrock_hip.h')).read()
    body = text[text.index('typedef struct {\n    int nitems;'):text.index('} rr_zscan_desc;')]
    body = re.sub(r'/\*.*?\*/', '', body, flags=re.S)
    fields = re.findall(r'(\w+)\s*;', body)
    assert fields == [f for f, _ in _hip.ZscanDesc._fields_]


def test_polish_descriptor_matches_header_field_order():
    from redrock_amd import _hip
    text = open(os.path.join(ROOT, 'include', 'redrock_hip.h')).read()
    body = text[text.index('typedef struct {\n    const double *d_tcache64;'):text.index('} rr_polish_desc;')]
    body = re.sub(r'/\*.*?\*/', '', body, flags=re.S)
    fields = re.findall(r'(\w+)\s*;', body)
    assert fields == [f for f, _ in _hip.PolishDesc._fields_]


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


def test_resid_u_operand_layout():
    """The matrix-core residual's template operand (zscan._u_operand, read by
    zscan_resid_mfma.hip): per (z, row) KS fp16 slots hi | lo | hi [| lo] of
    each basis value, zero beyond, hi + lo within 2^-22 of the float32 value
    (and 2^-25 absolute where lo is an fp16 subnormal);
    KS and the product count as rr_resid_u_slots states for every nb."""
    import torch
    from redrock_amd import _hip, zscan
    rng = np.random.default_rng(3)
    for nb in range(1, _hip.RR_MAX_NBASIS + 1):
        ks, nprod = _hip.resid_u_slots(nb)
        assert nprod == (4 if 4 * nb <= 32 else 3 if 3 * nb <= 32 else 4)
        assert ks == 32 * ((nprod * nb + 31) // 32)
        full = torch.from_numpy(rng.normal(size=(3, 7, nb)).astype(np.float32) * 100)
        buf = zscan._u_operand(full).reshape(3, 7, ks).double().numpy()
        f = full.numpy().astype(np.float64)
        hi, lo = buf[..., :nb], buf[..., nb:2 * nb]
        # 2^-22 relative, and half the fp16 subnormal spacing (2^-25) where lo is subnormal
        assert np.all(np.abs(hi + lo - f) <= 2.0 ** -22 * np.abs(f) + 2.0 ** -25)
        assert np.array_equal(buf[..., 2 * nb:3 * nb], hi)
        if nprod == 4:
            assert np.array_equal(buf[..., 3 * nb:4 * nb], lo)
        assert np.all(buf[..., nprod * nb:] == 0)
    with pytest.raises(Exception):
        _hip.resid_u_slots(0)


def test_gather_rows_copies_every_row():
    """rr_gather_rows (host code, no GPU): rows from separate arrays land in
    order in one buffer, for thread counts below, at and above the row count."""
    import ctypes
    import numpy as np
    from redrock_amd import _hip
    rng = np.random.default_rng(4)
    rows = [rng.normal(size=(3, 37)) for _ in range(23)]
    ptrs = np.array([r.ctypes.data for r in rows], dtype=np.uint64)
    for nthreads in (1, 4, 23, 100):
        out = np.zeros((23, 3, 37))
        rc = _hip.lib().rr_gather_rows(ptrs.ctypes.data, 23, rows[0].nbytes, out.ctypes.data, nthreads)
        assert rc == 0
        assert np.array_equal(out, np.stack(rows))
    assert _hip.lib().rr_gather_rows(ptrs.ctypes.data, 0, 8, None, 4) == 0
    bad = ptrs.copy()
    bad[5] = 0
    out = np.zeros((23, 3, 37))
    assert _hip.lib().rr_gather_rows(bad.ctypes.data, 23, rows[0].nbytes, out.ctypes.data, 4) < 0
    assert b'row 5' in _hip.lib().rr_last_error()
    assert _hip.lib().rr_gather_rows(ptrs.ctypes.data, -1, 8, out.ctypes.data, 4) < 0
