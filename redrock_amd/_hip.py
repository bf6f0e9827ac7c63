"""ctypes binding of libredrock_hip.so (declared in include/redrock_hip.h).

The library is built in-tree by ``python -m redrock_amd.build``.  There is no
fallback: if it is missing or cannot be loaded, every scan entry point raises.
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
# RR_LIB_PATH: load the library from another path (a build kept elsewhere)
LIB_PATH = os.environ.get('RR_LIB_PATH') or os.path.join(HERE, 'libredrock_hip.so')

RR_ABI_VERSION = 14
RR_F64, RR_F32, RR_F16, RR_BF16 = 0, 1, 2, 3
RR_PREC_F64, RR_PREC_F16, RR_PREC_BF16, RR_PREC_F32, RR_PREC_F16X3, RR_PREC_F16R = 0, 1, 2, 3, 4, 5
RR_STATUS_OK, RR_STATUS_SINGULAR = 0, 1
RR_MAX_NBASIS = 12
RR_MAX_LYMAN_LINES = 8

PRECISIONS = {'f64': RR_PREC_F64, 'f16': RR_PREC_F16, 'bf16': RR_PREC_BF16, 'f32': RR_PREC_F32,
              'f16x3': RR_PREC_F16X3, 'f16r': RR_PREC_F16R}

EXPORTS = ('rr_abi_version', 'rr_last_error', 'rr_rebin', 'rr_zscan',
           'rr_zscan_workspace_bytes', 'rr_pack_mixed', 'rr_zscan_tiles', 'rr_zscan_tile_table',
           'rr_zscan_grad', 'rr_arch_prep', 'rr_arch_workspace_bytes', 'rr_arch_fit',
           'rr_zscan_polish', 'rr_pack_lanes', 'rr_pack_lanes_bytes', 'rr_zscan_resid_tiles',
           'rr_zscan_noalloc', 'rr_grad_workspace_bytes', 'rr_resid_u_slots', 'rr_minima_head',
           'rr_gather_rows')
RR_NOALLOC_SCRATCH_BYTES = 64 << 20
RR_ARCH_MAX_NA = 4
RR_ARCH_MAX_NLEG = 12

_vp = ctypes.c_void_p
_i = ctypes.c_int
_ll = ctypes.c_longlong
_d = ctypes.c_double


class LymanLine(ctypes.Structure):
    _fields_ = [('line', _d), ('A', _d), ('B', _d)]


class ZscanDesc(ctypes.Structure):
    _fields_ = [
        ('nitems', _i), ('npix', _i), ('npix_pad', _i), ('nb', _i), ('precision', _i),
        ('halo', _i),
        ('d_tcache', _vp), ('cache_rows', _ll), ('cache_ldz', _ll), ('cache_nz', _i),
        ('tscale', _d),
        ('d_rband', _vp), ('d_w', _vp), ('d_wf', _vp), ('d_f', _vp),
        ('d_rtile', _vp), ('d_sfcol', _vp), ('d_scales', _vp),
        ('d_tcache_r', _vp), ('d_rlane', _vp), ('uniform_z', _i), ('tile', _i), ('resid_tile', _i),
        ('d_item_target', _vp), ('d_item_z0', _vp), ('d_item_nz', _vp), ('d_item_out', _vp),
        ('d_prior', _vp), ('d_oiisum', _vp),
        ('d_zchi2', _vp), ('d_zcoeff', _vp), ('d_penalty', _vp), ('d_status', _vp),
        ('max_item_nz', _i), ('total_rows', _ll),
        ('d_work', _vp), ('work_bytes', _ll),
        ('stages', _i), ('solve', _i), ('nonneg_mask', _i), ('d_bounds', _vp),
        ('d_tcache_u', _vp),
    ]


RR_STAGE_STATS, RR_STAGE_MAIN, RR_STAGE_FINISH, RR_STAGE_RESID = 1, 2, 4, 8
RR_SOLVE_PCA, RR_SOLVE_NNLS, RR_SOLVE_BVLS = 0, 1, 2

RR_POLISH_WINDOW = 4
RR_POLISH_MAX_WINDOWS = 64
RR_POLISH_MAX_NZ = 16384


class PolishDesc(ctypes.Structure):
    _fields_ = [
        ('d_tcache64', _vp), ('cache_ldz', _ll), ('tol', _d), ('max_windows', _i),
        ('d_windows', _vp), ('d_nwindows', _vp), ('d_queue', _vp),
    ]


class ArchDesc(ctypes.Structure):
    _fields_ = [
        ('npair', _i), ('nprob', _i), ('na', _i), ('nleg', _i), ('nbound', _i),
        ('npix_pad', _i), ('halo', _i),
        ('d_pair_target', _vp), ('d_acache', _vp), ('ld', _ll), ('narch', _i), ('d_columns', _vp),
        ('d_rband', _vp), ('d_w', _vp), ('d_wf', _vp),
        ('d_wrl', _vp), ('d_gll', _vp), ('d_yl', _vp), ('d_fwf', _vp),
        ('d_prior', _vp), ('d_wave', _vp), ('d_zigm', _vp),
        ('lines', ctypes.POINTER(LymanLine)), ('nlines', _i),
        ('d_coeff', _vp), ('d_chi2', _vp), ('d_work', _vp), ('work_bytes', _ll),
        ('d_x', _vp), ('nx', _i), ('d_yt', _vp), ('ldy', _ll), ('d_pe_lo', _vp), ('d_pe_hi', _vp),
        ('npix', _i), ('d_zrebin', _vp),
    ]


class HipError(RuntimeError):
    pass


_lib = None


def lib():
    """Load the library (once).  Raises if it is absent or of another ABI."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise HipError(f'{LIB_PATH} is missing: build it with `python -m redrock_amd.build`')
    L = ctypes.CDLL(LIB_PATH)
    L.rr_abi_version.restype = _i
    L.rr_abi_version.argtypes = []
    L.rr_last_error.restype = ctypes.c_char_p
    L.rr_last_error.argtypes = []
    L.rr_rebin.restype = _i
    L.rr_rebin.argtypes = [_vp, _i, _vp, _i, _ll, _vp, _i, _vp, _vp, _i,
                           ctypes.POINTER(LymanLine), _i, _vp, _i, _ll, _ll, _ll, _d, _vp]
    L.rr_zscan.restype = _i
    L.rr_zscan.argtypes = [ctypes.POINTER(ZscanDesc), _vp]
    L.rr_zscan_noalloc.restype = _i
    L.rr_zscan_noalloc.argtypes = [ctypes.POINTER(ZscanDesc), _vp]
    L.rr_zscan_polish.restype = _i
    L.rr_zscan_polish.argtypes = [ctypes.POINTER(ZscanDesc), ctypes.POINTER(PolishDesc), _vp]
    L.rr_zscan_workspace_bytes.restype = _ll
    L.rr_zscan_workspace_bytes.argtypes = [ctypes.POINTER(ZscanDesc)]
    L.rr_pack_mixed.restype = _i
    L.rr_pack_mixed.argtypes = [_vp, _vp, _vp, _i, _i, _i, _vp, _vp, _vp, _vp]
    L.rr_pack_lanes_bytes.restype = _ll
    L.rr_pack_lanes_bytes.argtypes = [_i, _i, _i]
    L.rr_pack_lanes.restype = _i
    L.rr_pack_lanes.argtypes = [_vp, _vp, _vp, _vp, _i, _i, _i, _vp, _vp]
    L.rr_zscan_resid_tiles.restype = _i
    L.rr_zscan_resid_tiles.argtypes = [ctypes.POINTER(_i), _i]
    L.rr_resid_u_slots.restype = _i
    L.rr_resid_u_slots.argtypes = [_i, ctypes.POINTER(_i)]
    L.rr_zscan_tiles.restype = _i
    L.rr_zscan_tiles.argtypes = [ctypes.POINTER(_i), _i]
    L.rr_zscan_tile_table.restype = _i
    L.rr_zscan_tile_table.argtypes = [_i, ctypes.POINTER(_i), _i]
    L.rr_arch_prep.restype = _i
    L.rr_arch_prep.argtypes = [_vp, _vp, _vp, _i, _i, _vp, _i, _vp, _vp, _vp, _vp, _vp]
    L.rr_arch_workspace_bytes.restype = _ll
    L.rr_arch_workspace_bytes.argtypes = [ctypes.POINTER(ArchDesc)]
    L.rr_arch_fit.restype = _i
    L.rr_arch_fit.argtypes = [ctypes.POINTER(ArchDesc), _vp]
    L.rr_minima_head.restype = _i
    L.rr_minima_head.argtypes = [_vp, _ll, _ll, _ll, _i, _vp, _vp, _vp]
    L.rr_gather_rows.restype = _i
    L.rr_gather_rows.argtypes = [_vp, _ll, _ll, _vp, _i]
    v = L.rr_abi_version()
    if v != RR_ABI_VERSION:
        raise HipError(f'libredrock_hip ABI {v}, expected {RR_ABI_VERSION}')
    _lib = L
    return L


def resid_u_slots(nb):
    """(KS, nprod) of the matrix-core residual's template operand cache:
    fp16 elements per (z, row) and the hi/lo products they hold (3 or 4)."""
    n = _i(0)
    ks = lib().rr_resid_u_slots(int(nb), ctypes.byref(n))
    if ks < 0:
        check(ks, 'rr_resid_u_slots')
    return ks, n.value


def lane_pixels(npix_pad):
    """Pixels of the lane-major per-target pack (rr_lane_pixels in the header)."""
    return (npix_pad + 21) // 22 * 22


def check(rc, what):
    if rc != 0:
        msg = lib().rr_last_error().decode(errors='replace')
        raise HipError(f'{what} failed ({rc}): {msg}')


def stream_handle(stream=None):
    """Raw hipStream_t of a torch stream (default: the current stream)."""
    import torch
    s = torch.cuda.current_stream() if stream is None else stream
    return ctypes.c_void_p(s.cuda_stream)


def ptr(t):
    """Device pointer of a contiguous device tensor (None -> NULL).

    The kernels index flat memory, so strided views are rejected rather than
    silently misread.
    """
    if t is None:
        return None
    if not t.is_contiguous():
        raise ValueError('device arrays passed to libredrock_hip must be contiguous')
    if not t.is_cuda:
        raise ValueError('libredrock_hip needs device (cuda) tensors')
    return ctypes.c_void_p(t.data_ptr())
