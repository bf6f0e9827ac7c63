"""Lyman-series IGM transmission (py/redrock/igm.py:321-439).

The scan path applies the transmission inside the rebin kernel (rr_rebin);
this module provides the reference's host-side function with the same
semantics and the line tables the kernel consumes.  Models are the
CaluraKamble family (Calura12, Kamble20) plus tables registered with
register_lyman_model; Inoue14 is not on the scan path's workloads and raises
NotImplementedError whenever it would actually attenuate.
"""
from __future__ import annotations

import numpy as np

from . import constants

_custom = dict()

igm_models = ('Calura12', 'Kamble20', 'Inoue14', 'None', None)


def register_lyman_model(name, lines):
    """Register a Lyman-series table: lines = [(line, A, B), ...] in multiply order."""
    _custom[name] = [(float(l), float(a), float(b)) for (l, a, b) in lines]


def known_models():
    return igm_models + tuple(_custom.keys())


def lyman_table(model):
    """Ordered (line, A, B) list of a model; None for 'None'/None."""
    if model is None or model == 'None':
        return None
    if model in _custom:
        return _custom[model]
    if model in ('Calura12', 'Kamble20'):
        return [(v['line'], v['A'], v['B']) for v in constants.Lyman_series[model].values()]
    if model == 'Inoue14':
        raise NotImplementedError('Inoue14 IGM is not implemented on the scan path')
    raise ValueError(f'Unrecognized model {model}; should be one of {known_models()}')


def igm_applies(zObj, lObs, model):
    """True when transmission_Lyman(..., always_return_array=False) is not None.

    Same test as igm.py:394-401: model not None and min(lObs)/(1+max(z)) <= LyA.
    """
    if model is None or model == 'None':
        return False
    min_wave = np.min(lObs) / (1. + np.max(zObj))
    return not (min_wave > constants.LyA_wavelength)


def transmission_Lyman_CaluraKamble(zObj, lObs, use_gpu=False, model='Calura12'):
    """igm.py:321-364."""
    lines = lyman_table(model)
    lRest = np.outer(1.0 / (1.0 + zObj), lObs)
    min_wave = lRest.min()
    lObs = np.tile(lObs, (zObj.size, 1))
    T = np.ones(lRest.shape)
    for (line, A, B) in lines:
        if min_wave > line:
            continue
        w = lRest < line
        zpix = lObs[w] / line - 1.
        tauEff = A * (1. + zpix) ** B
        T[w] *= np.exp(-tauEff)
    return T


def transmission_Lyman(zObj, lObs, use_gpu=False, model='Calura12', always_return_array=True):
    """Transmitted flux fraction (igm.py:369-439); same shapes and None rules."""
    if model not in known_models():
        raise ValueError(f'Unrecognized model {model}; should be one of {known_models()}')
    if (model == 'None' or model is None) and not always_return_array:
        return None
    if hasattr(zObj, 'cpu'):
        zObj = zObj.cpu().numpy()
    if hasattr(lObs, 'cpu'):
        lObs = lObs.cpu().numpy()
    min_wave = np.min(lObs) / (1. + np.max(zObj))
    if min_wave > constants.LyA_wavelength and not always_return_array:
        return None
    scalar_z = np.isscalar(zObj)
    scalar_l = np.isscalar(lObs)
    zObj = np.atleast_1d(zObj)
    lObs = np.atleast_1d(lObs)
    if always_return_array and (min_wave > constants.LyA_wavelength or model is None
                                or model == 'None'):
        T = np.ones((len(zObj), len(lObs)))
    else:
        T = transmission_Lyman_CaluraKamble(zObj, lObs, model=model)
    if scalar_z and scalar_l:
        T = T[0, 0]
    elif scalar_z:
        T = T[0]
    elif scalar_l:
        T = T[:, 0]
    if use_gpu:
        import torch
        T = torch.as_tensor(T, device='cuda')
    return T
