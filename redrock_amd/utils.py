"""Small helpers of redrock's archetype mode (py/redrock/utils.py)."""
from __future__ import annotations

import numpy as np


def native_endian(data):
    """`data` in native byte order: the same array when it already is, else a
    byte-swapped copy (utils.py:19-40)."""
    if data.dtype.isnative:
        return data
    return data.astype(data.dtype.newbyteorder('='))


def reduced_wavelength(wave):
    """wave mapped linearly onto [-1, 1], the domain of the Legendre
    polynomials of the archetype fit (utils.py:260-271)."""
    wave = np.asarray(wave, dtype=np.float64)
    wavemax = wave.max()
    wavemin = wave.min()
    return 2 * (wave - wavemin) / (wavemax - wavemin) - 1


def legendre_rows(wave, nleg):
    """(nleg, nwave): Legendre polynomials P_0 .. P_{nleg-1} of the reduced
    wavelength (targets.py:248)."""
    x = reduced_wavelength(wave)
    return np.array([np.polynomial.legendre.legval(x, np.eye(nleg)[i]) for i in range(nleg)]).reshape(
        nleg, len(x))


def extra_columns_in_archetype_mode():
    """Columns of the fit table that only archetype fits fill (utils.py:273-279):
    pca_* columns take the PCA values of other templates, archcoeff and
    legcoeff are zero for them."""
    pca_map = {"pca_coeff": "coeff", "pca_subtype": "subtype", "pca_spectype": "spectype"}
    zero_like_keys = {"archcoeff", "legcoeff"}
    return pca_map, zero_like_keys
