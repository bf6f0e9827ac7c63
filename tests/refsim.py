"""Fake templates and targets built as the reference's unit tests build them
(py/redrock/test/util.py:12-75): a 3-vector basis (constant, linear ramp, a
Gaussian line at mid-range plus a 0.1 sine), two cameras of 5 A pixels with
two noisy exposures each, and resolution matrices whose every column holds
the same sampled, normalised Gaussian, so the rows at the camera edges lose
part of it (their integral falls below 0.99 and Spectrum zeroes their ivar,
targets.py:33-35).  Data only: this module generates inputs for the tests.
"""
import numpy as np
import scipy.sparse

from redrock_amd.rebin import centers2edges
from redrock_amd.synth import antiderivative_rebin
from redrock_amd.targets import Spectrum, Target
from redrock_amd.templates import Template

CAMERAS = ((4000.0, 6700.0), (6502.0, 8001.0))


def template(spectype='GALAXY', subtype='', redshifts=None, wmin=100.0, wmax=9000.0, step=5.0):
    wave = np.arange(wmin, wmax + step / 2.0, step)
    ramp = np.arange(len(wave)) / len(wave)
    phase = np.pi * (2 * (wave - wmin) / (wmax - wmin) - 1)
    line = np.exp(-(wave - 0.5 * (wmin + wmax)) ** 2 / (2 * 20.0 ** 2)) + 0.1 * np.sin(phase)
    flux = np.vstack([np.ones_like(wave), ramp, line])
    z = np.linspace(0, 1, 20) if redshifts is None else np.asarray(redshifts, dtype=np.float64)
    return Template(spectype=spectype, redshifts=z, wave=wave, flux=flux, subtype=subtype)


def resolution(n, sigma=2.0):
    """dia_matrix whose every column is exp(-x^2 / (2 sigma)), x = -5..5, normalised
    (the test recipe's kernel: sigma enters unsquared)."""
    x = np.arange(-5, 6)
    k = np.exp(-x ** 2 / (2.0 * sigma))
    k /= k.sum()
    return scipy.sparse.dia_matrix((np.repeat(k[:, None], n, axis=1), x), shape=(n, n))


def target(tid, z, rng, coeff=(1.0, 2.0, 3.0), step=5.0):
    """Target at redshift z: per camera the template model (no R applied, as
    the recipe), sigma ~ N(1, 0.1) clipped to [0.5, 1.5], two exposures."""
    tm = template()
    model_flux = np.asarray(coeff) @ tm.flux
    spectra = []
    for lo, hi in CAMERAS:
        wave = np.arange(lo, hi, step)
        model = antiderivative_rebin(tm.wave * (1 + z), model_flux, centers2edges(wave))
        sigma = rng.normal(loc=1.0, scale=0.1, size=len(wave)).clip(0.5, 1.5)
        R = resolution(len(wave))
        ivar = 1.0 / sigma ** 2
        for _ in range(2):
            noisy = model + rng.normal(scale=sigma)
            s = Spectrum(wave, noisy, ivar.copy(), R, R.tocsr())
            s.raw_ivar = ivar            # before the resolution-integral mask
            spectra.append(s)
    return Target(tid, spectra)
