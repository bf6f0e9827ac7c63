"""Targets and spectra: the attribute contract the chi2 scan consumes.

Mirrors py/redrock/targets.py (Spectrum :18-142, Target :145-263,
DistTargets :390-489, DistTargetsCopy :521-577) for the single-process case.
Coadds, shared-memory packing and MPI distribution are outside the scan path
and are not provided.
"""
from __future__ import annotations

import numpy as np
import scipy.sparse

from . import constants

RBAND_HALF = 5          # resolution band half width (11 diagonals)
RBAND_NDIAG = 2 * RBAND_HALF + 1


def resolution_band(R, n=None):
    """Return the (n, 11) row band of a resolution matrix: out[p, k] = R[p, p+k-5].

    R may be any scipy.sparse matrix (the reference stores a dia_matrix whose
    ``data`` is column-indexed, ``data[d, j] = R[j - offsets[d], j]``,
    targets.py:25-27).  Duplicate entries are summed, as scipy does.
    Raises NotImplementedError if R has entries farther than 5 from the
    diagonal.
    """
    if scipy.sparse.issparse(R) and R.format == 'dia':
        return _dia_band(R, n)
    coo = scipy.sparse.coo_matrix(R)
    if n is None:
        n = coo.shape[0]
    k = coo.col.astype(np.int64) - coo.row.astype(np.int64) + RBAND_HALF
    keep = coo.data != 0
    if np.any((k[keep] < 0) | (k[keep] >= RBAND_NDIAG)):
        raise NotImplementedError('resolution matrix wider than 11 diagonals')
    band = np.zeros((n, RBAND_NDIAG), dtype=np.float64)
    np.add.at(band, (coo.row[keep], k[keep]), coo.data[keep])
    return band


def _dia_band(R, n=None):
    """resolution_band of a dia_matrix by slicing its diagonals: row p of
    diagonal offset `off` holds R[p, p+off] = data[d, p+off]."""
    nrow, ncol = R.shape
    if n is None:
        n = nrow
    band = np.zeros((n, RBAND_NDIAG), dtype=np.float64)
    data = np.asarray(R.data)
    L = min(ncol, data.shape[1])
    for d, off in enumerate(np.asarray(R.offsets).astype(np.int64)):
        p0, p1 = max(0, -off), min(n, nrow, L - off)
        if p1 <= p0:
            continue
        vals = data[d, p0 + off:p1 + off]
        if abs(off) > RBAND_HALF:
            if np.any(vals != 0):
                raise NotImplementedError('resolution matrix wider than 11 diagonals')
            continue
        band[p0:p1, off + RBAND_HALF] += vals
    return band


class Spectrum(object):
    """One camera's spectrum (targets.py:18-142).

    Args:
        wave, flux, ivar (array): wavelength grid, flux and inverse variance.
        R (scipy.sparse matrix): resolution matrix (dia_matrix in redrock).
        Rcsr (scipy.sparse.csr_matrix): optional CSR copy of R.
        band (str): camera name; used as the wavehash when given.

    As in the reference, ivar is zeroed in place where the row sum of R is
    below constants.min_resolution_integral (targets.py:33-35).
    """

    def __init__(self, wave, flux, ivar, R, Rcsr=None, band=None):
        if R is not None:
            w = np.asarray(R.sum(axis=1))[:, 0] < constants.min_resolution_integral
            ivar[w] = 0.
        self.nwave = wave.size
        self.wave = wave
        self.flux = flux
        self.ivar = ivar
        self.R = R
        self._Rcsr = Rcsr
        self._band = None
        self.band = band
        if band is not None:
            self.wavehash = band
        else:
            if hasattr(R, 'data'):
                self.wavehash = hash((len(wave), wave[0], wave[1], wave[-2], wave[-1],
                                      R.data.shape[0]))
            else:
                self.wavehash = hash((len(wave), wave[0], wave[1], wave[-2], wave[-1]))
        self.minedge = None
        self.maxedge = None

    @property
    def Rcsr(self):
        if self._Rcsr is None:
            self._Rcsr = self.R.tocsr()
        return self._Rcsr

    def rband(self):
        """Cached (nwave, 11) row band of R (see resolution_band)."""
        if self._band is None:
            self._band = resolution_band(self.R if self.R is not None else self._Rcsr,
                                         self.nwave)
        return self._band


class Target(object):
    """A single target (targets.py:145-263)."""

    def __init__(self, targetid, spectra, coadd=False, cosmics_nsig=0., meta=None):
        if coadd:
            raise NotImplementedError('coadds are outside the chi2-scan path')
        self.id = targetid
        self.spectra = spectra
        self.meta = dict() if meta is None else meta
        self.bands = [s.band for s in spectra]
        self._device = None   # cached device-side arrays, see zscan.TargetBatch
        self._legendre = None
        self.nleg = 0

    def legendre(self, nleg, use_gpu=False):
        """{wavehash: (nleg, nwave)} Legendre polynomials of degree 0..nleg-1
        of each camera's reduced wavelength, cached (targets.py:237-254)."""
        from .utils import legendre_rows
        if self._legendre is None or self.nleg != nleg:
            dwave = {s.wavehash: s.wave for s in self.spectra}
            self._legendre = {hs: legendre_rows(w, nleg) for hs, w in dwave.items()}
            self.nleg = nleg
        if use_gpu:
            import torch
            return {hs: torch.as_tensor(v, device='cuda') for hs, v in self._legendre.items()}
        return self._legendre

    def eval_model(self, bestfit, templates, archetypes=None, nleg=None):
        """Best-fit model of every camera, {wavehash: flux}, also kept as
        self.model (targets.py:184-234).

        bestfit: one fit row with SPECTYPE, SUBTYPE, COEFF, Z, FITMETHOD
        (upper or lower case; LEGCOEFF for archetype fits).  templates:
        {(spectype, subtype): Template}; archetypes: {spectype: Archetype}.
        All-zero coefficients (fully masked data) give zero models.
        """
        def field(name):
            try:
                return bestfit[name]
            except (KeyError, ValueError):
                return bestfit[name.lower()]

        spectype, subtype = field('SPECTYPE'), field('SUBTYPE')
        coeff, z, fitmethod = np.asarray(field('COEFF')), field('Z'), field('FITMETHOD')
        self.model = dict()
        if np.all(coeff == 0.0):
            for sp in self.spectra:
                self.model[sp.wavehash] = np.zeros(len(sp.wave), dtype=np.float32)
        elif fitmethod == 'ARCH':
            narch = len(subtype.split(';'))
            legcoeff = np.asarray(field('LEGCOEFF'))
            if legcoeff.ndim == 1:
                legcoeff = [legcoeff for _ in self.spectra]
            ax = archetypes[spectype]
            for i, sp in enumerate(self.spectra):
                self.model[sp.wavehash] = ax.eval(subtype, coeff[0:narch], sp.wave, z, R=sp.Rcsr,
                                                  legcoeff=legcoeff[i])
        else:
            tx = templates[(spectype, subtype)]
            for sp in self.spectra:
                self.model[sp.wavehash] = tx.eval(coeff, sp.wave, z, sp.Rcsr)
        return self.model

    def gpu_spectral_data(self):
        """Concatenated (weights, flux, wflux) as device tensors (targets.py:256-263)."""
        import torch
        if self._device is None:
            w = np.concatenate([s.ivar for s in self.spectra])
            f = np.concatenate([s.flux for s in self.spectra])
            dev = torch.device('cuda')
            wt = torch.as_tensor(w, dtype=torch.float64, device=dev)
            ft = torch.as_tensor(f, dtype=torch.float64, device=dev)
            self._device = (wt, ft, torch.as_tensor(w * f, dtype=torch.float64, device=dev))
        return self._device


class DistTargets(object):
    """Base class of target containers (targets.py:390-489)."""

    def __init__(self, targetids, comm=None):
        if comm is not None:
            raise NotImplementedError('MPI distribution is outside the chi2-scan path')
        self._comm = comm
        self._targetids = targetids
        self._dwave = None

    @property
    def comm(self):
        return self._comm

    @property
    def all_target_ids(self):
        return self._targetids

    def local_target_ids(self):
        return self._local_target_ids()

    def local(self):
        return self._local_data()

    def eval_models(self, bestfit, templates, archetypes=None, nleg=None):
        """Target.eval_model of each local target, with its row of bestfit
        found by TARGETID (targets.py:442-461)."""
        cols = bestfit.dtype.names if hasattr(bestfit, 'dtype') else list(bestfit.keys())
        tid = bestfit['TARGETID'] if 'TARGETID' in cols else bestfit['targetid']
        models = list()
        for tgt in self.local():
            i = np.where(np.asarray(tid) == tgt.id)[0][0]
            if hasattr(bestfit, 'dtype'):
                row = bestfit[i]
            else:
                row = {k: v[i] for k, v in bestfit.items()}
            models.append(tgt.eval_model(row, templates, archetypes, nleg))
        return models

    def wavegrids(self):
        """Dictionary of wavelength grids keyed by wavehash (targets.py:468-489)."""
        if self._dwave is None:
            d = dict()
            for t in self.local():
                for s in t.spectra:
                    if s.wavehash not in d:
                        d[s.wavehash] = s.wave.copy()
            self._dwave = d
        return self._dwave


class DistTargetsCopy(DistTargets):
    """All targets held by this process (targets.py:521-577, comm=None)."""

    def __init__(self, targets, comm=None, root=0):
        self._my_data = list(targets)
        self._my_targets = [t.id for t in targets]
        super().__init__(list(self._my_targets), comm=comm)

    def _local_target_ids(self):
        return self._my_targets

    def _local_data(self):
        return self._my_data
