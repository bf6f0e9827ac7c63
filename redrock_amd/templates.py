"""Templates and the device-resident rebinned template cache.

Template mirrors the in-memory constructor of py/redrock/templates.py:29-272
(FITS reading and template discovery are outside the scan path).
DistTemplate mirrors py/redrock/templates.py:525-663 for one process: it
rebins the template to every redshift on every wavelength grid once
(templates.py:584-597) — here on the GPU, with the IGM multiply fused — and
keeps the cache in HBM for the whole scan.
"""
from __future__ import annotations

import numpy as np

valid_template_methods = ('PCA', 'NMF')


def parse_fulltype(fulltype):
    """templates.py:274-284."""
    fulltype = fulltype.strip()
    if ':::' in fulltype:
        spectype, subtype = fulltype.split(':::')
    else:
        spectype, subtype = fulltype, ''
    return spectype, subtype


def make_fulltype(spectype, subtype):
    """templates.py:286-298."""
    spectype = spectype.strip()
    if subtype is not None:
        subtype = subtype.strip()
    if subtype is not None and subtype != '':
        return f'{spectype}:::{subtype}'
    return spectype


class Template(object):
    """A spectral template basis (templates.py:29-272, in-memory form).

    Args:
        spectype (str): GALAXY, QSO, STAR, ...
        redshifts (array): trial redshifts of the coarse scan.
        wave (array): rest-frame wavelengths, increasing.
        flux (array): (nbasis, nwave) basis vectors.
        subtype (str), method ('PCA' or 'NMF'), igm_model (str or None),
        version (str): as in the reference.
    """

    def __init__(self, filename=None, spectype=None, redshifts=None, wave=None,
                 flux=None, subtype=None, method='PCA', igm_model='Inoue14',
                 zscan_galaxy=None, zscan_qso=None, zscan_star=None, version=None):
        if filename is not None:
            raise NotImplementedError('template files are outside the chi2-scan path; '
                                      'construct Template from arrays')
        self._filename = None
        self._rrtype = spectype
        self._redshifts = np.asarray(redshifts, dtype=np.float64)
        self.wave = np.asarray(wave, dtype=np.float64)
        self.flux = np.asarray(flux, dtype=np.float64)
        self._subtype = subtype
        self._igm_model = igm_model
        self._method = method
        self._version = version
        self._nbasis = self.flux.shape[0]
        self._nwave = self.flux.shape[1]
        self.minwave = self.wave[0]
        self.maxwave = self.wave[-1]
        self.gpuwave = None
        self.gpuflux = None
        if self._method not in valid_template_methods:
            raise ValueError(f'Template method {self._method} unrecognized; '
                             f'should be one of {valid_template_methods}')
        # templates.py:172-174: [OII] slice for GALAXY PCA templates
        if self._rrtype == 'GALAXY' and self._method == 'PCA':
            isOII = (3724 <= self.wave) & (self.wave <= 3733)
            self.OIItemplate = self.flux[:, isOII]

    def __str__(self):
        zmin = round(self._redshifts[0], 4)
        zmax = round(self._redshifts[-1], 4)
        return (f'{self._rrtype} {self._subtype} {self._method} '
                f'IGM={self._igm_model} z={zmin}-{zmax}')

    __repr__ = __str__

    @property
    def nbasis(self):
        return self._nbasis

    @property
    def nwave(self):
        return self._nwave

    @property
    def template_type(self):
        return self._rrtype

    @property
    def sub_type(self):
        return self._subtype

    @property
    def full_type(self):
        return make_fulltype(self._rrtype, self._subtype)

    @property
    def version(self):
        return self._version

    @property
    def filename(self):
        return self._filename

    @property
    def redshifts(self):
        return self._redshifts

    @property
    def solve_matrices_algorithm(self):
        return self._method

    @property
    def method(self):
        return self._method

    @property
    def igm_model(self):
        return self._igm_model

    def device_arrays(self):
        """(wave, flux) as float64 device tensors, uploaded once."""
        import torch
        if self.gpuwave is None:
            dev = torch.device('cuda')
            self.gpuwave = torch.as_tensor(self.wave, dtype=torch.float64, device=dev)
            self.gpuflux = torch.as_tensor(np.ascontiguousarray(self.flux),
                                           dtype=torch.float64, device=dev)
        return self.gpuwave, self.gpuflux


class DistTemplatePiece(object):
    """The local redshift slice of a template cache (templates.py DistTemplatePiece)."""

    def __init__(self, index, redshifts, data):
        self.index = index
        self.redshifts = redshifts
        self.data = data


class DistTemplate(object):
    """Template rebinned to all redshifts, resident on the GPU.

    Mirrors templates.py:525-663 with comm=None: the whole redshift grid is
    local and cycle() is a no-op returning True.  ``local.data`` is the
    reference's dict {wavehash: (nz, nwave, nbasis)} view, materialised from
    the device cache on first access; the scan kernels read the device cache
    directly (``cache``).
    """

    def __init__(self, template, dwave, mp_procs=1, comm=None, use_gpu=True, gpu_mode=True):
        if comm is not None:
            raise NotImplementedError('MPI distribution is outside the chi2-scan path')
        from .zscan import TemplateCache
        self._comm = None
        self._template = template
        self._dwave = dwave
        self.cache = TemplateCache(template, dwave)
        self._piece = _LazyPiece(self)

    @property
    def comm(self):
        return self._comm

    @property
    def template(self):
        return self._template

    @property
    def local(self):
        return self._piece

    def cycle(self):
        return True


class _LazyPiece(object):
    def __init__(self, dt):
        self._dt = dt
        self.index = 0
        self.redshifts = dt.template.redshifts
        self._data = None

    @property
    def data(self):
        if self._data is None:
            self._data = self._dt.cache.reference_layout()
        return self._data
