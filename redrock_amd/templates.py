"""Templates and the device-resident rebinned template cache.

Template mirrors py/redrock/templates.py:29-272: the in-memory constructor
and redrock-format template files (read with redrock_amd.fits).
DistTemplate mirrors py/redrock/templates.py:525-663 for one process: it
rebins the template to every redshift on every wavelength grid once
(templates.py:584-597) — here on the GPU, with the IGM multiply fused — and
keeps the cache in HBM for the whole scan.
"""
from __future__ import annotations

import glob
import os

import numpy as np

valid_template_methods = ('PCA', 'NMF')


def get_template_dir(template_path=None):
    """template_path if given, else $RR_TEMPLATE_DIR, else the package's
    data/templates directory (templates.py:300-308).  The path is returned
    whether or not it exists; find_templates checks it."""
    if template_path is not None:
        return template_path
    d = os.environ.get('RR_TEMPLATE_DIR')
    if d is None:
        d = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'data', 'templates')
    return d


def find_templates(template_path=None):
    """List of template files to use (templates.py:311-380).

    template_path is None (get_template_dir()), a list/tuple/set/array of
    files (returned as given), a directory (its templates-default.txt if
    present, else every rrtemplate-*.fits in it, sorted), a .txt file naming
    one template per line (blank lines and '#' comments skipped, trailing
    text ignored, relative names taken from the .txt file's directory), or a
    single .fits / .fits.gz / .fits.fz file.  ValueError for a path that does
    not exist, an unrecognised suffix, or a listed file that is missing.
    """
    template_path = get_template_dir(template_path)
    if isinstance(template_path, (list, tuple, set, np.ndarray)):
        return template_path
    template_path = os.fspath(template_path)
    if not os.path.exists(template_path):
        raise ValueError(f"can't find template_path {template_path!r}, $RR_TEMPLATE_DIR "
                         'or the package template directory')
    if os.path.isdir(template_path):
        listfile = os.path.join(template_path, 'templates-default.txt')
        template_dir = template_path
    elif template_path.endswith('.txt'):
        listfile = template_path
        template_dir = os.path.dirname(os.path.abspath(template_path))
    elif template_path.endswith(('.fits', '.fits.gz', '.fits.fz')):
        return [template_path]
    else:
        raise ValueError(f'Unrecognized template_path={template_path!r}')

    if not os.path.exists(listfile):
        return sorted(glob.glob(os.path.join(template_dir, 'rrtemplate-*.fits')))
    files = list()
    with open(listfile) as fx:
        for line in fx:
            line = line.strip()
            if len(line) < 2 or line.startswith('#'):
                continue
            name = line.split()[0]
            if not name.startswith('/'):
                name = f'{template_dir}/{name}'
            if not os.path.exists(name):
                raise ValueError(f'missing {name} given in {listfile}')
            files.append(name)
    return files


def header2templatefiles(hdr, template_dir=None):
    """Template file names from header keywords (templates.py:383-423).

    For nn = 00..99: TEMFILnn names the file; otherwise TEMNAMnn (a fulltype)
    and TEMVERnn give rrtemplate-{spectype}-{subtype}-v{version}.fits, with
    'None' for a blank subtype and version 0.1 for an 'unknown' STAR version
    (ValueError for another spectype).  Each name is joined to template_dir
    (default get_template_dir()).
    """
    names = list()
    for i in range(100):
        filekey, typekey, verkey = f'TEMFIL{i:02d}', f'TEMNAM{i:02d}', f'TEMVER{i:02d}'
        if filekey in hdr:
            names.append(hdr[filekey])
        elif typekey in hdr and verkey in hdr:
            version = hdr[verkey].strip()
            spectype, subtype = parse_fulltype(hdr[typekey])
            if version == 'unknown':
                if spectype != 'STAR':
                    raise ValueError(f'unknown version for spectype={spectype!r}')
                version = '0.1'
            subtype = subtype if subtype != '' else None
            names.append(f'rrtemplate-{spectype}-{subtype}-v{version}.fits')
    if template_dir is None:
        template_dir = get_template_dir()
    return [os.path.join(template_dir, n) for n in names]


def load_templates_from_header(hdr, template_dir=None, asdict=False):
    """Templates named by header keywords (templates.py:426-438)."""
    return load_templates(header2templatefiles(hdr, template_dir=template_dir), asdict=asdict)


def _read_template_file(filename):
    """Fields of a redrock-format template file (templates.py:47-155).

    Returns dict(version, wave, flux, method, redshifts or None, rrtype,
    subtype, igm_model or None).
    """
    from . import fits
    if not os.path.exists(filename):
        alt = os.path.join(get_template_dir(), filename)
        if not os.path.exists(alt):
            raise IOError('unable to find ' + filename)
        filename = alt
    fx = fits.read(filename)
    hdu = fx['BASIS_VECTORS']
    hdr = hdu.header
    out = dict()
    out['version'] = hdr['VERSION'] if 'VERSION' in hdr else 'unknown'
    wave = np.asarray(hdr['CRVAL1'] + hdr['CDELT1'] * np.arange(hdr['NAXIS1']), dtype=np.float64)
    if 'LOGLAM' in hdr and hdr['LOGLAM'] != 0:
        wave = 10 ** wave
    out['wave'] = wave
    out['flux'] = np.asarray(hdu.data, dtype=np.float64)
    out['method'] = hdr['RRMETHOD'].upper() if 'RRMETHOD' in hdr else 'PCA'
    red = None
    if 'REDSHIFTS' in fx:
        red = fx['REDSHIFTS'].data
        if isinstance(red, dict):                     # a one-column table
            red = next(iter(red.values()))
        red = np.asarray(red, dtype=np.float64).ravel()
    out['redshifts'] = red
    out['rrtype'] = hdr['RRTYPE'].strip().upper()
    out['subtype'] = hdr['RRSUBTYP'].strip().upper() if 'RRSUBTYP' in hdr else ''
    out['igm_model'] = hdr['RRIGM'] if 'RRIGM' in hdr else None
    return out


def _default_redshifts(rrtype, zscan_galaxy=None, zscan_qso=None, zscan_star=None):
    """Redshift grid of an old-style template file without REDSHIFTS (templates.py:89-118)."""
    if rrtype == 'GALAXY':
        if zscan_galaxy is not None:
            zmin, zmax, dz = zscan_galaxy.split(',')
            return 10 ** np.arange(np.log10(1 + float(zmin)), np.log10(1 + float(zmax)), float(dz)) - 1
        return 10 ** np.arange(np.log10(1 - 0.005), np.log10(1 + 1.7), 3e-4) - 1
    if rrtype == 'STAR':
        if zscan_star is not None:
            zmin, zmax, dz = zscan_star.split(',')
            return np.arange(float(zmin), float(zmax), float(dz))
        return np.arange(-0.002, 0.00201, 4e-5)
    if rrtype == 'QSO':
        if zscan_qso is not None:
            zmin, zmax, dz = zscan_qso.split(',')
            return 10 ** np.arange(np.log10(1 + float(zmin)), np.log10(1 + float(zmax)), float(dz)) - 1
        return 10 ** np.arange(np.log10(1 + 0.05), np.log10(1 + 6.0), 5e-4) - 1
    raise ValueError('Unknown redshift range to use for template type {}'.format(rrtype))


def _default_igm(rrtype, version):
    """IGM model of a template file without RRIGM (templates.py:127-143)."""
    if rrtype == 'STAR':
        return 'None'
    if rrtype == 'GALAXY' and version == '2.6':
        return 'Calura12'
    if rrtype == 'QSO' and version in ('0.1', '1.0'):
        return 'Calura12'
    if rrtype == 'QSO' and version == '1.1':
        return 'Kamble20'
    raise ValueError('Missing keyword RRIGM specifying IGM model to use')


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
    """A spectral template basis (templates.py:29-272).

    Args:
        filename (str): a redrock-format template file (absolute, or relative
            to $RR_TEMPLATE_DIR); then the arguments below except the zscan_*
            ranges (for files without a REDSHIFTS HDU) are read from it.
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
        self._filename = filename
        if filename is not None:
            f = _read_template_file(filename)
            spectype, subtype, version = f['rrtype'], f['subtype'], f['version']
            wave, flux, method = f['wave'], f['flux'], f['method']
            redshifts = f['redshifts']
            if redshifts is None:
                redshifts = _default_redshifts(spectype, zscan_galaxy, zscan_qso, zscan_star)
            igm_model = f['igm_model'] if f['igm_model'] is not None else _default_igm(spectype, version)
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

    def eval(self, coeff, wave, z, R=None):
        """Template flux for coefficients coeff at redshift z on wavelengths
        wave, convolved with R if given (templates.py:247-272).  No (1+z)
        factor is applied and there is no IGM: model = rebin(flux'coeff).
        """
        from .rebin import trapz_rebin
        coeff = np.asarray(coeff, dtype=np.float64)[0:self._nbasis]
        flux = self.flux.T.dot(coeff)
        model = trapz_rebin(self.wave * (1 + z), flux, wave)
        if R is not None:
            model = R.dot(model)
        return model

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


def write_template(path, template, with_redshifts=True, loglam=False):
    """Write `template` as a redrock-format file (BASIS_VECTORS image with
    CRVAL1/CDELT1/RRTYPE/RRSUBTYP/RRIGM/RRMETHOD/VERSION, REDSHIFTS image).

    The wavelength grid must be uniform in lambda (or in log10 lambda with
    loglam=True), as the format stores only its start and step.
    """
    from . import fits
    w = np.asarray(template.wave, dtype=np.float64)
    x = np.log10(w) if loglam else w
    step = (x[-1] - x[0]) / (len(x) - 1)
    if not np.allclose(np.diff(x), step, rtol=1e-9, atol=0):
        raise ValueError('template wavelengths are not uniformly spaced')
    hdr = {'CRVAL1': float(x[0]), 'CDELT1': float(step), 'LOGLAM': 1 if loglam else 0,
           'RRTYPE': template.template_type, 'RRSUBTYP': template.sub_type or '',
           'RRMETHOD': template.solve_matrices_algorithm,
           'VERSION': str(template.version if template.version is not None else 'unknown')}
    if template.igm_model is not None:
        hdr['RRIGM'] = str(template.igm_model)
    hdus = [fits.image_hdu(template.flux, 'BASIS_VECTORS', hdr)]
    if with_redshifts:
        hdus.append(fits.image_hdu(template.redshifts, 'REDSHIFTS'))
    fits.write(path, hdus)


def load_templates(template_path=None, zscan_galaxy=None, zscan_star=None, zscan_qso=None,
                   asdict=False):
    """Template objects (templates.py:440-474).

    template_path: a list of template files, or anything find_templates
    takes (None, a directory, a .txt list, one .fits file).  zscan_*:
    'zmin,zmax,dz' redshift ranges for files without a REDSHIFTS HDU.
    Returns a list, or with asdict a dict keyed by (spectype, subtype).
    """
    if isinstance(template_path, (str, os.PathLike, type(None))):
        template_path = find_templates(template_path)
    templates = [Template(f, zscan_galaxy=zscan_galaxy, zscan_star=zscan_star, zscan_qso=zscan_qso)
                 for f in template_path]
    if asdict:
        return {(t.template_type, t.sub_type): t for t in templates}
    return templates


def load_dist_templates(dwave, templates=None, comm=None, mp_procs=1,
                        zscan_galaxy=None, zscan_qso=None, zscan_star=None,
                        redistribute=False, use_gpu=False, gpu_mode=False):
    """One DistTemplate per template, each rebinned to every grid of dwave
    (templates.py:715-778).

    templates is what load_templates takes.  With one process the whole
    redshift grid is local, so redistribute (the reference's allgather of
    the slices) changes nothing; the rebinned caches always live in HBM.
    """
    if comm is not None:
        raise NotImplementedError('MPI distribution is outside the chi2-scan path')
    tdict = load_templates(templates, asdict=True, zscan_galaxy=zscan_galaxy,
                           zscan_star=zscan_star, zscan_qso=zscan_qso)
    return [DistTemplate(t, dwave, mp_procs=mp_procs, comm=comm, use_gpu=use_gpu, gpu_mode=gpu_mode)
            for t in tdict.values()]


def eval_model(data, wave, R=None, templates=None):
    """Model spectra of a table of fits (templates.py:781-832).

    data: table-like with Z, COEFF, SPECTYPE, SUBTYPE columns (upper or
    lower case).  wave: wavelengths, or a dict of them per camera, in which
    case R must be None or a dict with the same keys and the result is a dict.
    R: per-row resolution matrices (anything with .dot), or None for no
    convolution.  templates: {(spectype, subtype): Template}, default
    load_templates(asdict=True).  Returns float32 (nrow, nwave).
    """
    if templates is None:
        templates = load_templates(asdict=True)
    if isinstance(wave, dict):
        if R is None:
            R = {k: None for k in wave}
        elif not isinstance(R, dict) or set(R.keys()) != set(wave.keys()):
            raise ValueError('R must be dict with same keys as wave')
        return {k: eval_model(data, wave[k], R=R[k], templates=templates) for k in wave}

    def col(name):
        return data[name] if name in _columns(data) else data[name.lower()]

    z, coeff, spectype, subtype = col('Z'), col('COEFF'), col('SPECTYPE'), col('SUBTYPE')
    models = np.zeros((len(z), len(wave)), dtype='f4')
    for i in range(len(z)):
        tx = templates[(spectype[i], subtype[i])]
        models[i] = tx.eval(np.asarray(coeff[i])[0:tx.nbasis], wave, z[i],
                            R=None if R is None else R[i])
    return models


def _columns(data):
    names = getattr(getattr(data, 'dtype', None), 'names', None)
    if names is not None:
        return names
    return data.keys() if hasattr(data, 'keys') else data.colnames
