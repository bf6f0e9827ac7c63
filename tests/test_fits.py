"""FITS reader/writer (redrock_amd.fits) and redrock-format template files.

tests/golden/redrock-test.fits and rrarchetype-galaxy-test.fits are the data
files of the reference's own tests (py/redrock/test/data): a redrock output
table (REDSHIFTS) and an archetype table, both written by astropy.
"""
import os

import numpy as np
import pytest

from redrock_amd import fits
from redrock_amd.templates import Template, write_template, load_templates

GOLDEN = os.path.join(os.path.dirname(__file__), 'golden')


def test_image_and_table_roundtrip(tmp_path):
    img = np.arange(24, dtype=np.float64).reshape(4, 6) / 7.0
    ints = np.array([3, -5, 2 ** 40])
    cols = {'TARGETID': np.array([1, 2, 3]), 'Z': np.array([0.5, 1.25, -0.001]),
            'COEFF': np.arange(30, dtype=float).reshape(3, 10), 'SPECTYPE': np.array(['GALAXY', 'QSO', 'STAR']),
            'OK': np.array([True, False, True])}
    path = str(tmp_path / 'x.fits')
    fits.write(path, [fits.image_hdu(img, 'A', {'NOTE': "it's", 'FLAG': True, 'N': 7, 'X': 0.25}),
                      fits.image_hdu(ints, 'B'), fits.table_hdu(cols, 'T')],
               primary_header={'RRVER': 'test'})
    h = fits.read(path)
    assert [x.name for x in h] == ['PRIMARY', 'A', 'B', 'T']
    assert h[0].header['RRVER'] == 'test' and h[0].data is None
    assert np.array_equal(h['A'].data, img) and h['A'].data.dtype == np.float64
    assert h['A'].header['NOTE'] == "it's" and h['A'].header['FLAG'] is True
    assert h['A'].header['N'] == 7 and h['A'].header['X'] == 0.25
    assert np.array_equal(h['B'].data, ints)
    t = h['T'].data
    assert list(t.keys()) == list(cols.keys())
    for k, v in cols.items():
        assert np.array_equal(t[k], v), k
    assert os.path.getsize(path) % 2880 == 0


def test_reference_redshift_table():
    """The reference's rrdesi output: 3 targets, GALAXY (10 coeff) and QSO (4)."""
    h = fits.read(os.path.join(GOLDEN, 'redrock-test.fits'))
    assert h[0].header['TEMNAM00'].strip() == 'GALAXY'
    assert h[0].header['TEMFIL00'] == 'rrtemplate-GALAXY-None-v2.6.fits'
    t = h['REDSHIFTS'].data
    assert list(t.keys()) == ['TARGETID', 'Z', 'ZERR', 'ZWARN', 'CHI2', 'COEFF', 'NPIXELS',
                              'SPECTYPE', 'SUBTYPE', 'NCOEFF', 'DELTACHI2']
    n = len(t['TARGETID'])
    assert n == 3 and t['COEFF'].shape == (3, 10)
    assert set(t['SPECTYPE']) <= {'GALAXY', 'QSO', 'STAR'}
    assert np.all(np.isfinite(t['Z'])) and np.all(t['CHI2'] > 0) and np.all(t['ZERR'] > 0)
    assert np.all(t['DELTACHI2'] >= 0) and np.all(t['NPIXELS'] > 0)
    # a template with fewer coefficients leaves the rest of COEFF zero
    for nc, c in zip(t['NCOEFF'], t['COEFF']):
        assert nc in (4, 10) and np.all(c[nc:] == 0) and np.any(c[:nc] != 0)
    assert np.all(t['TARGETID'] > 0)
    fm = h['FIBERMAP'].data
    assert len(fm['TARGETID']) == 3 and set(fm['TARGETID']) == set(t['TARGETID'])


def test_reference_archetype_table():
    h = fits.read(os.path.join(GOLDEN, 'rrarchetype-galaxy-test.fits'))
    hd = h['ARCHETYPES'].header
    assert hd['RRTYPE'].strip() == 'GALAXY' and hd['LOGLAM'] is False
    assert hd['CDELT1'] == 0.2 and abs(hd['CRVAL1'] - 1228.0701754386) < 1e-9
    a = h['ARCHETYPES'].data
    assert a['ARCHETYPE'].shape == (2, 48860) and a['ARCHETYPE'].dtype == np.float32
    flux = a['ARCHETYPE'].astype(np.float64)
    assert np.all(np.isfinite(flux)) and np.all(np.abs(flux).max(axis=1) > 0)
    assert all(len(s) > 0 for s in a['SUBTYPE'])
    # the wavelength grid the file defines spans the rest-frame optical
    wave = hd['CRVAL1'] + hd['CDELT1'] * np.arange(flux.shape[1])
    assert 1200 < wave[0] < 1300 and 10000 < wave[-1] < 11000


def _toy(spectype='GALAXY', subtype='', igm='Calura12', version='9.9', loglam=False):
    if loglam:
        wave = 10 ** (np.log10(1500.0) + 1e-4 * np.arange(300))
    else:
        wave = 1500.0 + 0.5 * np.arange(300)
    flux = np.vstack([np.sin(wave / (50.0 + 10 * k)) + k for k in range(3)])
    return Template(spectype=spectype, subtype=subtype, redshifts=np.linspace(0.1, 0.5, 11),
                    wave=wave, flux=flux, igm_model=igm, version=version)


@pytest.mark.parametrize('loglam', [False, True])
def test_template_file_roundtrip(tmp_path, loglam):
    t = _toy('STAR', 'K', 'None', '2.1', loglam)
    path = str(tmp_path / 'rrtemplate-star.fits')
    write_template(path, t, loglam=loglam)
    r = Template(path)
    assert r.template_type == 'STAR' and r.sub_type == 'K' and r.full_type == 'STAR:::K'
    assert r.igm_model == 'None' and r.version == '2.1' and r.solve_matrices_algorithm == 'PCA'
    assert np.allclose(r.wave, t.wave, rtol=1e-12, atol=0)
    assert np.array_equal(r.flux, t.flux)
    assert np.array_equal(r.redshifts, t.redshifts)
    assert r.nbasis == 3 and r.minwave == r.wave[0]


def test_template_file_defaults(tmp_path, monkeypatch):
    """Old-style files: no REDSHIFTS -> the type's default grid; no RRIGM ->
    the model implied by type and version (templates.py:89-143)."""
    from redrock_amd import templates as tmod
    for spectype, version, want_igm in (('GALAXY', '2.6', 'Calura12'), ('QSO', '1.1', 'Kamble20'),
                                        ('QSO', '1.0', 'Calura12'), ('STAR', 'x', 'None')):
        t = _toy(spectype, '', None, version)
        path = str(tmp_path / f'{spectype}-{version}.fits')
        write_template(path, t, with_redshifts=False)
        r = Template(path)
        assert r.igm_model == want_igm
        z = r.redshifts
        if spectype == 'GALAXY':
            want = 10 ** np.arange(np.log10(1 - 0.005), np.log10(1 + 1.7), 3e-4) - 1
        elif spectype == 'QSO':
            want = 10 ** np.arange(np.log10(1 + 0.05), np.log10(1 + 6.0), 5e-4) - 1
        else:
            want = np.arange(-0.002, 0.00201, 4e-5)
        assert np.array_equal(z, want)
    r = Template(str(tmp_path / 'GALAXY-2.6.fits'), zscan_galaxy='0.1,0.2,1e-3')
    assert np.array_equal(r.redshifts, 10 ** np.arange(np.log10(1.1), np.log10(1.2), 1e-3) - 1)
    r = Template(str(tmp_path / 'STAR-x.fits'), zscan_star='-0.001,0.001,1e-4')
    assert np.array_equal(r.redshifts, np.arange(-0.001, 0.001, 1e-4))
    # unknown version without RRIGM is an error, as in the reference
    t = _toy('GALAXY', '', None, '3.0')
    write_template(str(tmp_path / 'bad.fits'), t)
    with pytest.raises(ValueError):
        Template(str(tmp_path / 'bad.fits'))
    # relative names resolve through $RR_TEMPLATE_DIR
    monkeypatch.setenv('RR_TEMPLATE_DIR', str(tmp_path))
    assert load_templates(['STAR-x.fits'])[0].template_type == 'STAR'
    with pytest.raises(IOError):
        Template('no-such-template.fits')
    assert tmod.get_template_dir() == str(tmp_path)
