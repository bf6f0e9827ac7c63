"""Template discovery and model evaluation (the reference's
test_templates.py:15-217 and the eval helpers of templates.py / targets.py)."""
import os

import numpy as np
import pytest
import scipy.sparse

from redrock_amd import templates as tmod
from redrock_amd.templates import (Template, eval_model, find_templates, get_template_dir,
                                   header2templatefiles, load_templates,
                                   load_templates_from_header, make_fulltype, parse_fulltype,
                                   write_template)

DEFAULT = ('rrtemplate-galaxy-v2.fits', 'rrtemplate-qso-v1.fits')
ALTERNATE = ('rrtemplate-galaxy-v1.fits', 'rrtemplate-blat-v1.fits')


def _toy(spectype, subtype='', nbasis=3, version='9.9'):
    wave = 1000.0 + 0.5 * np.arange(18001)
    flux = np.vstack([np.sin(wave / (50.0 + 10 * k)) + 1.5 + k for k in range(nbasis)])
    return Template(spectype=spectype, subtype=subtype, redshifts=np.linspace(0.1, 0.5, 11),
                    wave=wave, flux=flux, igm_model='Calura12', version=version)


@pytest.fixture
def template_dir(tmp_path, monkeypatch):
    """Empty template files plus a default and an alternate list, as the
    reference's setUpClass makes them."""
    monkeypatch.delenv('RR_TEMPLATE_DIR', raising=False)
    for name in DEFAULT + ALTERNATE:
        (tmp_path / name).write_bytes(b'')
    for listname, names in (('templates-default.txt', DEFAULT),
                            ('templates-alternate.txt', ALTERNATE)):
        with open(tmp_path / listname, 'w') as fx:
            fx.write('# Header comment\n\n')
            for n in names:
                fx.write(f'{n}\n')
    return str(tmp_path)


def test_get_template_dir(monkeypatch):
    assert get_template_dir('blat') == 'blat'
    assert get_template_dir('/blat/foo') == '/blat/foo'
    monkeypatch.setenv('RR_TEMPLATE_DIR', '/biz/bat')
    assert get_template_dir() == '/biz/bat'
    monkeypatch.delenv('RR_TEMPLATE_DIR')
    pkg = os.path.dirname(os.path.abspath(tmod.__file__))
    assert get_template_dir() == os.path.join(pkg, 'data', 'templates')


def test_find_templates(template_dir, monkeypatch):
    got = find_templates(template_dir)
    assert sorted(os.path.basename(f) for f in got) == sorted(DEFAULT)
    assert find_templates(got[0:1]) == got[0:1]              # a list is returned as given
    monkeypatch.setenv('RR_TEMPLATE_DIR', template_dir)
    assert find_templates() == got
    alt = find_templates(os.path.join(template_dir, 'templates-alternate.txt'))
    assert [os.path.basename(f) for f in alt] == list(ALTERNATE)
    monkeypatch.chdir(template_dir)                           # a relative .txt path
    alt = find_templates('templates-alternate.txt')
    assert [os.path.basename(f) for f in alt] == list(ALTERNATE)
    assert all(os.path.exists(f) for f in alt)
    for bad in (os.path.join(template_dir, 'blat.txt'), 'blat.txt', 'blat.foo'):
        with pytest.raises(ValueError):
            find_templates(bad)
    assert find_templates(alt[0]) == [alt[0]]                # a single .fits file


def test_find_templates_old_style_and_listing_errors(tmp_path, monkeypatch):
    monkeypatch.delenv('RR_TEMPLATE_DIR', raising=False)
    for name in ('rrtemplate-b.fits', 'rrtemplate-a.fits', 'other.fits'):
        (tmp_path / name).write_bytes(b'')
    # no templates-default.txt: every rrtemplate-*.fits, sorted
    assert [os.path.basename(f) for f in find_templates(str(tmp_path))] == \
        ['rrtemplate-a.fits', 'rrtemplate-b.fits']
    # absolute paths and trailing comments in a list file
    with open(tmp_path / 'list.txt', 'w') as fx:
        fx.write(f'{tmp_path}/other.fits  # absolute\nrrtemplate-a.fits extra words\n')
    assert find_templates(str(tmp_path / 'list.txt')) == \
        [f'{tmp_path}/other.fits', f'{tmp_path}/rrtemplate-a.fits']
    with open(tmp_path / 'missing.txt', 'w') as fx:
        fx.write('rrtemplate-zzz.fits\n')
    with pytest.raises(ValueError):
        find_templates(str(tmp_path / 'missing.txt'))
    (tmp_path / 'odd.dat').write_bytes(b'')
    with pytest.raises(ValueError):
        find_templates(str(tmp_path / 'odd.dat'))


def test_header2templatefiles():
    tdir = '/some/dir'
    hdr = dict(TEMNAM00='GALAXY  ', TEMVER00='2.6     ',
               TEMNAM01='QSO:::HIZ', TEMVER01='1.0     ',
               TEMNAM02='STAR:::M', TEMVER02='unknown ')
    assert header2templatefiles(hdr, template_dir=tdir) == [
        '/some/dir/rrtemplate-GALAXY-None-v2.6.fits',
        '/some/dir/rrtemplate-QSO-HIZ-v1.0.fits',
        '/some/dir/rrtemplate-STAR-M-v0.1.fits']
    hdr = dict(TEMNAM00='GALAXY', TEMVER00='2.6', TEMNAM01='QSO:::HIZ', TEMVER01='1.0',
               TEMNAM02='STAR:::M', TEMVER02='3.14')
    got = header2templatefiles(hdr, template_dir=tdir)
    assert got[2].endswith('rrtemplate-STAR-M-v3.14.fits')
    hdr.update(TEMFIL00='blat.fits', TEMFIL01='foo.fits', TEMFIL02='bar.fits')
    assert [os.path.basename(f) for f in header2templatefiles(hdr, template_dir=tdir)] == \
        ['blat.fits', 'foo.fits', 'bar.fits']
    hdr = dict(TEMNAM00='GALAXY:::BGS', TEMVER00='3.0', TEMNAM02='GALAXY:::LRG', TEMVER02='3.0',
               TEMNAM01='GALAXY:::ELG', TEMVER01='3.0')
    assert [os.path.basename(f) for f in header2templatefiles(hdr, template_dir=tdir)] == \
        ['rrtemplate-GALAXY-BGS-v3.0.fits', 'rrtemplate-GALAXY-ELG-v3.0.fits',
         'rrtemplate-GALAXY-LRG-v3.0.fits']
    with pytest.raises(ValueError):
        header2templatefiles(dict(TEMNAM00='QSO', TEMVER00='unknown'), template_dir=tdir)
    # TEMNAM without TEMVER names nothing
    assert header2templatefiles(dict(TEMNAM00='QSO'), template_dir=tdir) == []


def test_fulltype_parsing():
    assert parse_fulltype('GALAXY') == ('GALAXY', '')
    assert parse_fulltype('QSO:::HIZ') == ('QSO', 'HIZ')
    assert make_fulltype('GALAXY', '') == 'GALAXY'
    assert make_fulltype('GALAXY', None) == 'GALAXY'
    assert make_fulltype('QSO', 'HIZ') == 'QSO:::HIZ'


def test_load_templates(tmp_path, monkeypatch):
    monkeypatch.delenv('RR_TEMPLATE_DIR', raising=False)
    files = []
    for spectype, subtype, version in (('GALAXY', '', '2.6'), ('QSO', 'HIZ', '1.0'),
                                       ('STAR', 'M', '3.14')):
        name = f'rrtemplate-{spectype}-{subtype or None}-v{version}.fits'
        write_template(str(tmp_path / name), _toy(spectype, subtype, version=version))
        files.append(str(tmp_path / name))
    got = load_templates(str(tmp_path))                      # directory, old style
    assert [t.filename for t in got] == sorted(files)
    got = load_templates(files[0:2])
    assert [t.filename for t in got] == files[0:2]
    d = load_templates(files, asdict=True)
    assert sorted(d.keys()) == [('GALAXY', ''), ('QSO', 'HIZ'), ('STAR', 'M')]
    assert d[('QSO', 'HIZ')].filename == files[1]
    # the same files named by header keywords
    hdr = dict(TEMNAM00='GALAXY', TEMVER00='2.6', TEMNAM01='QSO:::HIZ', TEMVER01='1.0',
               TEMNAM02='STAR:::M', TEMVER02='3.14')
    d2 = load_templates_from_header(hdr, template_dir=str(tmp_path), asdict=True)
    assert {k: t.filename for k, t in d2.items()} == {k: t.filename for k, t in d.items()}
    # zscan ranges reach templates without a REDSHIFTS HDU, by keyword in the
    # reference's order (galaxy, star, qso)
    write_template(str(tmp_path / 'nz.fits'), _toy('STAR', 'K'), with_redshifts=False)
    t = load_templates(str(tmp_path / 'nz.fits'), None, '-0.001,0.001,1e-4')[0]
    assert np.array_equal(t.redshifts, np.arange(-0.001, 0.001, 1e-4))


def test_eval_model_argument_checks():
    tx = _toy('GALAXY')
    data = dict(Z=np.array([0.2]), COEFF=np.ones((1, 10)), SPECTYPE=np.array(['GALAXY']),
                SUBTYPE=np.array(['']))
    wave = {'b': np.linspace(3600, 5800, 50), 'r': np.linspace(5700, 7600, 50)}
    with pytest.raises(ValueError):
        eval_model(data, wave, R={'b': None}, templates={('GALAXY', ''): tx})
    with pytest.raises(ValueError):
        eval_model(data, wave, R=[None], templates={('GALAXY', ''): tx})


def test_target_eval_model_zero_coeff():
    from redrock_amd.targets import Spectrum, Target
    wave = np.linspace(3600, 5800, 40)
    R = scipy.sparse.identity(40, format='dia')
    t = Target(7, [Spectrum(wave, np.ones(40), np.ones(40), R, band='b')])
    model = t.eval_model(dict(SPECTYPE='GALAXY', SUBTYPE='', COEFF=np.zeros(10), Z=0.5,
                              FITMETHOD='PCA'), templates={})
    assert list(model.keys()) == ['b'] and model['b'].dtype == np.float32
    assert np.all(model['b'] == 0) and t.model is model


# ---------------------------------------------------------------- on the GPU

@pytest.mark.gpu
def test_template_eval(gpu):
    """Template.eval = rebin of flux'coeff onto wave, no (1+z), then R."""
    from oracle import rr_oracle as orc
    tx = _toy('GALAXY')
    wave = np.linspace(3600, 5800, 700)
    coeff = np.array([0.5, -1.0, 2.0, 99.0])                  # extra entries ignored
    z = 0.37
    got = tx.eval(coeff, wave, z)
    want = orc.trapz_rebin(tx.wave * (1 + z), tx.flux.T.dot(coeff[:3]), wave)
    assert np.allclose(got, want, rtol=1e-12, atol=1e-12)
    # a constant basis rebins to the same constant exactly: no (1+z) factor
    flat = Template(spectype='STAR', subtype='X', redshifts=[0.0], wave=tx.wave,
                    flux=np.ones((1, tx.nwave)), version='1')
    assert np.allclose(flat.eval([3.0], wave, z), 3.0, rtol=1e-13)
    band = np.random.default_rng(0).uniform(0.1, 0.3, size=(5, len(wave)))
    R = scipy.sparse.dia_matrix((band, [-2, -1, 0, 1, 2]), shape=(len(wave), len(wave)))
    assert np.allclose(tx.eval(coeff, wave, z, R=R), R.dot(want), rtol=1e-12, atol=1e-12)


@pytest.mark.gpu
def test_eval_model_rows_and_cameras(gpu):
    tg, tq = _toy('GALAXY'), _toy('QSO', 'HIZ', nbasis=2)
    templates = {('GALAXY', ''): tg, ('QSO', 'HIZ'): tq}
    data = dict(Z=np.array([0.2, 0.9]), COEFF=np.array([[1.0, 2.0, -0.5, 0, 0], [0.3, 4.0, 0, 0, 0]]),
                SPECTYPE=np.array(['GALAXY', 'QSO']), SUBTYPE=np.array(['', 'HIZ']))
    wb, wr = np.linspace(3600, 5800, 300), np.linspace(5700, 7000, 260)
    rng = np.random.default_rng(1)
    Rb = [scipy.sparse.dia_matrix((rng.uniform(0, 0.4, (3, 300)), [-1, 0, 1]), shape=(300, 300))
          for _ in range(2)]
    got = eval_model(data, wb, R=Rb, templates=templates)
    assert got.dtype == np.float32 and got.shape == (2, 300)
    for i, tx in enumerate((tg, tq)):
        want = tx.eval(data['COEFF'][i], wb, data['Z'][i], R=Rb[i]).astype(np.float32)
        assert np.array_equal(got[i], want)
    # a dict of cameras, R None; lower-case columns as zfind writes them
    lower = {k.lower(): v for k, v in data.items()}
    got = eval_model(lower, {'b': wb, 'r': wr}, templates=templates)
    assert sorted(got.keys()) == ['b', 'r'] and got['r'].shape == (2, 260)
    assert np.array_equal(got['r'][1], tq.eval(data['COEFF'][1], wr, 0.9).astype(np.float32))


@pytest.mark.gpu
def test_target_eval_model(gpu):
    from redrock_amd import archetypes
    from redrock_amd.synth import archetypes_from
    from redrock_amd.targets import DistTargetsCopy, Spectrum, Target
    rng = np.random.default_rng(2)
    spectra = []
    for band, (lo, hi) in (('b', (3600, 5800)), ('r', (5700, 7600))):
        wave = np.linspace(lo, hi, 400)
        R = scipy.sparse.dia_matrix((rng.uniform(0, 0.4, (3, 400)), [-1, 0, 1]), shape=(400, 400))
        spectra.append(Spectrum(wave, rng.normal(size=400), np.ones(400), R, band=band))
    tgt = Target(11, spectra)
    tx = _toy('GALAXY')
    row = dict(SPECTYPE='GALAXY', SUBTYPE='', COEFF=np.array([1.0, 0.5, -0.2, 0, 0]), Z=0.3,
               FITMETHOD='PCA')
    model = tgt.eval_model(row, {('GALAXY', ''): tx})
    for sp in spectra:
        assert np.array_equal(model[sp.wavehash], tx.eval(row['COEFF'], sp.wave, 0.3, sp.Rcsr))
    # an archetype fit: COEFF holds the archetype weights, LEGCOEFF per camera
    arch = archetypes_from(tx, narch=3, seed=4, subtypes=['A', 'B', 'C'])
    legc = np.array([[0.1, 0.02], [-0.1, 0.03]])
    row = dict(SPECTYPE='GALAXY', SUBTYPE='B_1;C_2', COEFF=np.array([2.0, 0.5, 0, 0]), Z=0.3,
               FITMETHOD='ARCH', LEGCOEFF=legc)
    model = tgt.eval_model(row, {}, archetypes={'GALAXY': arch})
    for i, sp in enumerate(spectra):
        want = arch.eval('B_1;C_2', np.array([2.0, 0.5]), sp.wave, 0.3, R=sp.Rcsr, legcoeff=legc[i])
        assert np.array_equal(model[sp.wavehash], want)
    # DistTargets.eval_models finds each target's row by TARGETID
    bestfit = dict(TARGETID=np.array([5, 11]), SPECTYPE=np.array(['GALAXY', 'GALAXY']),
                   SUBTYPE=np.array(['', '']), COEFF=np.array([[0.0] * 5, [1.0, 0.5, -0.2, 0, 0]]),
                   Z=np.array([0.1, 0.3]), FITMETHOD=np.array(['PCA', 'PCA']))
    models = DistTargetsCopy([tgt]).eval_models(bestfit, {('GALAXY', ''): tx})
    assert len(models) == 1
    for sp in spectra:
        assert np.array_equal(models[0][sp.wavehash],
                              tx.eval(bestfit['COEFF'][1], sp.wave, 0.3, sp.Rcsr))
    assert isinstance(arch, archetypes.Archetype)


@pytest.mark.gpu
def test_load_dist_templates(gpu, tmp_path, monkeypatch):
    """test_templates.py:146-156 of the reference, on files written here."""
    monkeypatch.delenv('RR_TEMPLATE_DIR', raising=False)
    write_template(str(tmp_path / 'rrtemplate-GALAXY-None-v1.fits'), _toy('GALAXY'))
    write_template(str(tmp_path / 'rrtemplate-QSO-None-v1.fits'), _toy('QSO', nbasis=2))
    dwave = {'b': np.linspace(3600, 5800, 200), 'r': np.linspace(5700, 7000, 150)}
    dts = tmod.load_dist_templates(dwave, templates=str(tmp_path))
    assert [d.template.template_type for d in dts] == ['GALAXY', 'QSO']
    for d in dts:
        assert d.template.wave.ndim == 1 and d.template.flux.ndim == 2
        assert d.template.wave.shape[0] == d.template.flux.shape[1]
        assert d.cycle()
        data = d.local.data
        assert data['b'].shape == (11, 200, d.template.nbasis)
