"""The reference's own recovery tests on the GPU path.

* test_zscan.py:52-99: two fake targets at z = 0.2 and 0.25, the fake
  3-vector template on 50 redshifts in [0.15, 0.3]: zfind gives zerr <
  0.002, and with a Gaussian prior (sigma 0.01 at the true z, written to a
  PRIORS FITS file) |z - ztrue| / zerr < 5 as well.
* test_zscan.py:174-202: two STAR subtypes (F, M) on 25 redshifts in
  [-1e-3, 1e-3]: every row of the zfind table has nminima coefficients and
  spectype STAR.
* The truncated resolution rows of those fake targets (targets.py:33-35):
  the product's masked ivar equals the oracle's, and the FP64 scan of the
  targets equals the oracle's scan of the oracle-masked spectra (<= 1e-9,
  same best z).

Fake data: tests/refsim.py (the reference's test/util.py recipe, fixed seeds).
"""
import numpy as np
import pytest

from oracle import rr_oracle as orc
from tests import refsim

pytestmark = pytest.mark.gpu


def _dist(targets):
    from redrock_amd.targets import DistTargetsCopy
    return DistTargetsCopy(targets)


def _best(zfit, tid):
    i = [k for k, t in enumerate(zfit['targetid']) if t == tid and zfit['znum'][k] == 0][0]
    return zfit['z'][i], zfit['zerr'][i]


@pytest.mark.parametrize('precision', ['f64', 'f16r'])
def test_zfind_recovers_redshift(gpu, tmp_path, precision):
    from redrock_amd import fits
    from redrock_amd.templates import DistTemplate
    from redrock_amd.zfind import zfind
    z1, z2 = 0.2, 0.25
    rng = np.random.default_rng(20240611)
    t1 = refsim.target(111, z1, rng)
    t2 = refsim.target(222, z2, rng)
    dtarg = _dist([t1, t2])
    dwave = dtarg.wavegrids()
    dtemp = DistTemplate(refsim.template(redshifts=np.linspace(0.15, 0.3, 50)), dwave)
    print(f'case: zfind recovery {precision}, z = {z1}, {z2}', flush=True)
    _, zfit = zfind(dtarg, [dtemp], precision=precision)
    for tid, ztrue in ((111, z1), (222, z2)):
        z, zerr = _best(zfit, tid)
        print(f'target {tid}: z {z:.6f} (true {ztrue}), zerr {zerr:.2e}')
        assert zerr < 0.002
    # Gaussian priors from a PRIORS FITS file, as the reference writes it
    path = str(tmp_path / 'priors.fits')
    fits.write(path, [fits.table_hdu({'TARGETID': np.array([111, 222]), 'Z': np.array([z1, z2]),
                                      'SIGMA': np.array([0.01, 0.01]),
                                      'FUNCTION': np.array(['gaussian', 'gaussian'])},
                                     name='PRIORS')])
    _, zfit = zfind(dtarg, [dtemp], priors=path, precision=precision)
    for tid, ztrue in ((111, z1), (222, z2)):
        z, zerr = _best(zfit, tid)
        print(f'target {tid} with prior: z {z:.6f}, |dz|/zerr {abs(z - ztrue) / zerr:.2f}')
        assert abs(z - ztrue) / zerr < 5
        assert zerr < 0.002


@pytest.mark.parametrize('precision', ['f64', 'f16r'])
def test_zfind_star_subtypes(gpu, precision):
    from redrock_amd.templates import DistTemplate
    from redrock_amd.zfind import zfind
    rng = np.random.default_rng(7)
    t1 = refsim.target(111, 0.0, rng)
    t2 = refsim.target(222, 1e-4, rng)
    dtarg = _dist([t1, t2])
    dwave = dtarg.wavegrids()
    zg = np.linspace(-1e-3, 1e-3, 25)
    F = DistTemplate(refsim.template(spectype='STAR', subtype='F', redshifts=zg), dwave)
    M = DistTemplate(refsim.template(spectype='STAR', subtype='M', redshifts=zg), dwave)
    print(f'case: zfind STAR F/M subtypes {precision}', flush=True)
    nminima = 3
    _, zfit = zfind(dtarg, [F, M], mp_procs=1, nminima=nminima, precision=precision)
    coeff = np.asarray(zfit['coeff'])
    assert coeff.ndim == 2 and coeff.shape[1] == nminima
    assert all(str(s).strip() == 'STAR' for s in zfit['spectype'])


def test_truncated_resolution_rows_match_oracle(gpu):
    from types import SimpleNamespace
    from redrock_amd import zscan
    from redrock_amd.templates import DistTemplate
    rng = np.random.default_rng(3)
    targets = [refsim.target(i, z, rng) for i, z in enumerate((0.18, 0.22, 0.27))]
    dwave = {s.wavehash: s.wave for s in targets[0].spectra}
    tm = refsim.template(redshifts=np.linspace(0.15, 0.3, 50))
    # the oracle's spectra: the raw ivar masked by its own restatement of the rule
    ora = []
    for tg in targets:
        sp = []
        for s in tg.spectra:
            ivar = orc.masked_ivar(s.R, s.raw_ivar)
            assert np.count_nonzero(ivar == 0) == 6           # rows 0-2 and n-3..n-1
            assert np.array_equal(ivar, s.ivar)
            sp.append(SimpleNamespace(wave=s.wave, flux=s.flux, ivar=ivar, R=s.R, wavehash=s.wavehash))
        ora.append(SimpleNamespace(id=tg.id, spectra=sp))
    print('case: truncated R rows, f64 scan vs oracle', flush=True)
    dt = DistTemplate(tm, dwave)
    chi2, coeff, pen = zscan.calc_zchi2([t.id for t in targets], targets, dt, precision='f64')
    # the in-memory template's IGM model (Inoue14) cannot attenuate these rest
    # wavelengths (> 3000 A): no line table, no transmission
    tdata = orc.template_cache(tm, dwave, lines=[])
    rchi2, rcoeff, rpen = orc.calc_zchi2(ora, tdata, tm.nbasis, tm.OIItemplate)
    rel = np.abs(chi2 - rchi2) / np.abs(rchi2)
    print('max rel chi2 err', rel.max())
    assert rel.max() <= 1e-9
    assert np.array_equal(chi2.argmin(axis=1), rchi2.argmin(axis=1))
