'''
Host-side logic of the product (CPU): geometry, set-up matrices, spin
groups, workloads and seeded draws agree with the oracle exactly.
'''
import numpy as np
import pytest

from oracle import geometry as og, ops as oops
from optweight_amd import map_utils, mat_utils, alm_utils, sht, workloads, noise_utils, \
    preconditioners


def test_gl_minfo_matches_oracle():
    m = map_utils.MapInfo.map_info_gauss_legendre(33, 65)
    g = og.gauss_legendre(33, 65)
    np.testing.assert_array_equal(m.theta, g.theta)
    np.testing.assert_array_equal(m.weight, g.weight)
    np.testing.assert_array_equal(m.nphi, g.nphi)
    np.testing.assert_array_equal(m.offsets, g.offsets)


def test_equal_area_and_fullsky_match_oracle():
    m = map_utils.get_gauss_minfo(2 * 40)
    g = og.gauss_minfo(2 * 40)
    mf = map_utils.get_fullsky_geometry(m)
    gf = og.fullsky_geometry(g)
    np.testing.assert_array_equal(mf.theta, gf.theta)
    me = map_utils.get_equal_area_minfo(mf, orig_band=0.4)
    ge = og.equal_area(gf, orig_band=0.4)
    np.testing.assert_array_equal(me.nphi, ge.nphi)
    np.testing.assert_array_equal(me.weight, ge.weight)
    assert map_utils.minfo2lmax(m) == 40


def test_spin_groups():
    assert sht.spin_groups([0, 2], 3) == [(0, 0, 1), (2, 1, 3)]
    assert sht.spin_groups(0, 1) == [(0, 0, 1)]
    assert sht.spin_groups(2, 2) == [(2, 0, 2)]
    assert sht.default_spin((3, 10)) == [0, 2]
    with pytest.raises(ValueError):
        sht.spin_groups(2, 3)


def test_matpow_matches_oracle():
    rng = np.random.default_rng(1)
    a = rng.normal(size=(3, 3, 50))
    a = np.einsum('ijn,kjn->ikn', a, a)
    for p in (-1, 0.5, -0.5):
        np.testing.assert_allclose(mat_utils.matpow(a, p), oops.matpow(a, p), rtol=1e-12,
                                   atol=1e-12)
    d = np.abs(rng.normal(size=(3, 40)))
    d[0, 3] = 0
    np.testing.assert_allclose(mat_utils.matpow(d, -1), oops.matpow(d, -1))


def test_eigpow_lim_behaviour():
    '''test_mat_c_utils.py:62-92: eigenvalues under lim * max and negative
    eigenvalues are zeroed; all zero under lim0.'''
    m = np.diag([1., 1e-8, -1.])[:, :, None]
    out = mat_utils.eigpow(m, 1, lim=1e-6)
    np.testing.assert_allclose(out[:, :, 0], np.diag([1., 0, 0]), atol=1e-15)
    out = mat_utils.eigpow(m * 1e-200, 1, lim0=1e-100)
    assert not np.any(out)


def test_unit_var_alm_matches_oracle():
    ainfo = alm_utils.AlmInfo(20)
    a = alm_utils.unit_var_alm(ainfo, (3,), 5)
    b = oops.unit_var_alm(20, 3, 5)
    np.testing.assert_array_equal(a, b)
    assert np.all(a[:, :21].imag == 0)


def test_isotropic_ivar_matches_oracle():
    m = map_utils.MapInfo.map_info_gauss_legendre(17, 33)
    g = og.gauss_legendre(17, 33)
    icov = np.random.default_rng(2).uniform(size=(3, m.npix))
    np.testing.assert_allclose(map_utils.get_isotropic_ivar(icov, m),
                               oops.get_isotropic_ivar(icov, g), rtol=1e-14)


def test_beam_golden():
    assert abs(noise_utils._band_limit_gauss_beam(10)[-1] / 0.019927032253073726 - 1) < 1e-7


def test_itau_ell_harm_golden():
    '''test_preconditioners.py:50-72.'''
    w_ell = np.zeros((2, 4), dtype=np.float32)
    w_ell[0, :2] = 1
    w_ell[1, 2:] = 0.1
    icov = np.ones((3, 4), dtype=np.float32)
    out = preconditioners.get_itau_ell_harm(icov, w_ell)
    exp = np.asarray([1, 1, 0.01, 0.01]) * np.eye(3)[:, :, None]
    np.testing.assert_allclose(out, exp, rtol=1e-6)


def test_workload_mask_and_spectra():
    m = workloads.gl_minfo(63)
    obs = workloads.make_mask(m, 0, ndisk=20)
    w = np.repeat(m.weight, m.nphi)
    frac = np.sum(w * obs) / (4 * np.pi)
    assert 0.80 < frac < 0.8264  # |lat| < 10 deg removes sin(10 deg) = 17.4% of the area
    cov = workloads.cov_ell_TQU(63)
    assert np.all(cov[:, :, :2] == 0)
    np.testing.assert_allclose(cov[0, 1], 0.3 * np.sqrt(cov[0, 0] * cov[1, 1]))
    icov = workloads.pinv_ell(cov)
    np.testing.assert_allclose(np.einsum('ijl,jkl->ikl', icov[:, :, 2:], cov[:, :, 2:]),
                               np.eye(3)[:, :, None] * np.ones(62), atol=1e-9)
    d = workloads.depth_map(m, 1)
    assert abs(d.min() - 0.5) < 1e-12 and abs(d.max() - 2.0) < 1e-12


def test_log_bands():
    from optweight_amd import workloads
    e = workloads.log_bands(2, 1023, 30)
    assert e.size == 31 and e[0] == 2 and e[-1] == 1024 and np.all(np.diff(e) >= 1)
    # rounding collisions at low l are pushed up by one
    e = workloads.log_bands(2, 40, 20)
    assert e.size == 21 and np.all(np.diff(e) >= 1) and e[-1] == 41
    with pytest.raises(ValueError):
        workloads.log_bands(2, 20, 30)
