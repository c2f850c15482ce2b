'''
Pins of the oracle (CPU). Each check is independent of the formula it
pins: exact Wigner-d sums (mpmath), textbook closed forms, an independent
closed form of the spin-2 functions (HEALPix / Zaldarriaga-Seljak), brute
force synthesis, the reference's own compiled C (oracle/_ref) and the
literal answers of the reference test-suite.
'''
import ctypes
import os

import mpmath
import numpy as np
import pytest

from oracle import legendre as L, sht as osht, geometry as og, layout, ops, cg as ocg, \
    wiener as ow

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_LIB = os.path.join(ROOT, 'oracle', '_ref', 'libref_c_utils.so')


# ------------------------------------------------------------- Legendre

@pytest.mark.parametrize('s', [0, 2, -2])
@pytest.mark.parametrize('m', [0, 1, 2, 3, 7])
def test_recurrence_vs_exact_wigner(s, m):
    th = np.array([0.05, 0.7, 1.3, 2.2, 3.0])
    lam = L.lam_m(m, 14, th, s)
    for l in range(m, 15):
        for i, t in enumerate(th):
            ex = float(L.lam_exact(l, m, s, t))
            assert abs(float(lam[l - m, i]) - ex) < 1e-15 * max(1, abs(ex))


def test_textbook_values():
    t = 0.4
    assert abs(float(L.lam_exact(0, 0, 0, t)) - 1 / np.sqrt(4 * np.pi)) < 1e-16
    assert abs(float(L.lam_exact(1, 1, 0, t)) + np.sqrt(3 / 8 / np.pi) * np.sin(t)) < 1e-16
    assert abs(float(L.lam_exact(2, 0, 2, t)) - np.sqrt(15 / 32 / np.pi) * np.sin(t) ** 2) < 1e-16
    assert abs(float(L.lam_exact(2, 0, -2, t)) - np.sqrt(15 / 32 / np.pi) * np.sin(t) ** 2) < 1e-16


@pytest.mark.parametrize('m', [0, 1, 2, 3, 5])
def test_spin2_vs_healpix_closed_form(m):
    '''F1 = (2lam + -2lam)/2, F2 = (2lam - -2lam)/2 from Wigner d agree with
    the independent P_l^m closed form, including the sign of F2.'''
    for l in range(max(2, m), 10):
        for t in (0.3, 0.9, 2.0):
            lp = float(L.lam_exact(l, m, 2, t))
            lm = float(L.lam_exact(l, m, -2, t))
            F1, F2 = L.f1f2_healpix(l, m, t)
            assert abs((lp + lm) / 2 - float(F1)) < 1e-14
            assert abs((lp - lm) / 2 - float(F2)) < 1e-14


def test_healpix_pin_catches_flipped_sign():
    '''A planted mistake (F2 sign flipped in the oracle's spin-2 functions)
    is caught by the closed-form pin.'''
    l, m, t = 4, 2, 0.8
    lp = float(L.lam_exact(l, m, 2, t))
    lm = float(L.lam_exact(l, m, -2, t))
    F1, F2 = L.f1f2_healpix(l, m, t)
    wrong_f2 = -(lp - lm) / 2
    assert abs(wrong_f2 - float(F2)) > 1e-3


def test_high_lmax_recurrence_against_legenp():
    """lmax 4095, m 3000, theta 0.85: the seed slam_mm ~ 6e-373 lies below
    the float64 range, yet slam_{4095,m} is O(1). The longdouble recurrence
    matches mpmath's associated Legendre function (60 digits)."""
    th, m, l = 0.85, 3000, 4095
    lam = L.lam_m(m, l, np.array([th]), 0)
    assert 0 < abs(lam[0, 0]) < np.longdouble('1e-330')
    with mpmath.workdps(60):
        x = mpmath.cos(mpmath.mpf(th))
        norm = mpmath.sqrt(mpmath.mpf(2 * l + 1) / (4 * mpmath.pi)
                           * mpmath.factorial(l - m) / mpmath.factorial(l + m))
        ex = float(norm * mpmath.legenp(l, m, x))
    assert abs(float(lam[-1, 0]) - ex) < 1e-15 * abs(ex)


# ------------------------------------------------------------------ SHT

def _brute_synth(a, g, lmax):
    ell, em = layout.lm_arrays(lmax)
    phi = np.concatenate([g.phi0[r] + 2 * np.pi * np.arange(g.nphi[r]) / g.nphi[r]
                          for r in range(g.nrow)])
    rr = np.concatenate([np.full(g.nphi[r], r) for r in range(g.nrow)])
    bt = np.zeros(g.npix)
    bq = np.zeros(g.npix)
    bu = np.zeros(g.npix)
    for i, (l, m) in enumerate(zip(ell, em)):
        lam0 = np.array([float(L.lam_exact(l, m, 0, t)) for t in g.theta])
        lp = np.array([float(L.lam_exact(l, m, 2, t)) for t in g.theta])
        ln = np.array([float(L.lam_exact(l, m, -2, t)) for t in g.theta])
        e = np.exp(1j * m * phi)
        fac = 1 if m == 0 else 2
        bt += fac * (a[0, i] * lam0[rr] * e).real
        F1, F2 = (lp + ln) / 2, (lp - ln) / 2
        bq += fac * (-(a[1, i] * F1[rr] + 1j * a[2, i] * F2[rr]) * e).real
        bu += fac * (-(a[2, i] * F1[rr] - 1j * a[1, i] * F2[rr]) * e).real
    return np.stack([bt, bq, bu])


def _rand_alm(lmax, npol, seed):
    rng = np.random.default_rng(seed)
    n = layout.nelem(lmax)
    a = rng.normal(size=(npol, n)) + 1j * rng.normal(size=(npol, n))
    ell, em = layout.lm_arrays(lmax)
    a[:, em == 0] = a[:, em == 0].real
    return a


def test_sht_brute_force():
    lmax = 5
    g = og.gauss_legendre(lmax + 1, 2 * lmax + 1)
    a = _rand_alm(lmax, 3, 1)
    out = np.zeros((3, g.npix))
    osht.alm2map(a, out, lmax, g, [0, 2])
    ref = _brute_synth(a, g, lmax)
    assert np.abs(out - ref).max() < 1e-13


@pytest.mark.parametrize('spin,npol', [(0, 1), (2, 2), ([0, 2], 3)])
def test_sht_roundtrip_and_adjoint(spin, npol):
    lmax = 20
    g = og.gauss_legendre(lmax + 1, 2 * lmax + 1)
    a = _rand_alm(lmax, npol, 2)
    ell, _ = layout.lm_arrays(lmax)
    if npol >= 2:
        a[npol - 2:, ell < 2] = 0
    mp = np.zeros((npol, g.npix))
    osht.alm2map(a, mp, lmax, g, spin)
    back = np.zeros_like(a)
    osht.map2alm(mp, back, g, lmax, spin)
    assert np.linalg.norm(back - a) / np.linalg.norm(a) < 1e-13
    f = np.random.default_rng(3).normal(size=(npol, g.npix))
    at = np.zeros_like(a)
    osht.map2alm(f, at, g, lmax, spin, adjoint=True)
    lhs = np.sum(mp * f)
    rhs = ops.contract_almxblm(a, at)
    assert abs(lhs - rhs) < 1e-12 * abs(lhs)


def test_sht_adjoint_equal_area():
    lmax = 16
    g = og.equal_area(og.gauss_legendre(2 * lmax + 2, 4 * lmax + 1), orig_band=0.4)
    assert np.any(g.nphi < 2 * lmax + 1)
    a = _rand_alm(lmax, 1, 4)
    mp = np.zeros((1, g.npix))
    osht.alm2map(a, mp, lmax, g, 0)
    f = np.random.default_rng(5).normal(size=(1, g.npix))
    at = np.zeros_like(a)
    osht.map2alm(f, at, g, lmax, 0, adjoint=True)
    assert abs(np.sum(mp * f) - ops.contract_almxblm(a, at)) < 1e-12 * abs(np.sum(mp * f))


def test_y00_map():
    '''a_00 = sqrt(4 pi) gives the constant map 1.'''
    lmax = 4
    g = og.gauss_legendre(lmax + 1, 2 * lmax + 1)
    a = np.zeros((1, layout.nelem(lmax)), dtype=np.complex128)
    a[0, 0] = np.sqrt(4 * np.pi)
    mp = np.zeros((1, g.npix))
    osht.alm2map(a, mp, lmax, g, 0)
    np.testing.assert_allclose(mp, 1, rtol=1e-14)


# ------------------------------------------------------ geometry / setup

def test_gl_geometry():
    '''test_map_utils.py:87-93 (sum w nphi = 4 pi) and :215-234.'''
    from scipy.special import roots_legendre
    g = og.gauss_minfo(20)
    assert abs(np.sum(g.weight * g.nphi) - 4 * np.pi) < 1e-13
    x, w = roots_legendre(11)
    np.testing.assert_allclose(np.cos(g.theta), x[::-1], atol=1e-15)
    np.testing.assert_allclose(g.weight * g.nphi, w[::-1] * 2 * np.pi, rtol=1e-14)


def test_mirror_pairs():
    g = og.gauss_legendre(9, 17)
    n, s = og.mirror_pairs(g)
    assert sorted(np.concatenate([n, s[s >= 0]]).tolist()) == list(range(9))
    assert np.sum(s < 0) == 1  # the equator ring


# ------------------------------------------------------ streaming ops

def test_contract_golden():
    '''test_alm_utils.py:411-448.'''
    a = np.asarray([1, 1, 1, 1, 2j, 2j, 2j, 3j, 3j, 4j])
    b = np.asarray([2, 2, 2, 2, 3, 3, 3, -4j, -4j, 5])
    assert ops.contract_almxblm(a, b) == -40
    assert ops.contract_almxblm(np.ones((3, 10)) * a, np.ones((3, 10)) * b) == -120
    assert ops.contract_almxblm(np.ones((2, 3, 10)) * a, np.ones((2, 3, 10)) * b) == -240


def test_trunc_wlm2alm_golden():
    '''test_alm_c_utils.py:10-59 and 387-441.'''
    alm = np.zeros((2, 6), dtype=np.complex128)
    alm[0] = [1, 2, 3, 4, 5, 6]
    alm[1] = np.asarray([1, 2, 3, 4, 5, 6]) * 2
    np.testing.assert_equal(ops.trunc_alm(alm, 2, 1), [[1, 2, 4], [2, 4, 8]])
    alm = np.ones((3, 10), dtype=np.complex128)
    alm[1] *= 2
    alm[2] *= 3
    wlm = np.ones((3, 6), dtype=np.complex128) * 1j
    wlm[1] *= 10
    wlm[2] *= 100
    wlm *= np.arange(6)
    w_ell = np.arange(4, dtype=np.float64)
    w_ell[-1] = 0
    ops.wlm2alm(w_ell, wlm, alm, 2, 3)
    np.testing.assert_allclose(alm[0], [1, 1 + 1j, 1 + 4j, 1, 1 + 3j, 1 + 8j, 1, 1 + 10j, 1, 1])
    np.testing.assert_allclose(alm[2], [3, 3 + 100j, 3 + 400j, 3, 3 + 300j, 3 + 800j, 3,
                                        3 + 1000j, 3, 3])


def test_eigpow_golden():
    '''test_mat_c_utils.py:16-46.'''
    m = np.asarray([[5, -4, 0], [-4, 5, 0], [0, 0, 9]], dtype=np.float64)[:, :, None]
    out = ops.eigpow(m, 0.5)
    np.testing.assert_allclose(out[:, :, 0], [[2, -1, 0], [-1, 2, 0], [0, 0, 3]], atol=1e-12)


def test_beam_and_lowpass_golden():
    '''test_noise_utils.py:120-135, test_multigrid.py:12-18.'''
    b = ops.band_limit_gauss_beam(10)
    assert abs(b[-1] / 0.019927032253073726 - 1) < 1e-7
    b = ops.band_limit_gauss_beam(10, fwhm=np.radians(10))
    assert abs(b[-1] / 0.73923778 - 1) < 1e-7
    r = ops.lowpass_filter(10)
    assert abs(r[0] - 1) < 1e-15
    assert abs(r[5] ** 2 - 0.05) < 1e-12


@pytest.mark.skipif(not os.path.exists(REF_LIB), reason='reference C library not built')
class TestAgainstReferenceC:
    '''The oracle's lmul / trunc / wlm2alm / ring weight against the
    reference's own C sources compiled by oracle/Makefile.'''

    @pytest.fixture(scope='class')
    def ref(self):
        return ctypes.CDLL(REF_LIB)

    @pytest.mark.parametrize('ncomp', [1, 3])
    @pytest.mark.parametrize('diag', [False, True])
    def test_lmul(self, ref, ncomp, diag):
        lmax = 23
        rng = np.random.default_rng(7)
        n = layout.nelem(lmax)
        a = rng.normal(size=(ncomp, n)) + 1j * rng.normal(size=(ncomp, n))
        mat = rng.normal(size=(ncomp, lmax + 1) if diag else (ncomp, ncomp, lmax + 1))
        out = np.zeros_like(a)
        fn = ref.lmul_diag_dp if diag else ref.lmul_dp
        fn(mat.ctypes.data_as(ctypes.c_void_p), a.ctypes.data_as(ctypes.c_void_p),
           out.ctypes.data_as(ctypes.c_void_p), ctypes.c_int(lmax), ctypes.c_int(ncomp))
        np.testing.assert_allclose(ops.lmul(a, mat, lmax), out, rtol=1e-15, atol=1e-14)

    def test_lmul_sp(self, ref):
        lmax, ncomp = 17, 3
        rng = np.random.default_rng(8)
        n = layout.nelem(lmax)
        a = (rng.normal(size=(ncomp, n)) + 1j * rng.normal(size=(ncomp, n))).astype(np.complex64)
        mat = rng.normal(size=(ncomp, ncomp, lmax + 1)).astype(np.float32)
        out = np.zeros_like(a)
        ref.lmul_sp(mat.ctypes.data_as(ctypes.c_void_p), a.ctypes.data_as(ctypes.c_void_p),
                    out.ctypes.data_as(ctypes.c_void_p), ctypes.c_int(lmax), ctypes.c_int(ncomp))
        np.testing.assert_allclose(ops.lmul(a, mat, lmax), out, rtol=1e-5, atol=1e-5)

    def test_trunc_and_wlm2alm(self, ref):
        lmax, lmax_out, ncomp = 19, 8, 2
        rng = np.random.default_rng(9)
        a = rng.normal(size=(ncomp, layout.nelem(lmax))) + 0j
        out = np.zeros((ncomp, layout.nelem(lmax_out)), dtype=np.complex128)
        ref.trunc_alm_dp(a.ctypes.data_as(ctypes.c_void_p), out.ctypes.data_as(ctypes.c_void_p),
                         ctypes.c_int(lmax), ctypes.c_int(lmax_out), ctypes.c_int(ncomp))
        np.testing.assert_array_equal(ops.trunc_alm(a, lmax, lmax_out), out)
        w = rng.normal(size=lmax_out + 1)
        wlm = rng.normal(size=(ncomp, layout.nelem(lmax_out))) + 1j
        acc = a.copy()
        ref.wlm2alm_dp(w.ctypes.data_as(ctypes.c_void_p), wlm.ctypes.data_as(ctypes.c_void_p),
                       acc.ctypes.data_as(ctypes.c_void_p), ctypes.c_int(lmax_out),
                       ctypes.c_int(lmax), ctypes.c_int(ncomp))
        mine = ops.wlm2alm(w, wlm, a.copy(), lmax_out, lmax)
        np.testing.assert_allclose(mine, acc, rtol=1e-15, atol=1e-15)

    def test_ringweight(self, ref):
        g = og.gauss_legendre(9, 17)
        f = np.random.default_rng(10).normal(size=g.npix)
        out = f.copy()
        nphi = g.nphi.astype(np.int64)
        offs = g.offsets.astype(np.int64)
        stride = np.ones(g.nrow, dtype=np.int64)
        ref._apply_ringweight_core_dp(out.ctypes.data_as(ctypes.c_void_p),
                                      g.weight.ctypes.data_as(ctypes.c_void_p),
                                      nphi.ctypes.data_as(ctypes.c_void_p),
                                      offs.ctypes.data_as(ctypes.c_void_p),
                                      stride.ctypes.data_as(ctypes.c_void_p),
                                      ctypes.c_long(g.nrow))
        np.testing.assert_array_equal(osht.ringweight(f.copy(), g), out)


# ------------------------------------------------------------------ CG

def test_cg_closed_forms():
    '''test_solvers.py:45-111 with the oracle solver.'''
    alm = np.asarray([1, 2, 3, 4j, 5j, 6])
    s = ow.WienerMap(alm, lambda a: 2 * a, lambda a: 3 * a, (lambda a: a, lambda a: a))
    cg = s.init_solver()
    b_in = cg.b.copy()
    cg.step()
    np.testing.assert_allclose(cg.x, alm * 3 / 5, atol=1e-14)
    np.testing.assert_allclose(cg.A(cg.x), b_in, atol=1e-13)
    assert abs(s.get_qform() - (0.5 * 168 * (3 / 5) ** 2 * 5 - 168 * 3 * (3 / 5))) < 1e-10
    s = ow.WienerMap(np.ones(6), lambda a: 2 * a, lambda a: 3 * a, (lambda a: a, lambda a: a),
                     beam=lambda a: 0.1 * a, rand_isignal=np.arange(6.),
                     rand_inoise=np.arange(6.)[::-1])
    cg = s.init_solver()
    cg.step()
    np.testing.assert_allclose(cg.x, (0.3 + np.arange(6) + 0.1 * np.arange(6)[::-1]) / 2.03,
                               atol=1e-14)


def test_pinv_exact_fullsky():
    '''Full-sky white noise: the pseudo-inverse equals A^-1 (one CG step).'''
    lmax = 12
    g = og.gauss_legendre(lmax + 1, 2 * lmax + 1)
    ell = np.arange(lmax + 1)
    cl = np.zeros(lmax + 1)
    cl[2:] = 1e3 * (ell[2:] + 10.) ** -2.5
    icov_ell = np.zeros((1, lmax + 1))
    icov_ell[0, 2:] = 1 / cl[2:]
    icov_pix = np.repeat(g.weight, g.nphi)[None] / 4.
    imap = np.random.default_rng(11).normal(size=(1, g.npix))
    s = ow.WienerMap.from_arrays(imap, g, lmax, icov_ell, icov_pix)
    s.add_preconditioner(ow.PseudoInvPreconditioner(lmax, icov_ell, icov_pix, g, 0))
    s.init_solver()
    s.cg.step()
    b = np.sqrt(ops.contract_almxblm(s.b0, s.b0))
    assert s.get_residual() / b < 1e-12


def test_multigrid_levels():
    '''test_multigrid.py:20-84 properties on the oracle.'''
    lmax = 22
    ells = np.arange(lmax + 1)
    icov = np.ones((3, 3, lmax + 1)) * np.asarray(
        [[1, 0.1, 0.1], [0.1, 1, 0.1], [0.1, 0.1, 1]])[:, :, None]
    icov *= 1e-4 * (ells + 1) ** 2
    icov *= ops.gauss_beam(2 * np.pi / lmax, lmax)
    icov[:, :, :2] = 0
    g = og.gauss_minfo(2 * lmax)
    mask = np.ones(g.npix, dtype=bool)
    m2 = og.view_2d(mask, g)
    m2[0:5, :] = False
    m2[21:, :] = False
    levels = ow.get_levels(mask, g, icov, [0, 2], min_pix=10, lmax_r_ell=lmax)
    assert len(levels) == 3
    assert [og.minfo2lmax(lv.geom) for lv in levels] == [lmax, lmax // 2, lmax // 4]
    assert levels[0].pinv_g is None and levels[2].pinv_g is not None


# ------------------------------------------- masked CG and two-level

def test_add_to_alm_golden():
    '''test_alm_utils.py:191-245 (literal answers).'''
    alm = np.asarray([1, 2, 3, 4, 5, 6, 7, 8, 9, 10], dtype=np.complex64)
    blm = np.asarray([1, 2, 3, 4, 5j, 6j, 7j, 8j, 9j, 10j], dtype=np.complex64)
    ops.add_to_alm(alm, blm, 3, 3)
    np.testing.assert_array_equal(
        alm, np.asarray([2, 4, 6, 8, 5 + 5j, 6 + 6j, 7 + 7j, 8 + 8j, 9 + 9j, 10 + 10j]))
    alm = np.asarray([1, 2, 3, 4, 5, 6, 7, 8, 9, 10], dtype=np.complex64)
    blm = np.asarray([1, 2, 3, 4j, 5j, 6j], dtype=np.complex64)
    ops.add_to_alm(alm, blm, 3, 2)
    np.testing.assert_array_equal(alm, np.asarray([2, 4, 6, 4, 5 + 4j, 6 + 5j, 7, 8 + 6j, 9, 10]))
    alm = np.asarray([1, 2, 3, 4, 5, 6, 7, 8, 9, 10], dtype=np.complex64)
    ops.add_to_alm(alm, blm, 3, 2, overwrite=True)
    np.testing.assert_array_equal(alm, np.asarray([1, 2, 3, 4, 4j, 5j, 7, 6j, 9, 10]))
    with pytest.raises(ValueError):
        ops.add_to_alm(blm.copy(), alm, 2, 3)


def _spd_icov(npol, lmax, seed):
    rng = np.random.default_rng(seed)
    ell = np.arange(lmax + 1)
    icov = np.zeros((npol, npol, lmax + 1))
    for l in range(lmax + 1):
        b = rng.normal(size=(npol, npol)) * 0.2
        icov[:, :, l] = (np.eye(npol) + b @ b.T) * (ell[l] + 2.0) ** 2
    return icov


def test_pix_ell_pix_inverse_closed_form():
    '''On a Gauss-Legendre grid (Y^T W Y = I at the band limit) the inner
    preconditioner P = W Y icov^-1 Y^T W of MaskedPreconditionerCG sends
    Y icov a to W Y a. A planted mistake (weights on the wrong side, or the
    power +1 instead of -1) fails the same check.'''
    lmax = 12
    g = og.gauss_legendre(lmax + 1, 2 * lmax + 1)
    icov = _spd_icov(3, lmax, 5)
    # no T-E/B coupling at l < 2, where the spin-2 harmonics vanish
    icov[0, 1:, :2] = 0
    icov[1:, 0, :2] = 0
    a = np.random.default_rng(6).normal(size=(3, layout.nelem(lmax))) * (1 + 1j)
    ell, em = layout.lm_arrays(lmax)
    a[:, em == 0] = a[:, em == 0].real
    a[1:, ell < 2] = 0
    x = np.zeros((3, g.npix))
    ia = ops.lmul(a, icov, lmax)
    ia[1:, ell < 2] = 0
    osht.alm2map(ia, x, lmax, g, [0, 2])
    want = np.zeros((3, g.npix))
    osht.alm2map(a, want, lmax, g, [0, 2], adjoint=True)
    ones = np.ones((3, g.npix), dtype=bool)
    P = ow.pix_ell_pix_op(ones, icov, g, [0, 2], lmax=lmax, power_x=-1, adjoint=True)
    got = P(x)
    assert np.linalg.norm(got - want) / np.linalg.norm(want) < 1e-12
    for bad in (ow.pix_ell_pix_op(ones, icov, g, [0, 2], lmax=lmax, power_x=-1),
                ow.pix_ell_pix_op(ones, icov, g, [0, 2], lmax=lmax, adjoint=True)):
        assert np.linalg.norm(bad(x) - want) / np.linalg.norm(want) > 1e-3


def _dense_masked_solve(P, alm, lmax_in):
    '''r Y^T M_u x with x the least-squares solution of G x = M_u Y r a_lo
    (dense G over the unobserved pixels): the converged masked CG.'''
    mu = P.mask_unobs
    npol, npix = mu.shape
    idx = [(c, k) for c in range(npol) for k in np.nonzero(mu[c])[0]]
    G = np.zeros((len(idx), len(idx)))
    for j, (c, k) in enumerate(idx):
        e = np.zeros((npol, npix))
        e[c, k] = 1
        ge = P.g_op(e)
        G[:, j] = [ge[cc, kk] for cc, kk in idx]
    a_lo = ops.trunc_alm(alm, lmax_in, P.lmax)
    if P.r_ell is not None:
        a_lo = ops.lmul(a_lo, P.r_ell[:P.lmax + 1], P.lmax)
    imap = np.zeros((npol, npix))
    osht.alm2map(a_lo, imap, P.lmax, P.geom, P.spin)
    b = np.asarray([imap[c, k] for c, k in idx])
    x = np.linalg.lstsq(G, b, rcond=1e-13)[0]
    xm = np.zeros((npol, npix))
    for v, (c, k) in zip(x, idx):
        xm[c, k] = v
    out_lo = np.zeros_like(a_lo)
    osht.map2alm(xm, out_lo, P.geom, P.lmax, P.spin, adjoint=True)
    if P.r_ell is not None:
        out_lo = ops.lmul(out_lo, P.r_ell[:P.lmax + 1], P.lmax)
    out = np.zeros_like(alm)
    for m in range(P.lmax + 1):
        i0 = layout.mstart(m, lmax_in)
        j0 = layout.mstart(m, P.lmax)
        out[..., i0 + m:i0 + P.lmax + 1] = out_lo[..., j0 + m:j0 + P.lmax + 1]
    return out


@pytest.mark.parametrize('spin,npol,lmax_r_ell', [(0, 1, None), ([0, 2], 3, None), ([0, 2], 3, 30)])
def test_masked_cg_brute_force(spin, npol, lmax_r_ell):
    '''MaskedPreconditionerCG run for more CG steps than the masked system
    has unknowns equals the dense least-squares solve of that system (the
    output r Y^T M_u x is the same for every solution x), with truncation
    to the preconditioner's lmax and the low-pass filter.'''
    lmax_in, lmax = 12, 9
    g = og.gauss_legendre(lmax_in + 1, 2 * lmax_in + 1)
    th = np.repeat(g.theta, g.nphi)
    mask = th > 0.8
    icov = _spd_icov(npol, lmax_in, 7) if npol == 3 else \
        ((np.arange(lmax_in + 1) + 2.0) ** 2)[None]
    P = ow.MaskedPreconditionerCG(lmax_in, icov, spin, mask, g, lmax_r_ell=lmax_r_ell,
                                  lmax=lmax, nsteps=1)
    nunk = int(P.mask_unobs.sum())
    P.nsteps = nunk + 2
    a = np.random.default_rng(8).normal(size=(npol, layout.nelem(lmax_in))) * (1 - 1j)
    ell, em = layout.lm_arrays(lmax_in)
    a[:, em == 0] = a[:, em == 0].real
    if npol == 3:
        a[1:, ell < 2] = 0
    got = P(a)
    ref = _dense_masked_solve(P, a, lmax_in)
    assert np.linalg.norm(got - ref) / np.linalg.norm(ref) < 1e-9
    hi = ell > lmax
    assert np.all(got[:, hi] == 0)


def test_adef_closed_forms():
    '''Two-level preconditioners on a dense SPD system: with Q = 0 both
    reduce to the main preconditioner M; with Q = A^-1 both give A^-1 b
    exactly; with Q on a slice only, the other rows see M alone.'''
    rng = np.random.default_rng(9)
    n = 8
    B = rng.normal(size=(n, n))
    A = B @ B.T + n * np.eye(n)
    C = rng.normal(size=(n, n))
    M = C @ C.T + np.eye(n)
    Ainv = np.linalg.inv(A)
    a_op = lambda v: A @ v
    m_op = lambda v: M @ v
    b = rng.normal(size=n)
    for method in ('ADEF-1', 'ADEF-2'):
        f0 = ow.get_2level_prec(m_op, lambda v: 0 * v, a_op, method)
        np.testing.assert_allclose(f0(b), M @ b, rtol=1e-13)
        fi = ow.get_2level_prec(m_op, lambda v: Ainv @ v, a_op, method)
        np.testing.assert_allclose(fi(b), Ainv @ b, rtol=1e-10, atol=1e-12)
    # ADEF-2 with a block-diagonal A and Q = (A^-1 on the first block)
    A2 = A.copy()
    A2[:4, 4:] = 0
    A2[4:, :4] = 0
    Q1 = np.linalg.inv(A2[:4, :4])
    a2 = lambda v: A2 @ v
    for method in ('ADEF-1', 'ADEF-2'):
        f = ow.get_2level_prec(m_op, lambda v: Q1 @ v, a2, method, sel_masked=np.s_[:4])
        out = f(b)
        if method == 'ADEF-2':
            mb = M @ b
            am = A2 @ mb
            want = mb.copy()
            want[:4] += Q1 @ b[:4] - Q1 @ am[:4]
        else:
            q = np.zeros(n)
            q[:4] = Q1 @ b[:4]
            want = M @ (b - A2 @ q) + q
        np.testing.assert_allclose(out, want, rtol=1e-12)
    with pytest.raises(ValueError):
        ow.get_2level_prec(m_op, m_op, a_op, 'ADEF-3')


# ------------------------------------------- constant-correlation noise

def _gl_pol_setup(lmax, seed):
    g = og.gauss_legendre(lmax + 1, 2 * lmax + 1)
    x = _spd_icov(3, lmax, seed)
    x[0, 1:, :2] = 0
    x[1:, 0, :2] = 0
    z = np.random.default_rng(seed + 1).normal(size=(3, layout.nelem(lmax))) * (1 - 1j)
    ell, em = layout.lm_arrays(lmax)
    z[:, em == 0] = z[:, em == 0].real
    z[1:, ell < 2] = 0
    return g, x, z


def test_inv_pix_ell_pix_closed_form():
    '''With M_pix = c (no mask) on a GL grid, A = c^2p Y X^q Y^T and its
    preconditioner c^-2p W Y X^-q Y^T W invert each other on range(W Y), so
    one CG step gives A^-1 (Y z) = c^-2p W Y X^-q z exactly (here p = -1/2,
    q = -1 as in the constant-correlation noise model), and the inverse
    square root gives c^-p W Y X^-q/2 a. A preconditioner without the
    flipped weight side misses after one step.'''
    lmax = 10
    g, x, z = _gl_pol_setup(lmax, 20)
    c = 2.5
    m_pix = np.full((3, g.npix), c)
    yz = np.zeros((3, g.npix))
    osht.alm2map(z, yz, lmax, g, [0, 2])
    inv = ow.inv_pix_ell_pix_op(m_pix, x, g, [0, 2], power_m=-0.5, power_x=-1, nsteps=1)
    want = np.zeros((3, g.npix))
    osht.alm2map(ops.lmul(z, x, lmax), want, lmax, g, [0, 2], adjoint=True)
    want *= c
    assert np.linalg.norm(inv(yz) - want) / np.linalg.norm(want) < 1e-12
    sq = ow.inv_sqrt_pix_ell_pix_op(m_pix, x, g, [0, 2], power_m=-0.5, power_x=-1, nsteps=1)
    want2 = np.zeros((3, g.npix))
    osht.alm2map(ops.lmul(z, ops.matpow(x, 0.5, return_diag=True), lmax), want2, lmax, g,
                 [0, 2], adjoint=True)
    want2 *= np.sqrt(c)
    assert np.linalg.norm(sq(z) - want2) / np.linalg.norm(want2) < 1e-12
    # planted mistake: preconditioner with the weights on the same side as A
    am = ow.pix_ell_pix_op(m_pix, x, g, [0, 2], power_x=-1, power_m=-0.5)
    bad_p = ow.pix_ell_pix_op(m_pix, x, g, [0, 2], power_x=1, power_m=0.5)
    cg = ocg.CG(am, yz.copy(), M=bad_p, dot=lambda u, v: np.sum(u * v))
    cg.step()
    assert np.linalg.norm(cg.x - want) / np.linalg.norm(want) > 1e-3


def test_inv_pix_ell_pix_masked_brute_force():
    '''With a mask and a varying pixel matrix, 50 CG steps (more than the
    34 observed unknowns) solve the dense masked system M' A M' x = b on the
    observed pixels (b in its range: the system is rank deficient, two rings
    share the m = lmax mode), and leaves the masked pixels at zero.'''
    lmax = 8
    g = og.gauss_legendre(lmax + 1, 2 * lmax + 1)
    rng = np.random.default_rng(21)
    m_pix = rng.uniform(0.5, 2.0, size=(1, g.npix))
    x_ell = ((np.arange(lmax + 1) + 1.0) ** -2)[None]
    th = np.repeat(g.theta, g.nphi)
    mask = (th < 0.9)[None].astype(float)
    idx = np.nonzero(mask[0])[0]
    a_mat = ow.pix_ell_pix_op(m_pix, x_ell, g, 0, power_x=-1, power_m=-0.5)
    A = np.zeros((len(idx), len(idx)))
    for j, k in enumerate(idx):
        e = np.zeros((1, g.npix))
        e[0, k] = 1
        A[:, j] = a_mat(e)[0, idx]
    b = mask * a_mat(mask * rng.normal(size=(1, g.npix)))
    inv = ow.inv_pix_ell_pix_op(m_pix, x_ell, g, 0, power_m=-0.5, power_x=-1, mask=mask,
                                nsteps=50)
    got = inv(b)
    res = A @ got[0, idx] - b[0, idx]
    # residual 3e-3 after 10 steps, 1e-7 after 50 (rank-deficient system)
    assert np.linalg.norm(res) / np.linalg.norm(b) < 1e-6
    assert np.all(got[0, mask[0] == 0] == 0)
