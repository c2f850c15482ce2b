'''
Pins of the oracle's wavelet noise model and harmonic-space Wiener solver
(oracle/wavelet.py) by closed forms, no GPU needed.

With per-scale inverse covariance maps proportional to the quadrature
weights, icov_j = c_j W on the Gauss-Legendre grid of lmax_j, the exact
quadrature Y^T W Y = I turns the operators into per-multipole weights:
  WavMatVecAlm                      = sum_j c_j w_j(l)^2
  WavMatVecAlm(power=-1, adjoint)   = sum_j w_j(l)^2 / c_j
  get_ivar_ell                      = sum_j c_j w_j(l)^2
(the same reduction the reference's noisebox test uses,
test_operators.py:291-335). A dropped quadrature weight, a kernel applied
once instead of twice or a wrong scale lmax breaks these.
'''
import numpy as np
import pytest

from oracle import geometry as og, layout, ops, wavelet as ow
from oracle.wiener import HarmonicPreconditioner


def kernels(lmax, nwav=3):
    '''Cosine-squared partition of unity over l (sum_j w_j^2 = 1).'''
    ells = np.arange(lmax + 1)
    edges = np.linspace(0, lmax, nwav)
    w = np.zeros((nwav, lmax + 1))
    for j in range(nwav):
        if j == 0:
            x = np.clip((ells - edges[0]) / (edges[1] - edges[0]), 0, 1)
            w[j] = np.cos(np.pi / 2 * x)
        elif j == nwav - 1:
            x = np.clip((ells - edges[j - 1]) / (edges[j] - edges[j - 1]), 0, 1)
            w[j] = np.sin(np.pi / 2 * x)
        else:
            x1 = np.clip((ells - edges[j - 1]) / (edges[j] - edges[j - 1]), 0, 1)
            x2 = np.clip((ells - edges[j]) / (edges[j + 1] - edges[j]), 0, 1)
            w[j] = np.where(ells <= edges[j], np.sin(np.pi / 2 * x1), np.cos(np.pi / 2 * x2))
    w[np.abs(w) < 1e-15] = 0
    return w


def uniform_wav(lmax, w_ell, cs, npol):
    maps, geoms = {}, {}
    for j, c in enumerate(cs):
        lj = ow.wav_lmax(w_ell, j)
        g = og.gauss_legendre(lj + 1, 2 * lj + 1)
        wpix = np.repeat(g.weight, g.nphi)
        maps[(j, j)] = c * wpix * np.ones((npol, 1))
        geoms[(j, j)] = g
    return maps, geoms


def rand_alm(lmax, npol, seed, spin2=False):
    rng = np.random.default_rng(seed)
    n = layout.nelem(lmax)
    a = rng.normal(size=(npol, n)) + 1j * rng.normal(size=(npol, n))
    ell, em = layout.lm_arrays(lmax)
    a[:, em == 0] = a[:, em == 0].real
    if spin2:
        a[1:, ell < 2] = 0
    return a


@pytest.mark.parametrize('npol,spin', [(1, 0), (3, [0, 2])])
def test_wav_op_closed_form(npol, spin):
    lmax = 14
    w = kernels(lmax)
    cs = [3.0, 0.5, 7.0]
    maps, geoms = uniform_wav(lmax, w, cs, npol)
    a = rand_alm(lmax, npol, 1, spin2=npol == 3)
    ell, _ = layout.lm_arrays(lmax)
    weight = np.sum(np.asarray(cs)[:, None] * w ** 2, axis=0)
    got = ow.wav_op(lmax, maps, geoms, w, spin)(a)
    np.testing.assert_allclose(got, weight[ell] * a, rtol=1e-11, atol=1e-12)
    inv = np.sum(w ** 2 / np.asarray(cs)[:, None], axis=0)
    got = ow.wav_op(lmax, maps, geoms, w, spin, power=-1, adjoint=True)(a)
    np.testing.assert_allclose(got, inv[ell] * a, rtol=1e-11, atol=1e-12)


def test_wav_op_catches_missing_weight():
    '''Without the quadrature weights (adjoint=False on power -1) the
    closed form of the weighted operator does not hold.'''
    lmax = 10
    w = kernels(lmax)
    maps, geoms = uniform_wav(lmax, w, [2.0, 1.0, 4.0], 1)
    a = rand_alm(lmax, 1, 2)
    ell, _ = layout.lm_arrays(lmax)
    inv = np.sum(w ** 2 / np.asarray([2.0, 1.0, 4.0])[:, None], axis=0)
    got = ow.wav_op(lmax, maps, geoms, w, 0, power=-1, adjoint=False)(a)
    assert np.max(np.abs(got - inv[ell] * a)) > 1e-3


def test_get_ivar_ell_closed_form():
    lmax = 12
    w = kernels(lmax)
    cs = [2.0, 5.0, 0.25]
    maps, geoms = uniform_wav(lmax, w, cs, 1)
    got = ow.get_ivar_ell(maps, geoms, w)
    np.testing.assert_allclose(got[0, 0], np.sum(np.asarray(cs)[:, None] * w ** 2, axis=0),
                               rtol=1e-12)


def test_gauss2gauss_identity_on_same_grid():
    g = og.gauss_legendre(9, 17)
    rng = np.random.default_rng(3)
    m = rng.normal(size=(2, g.npix))
    for order in (1, 3):
        np.testing.assert_allclose(ow.gauss2gauss(m, g, g, order=order), m, atol=1e-12)


def test_wiener_alm_one_step_exact():
    '''Full sky, uniform wavelet noise: the harmonic preconditioner with
    itau = get_ivar_ell is the exact inverse of A, so CG converges in one
    step (the closed form of the reference's solver tests, re-expressed).'''
    lmax = 12
    w = kernels(lmax)
    cs = [4.0, 1.0, 9.0]
    maps, geoms = uniform_wav(lmax, w, cs, 1)
    ells = np.arange(lmax + 1)
    icov_ell = np.zeros((1, lmax + 1))
    icov_ell[0, 1:] = (ells[1:] + 1.0) ** 2 / 10
    a = rand_alm(lmax, 1, 4)
    sol = ow.WienerAlm.from_arrays_wav(a, lmax, icov_ell, maps, geoms, w, 0)
    itau = ow.get_ivar_ell(maps, geoms, w)
    sol.add_preconditioner(HarmonicPreconditioner(lmax, icov_ell, itau=itau))
    cg = sol.init_solver()
    cg.step()
    b0 = np.sqrt(ops.contract_almxblm(sol.b0, sol.b0))
    assert sol.get_residual() / b0 < 1e-12
    ell, _ = layout.lm_arrays(lmax)
    tot = np.sum(np.asarray(cs)[:, None] * w ** 2, axis=0)
    exp = tot[ell] * a[0] / (icov_ell[0][ell] + tot[ell])
    np.testing.assert_allclose(cg.x[0], exp, rtol=1e-10, atol=1e-13)


def test_pinv_wav_closed_form():
    '''M = H [S^-1 + tau^2 sum_j w_j^2 / c_j] H with H = (S^-1 + tau)^-1,
    tau = sum_j c_j w_j^2, on full sky with uniform wavelet noise.'''
    lmax = 12
    w = kernels(lmax)
    cs = [4.0, 1.0, 9.0]
    maps, geoms = uniform_wav(lmax, w, cs, 1)
    ells = np.arange(lmax + 1)
    icov_ell = np.zeros((1, lmax + 1))
    icov_ell[0, 2:] = (ells[2:] + 1.0) ** 2 / 10
    prec = ow.PseudoInvPreconditionerWav(lmax, icov_ell, maps, geoms, w, 0)
    a = rand_alm(lmax, 1, 5)
    tau = np.sum(np.asarray(cs)[:, None] * w ** 2, axis=0)
    inv = np.sum(w ** 2 / np.asarray(cs)[:, None], axis=0)
    h = np.zeros(lmax + 1)
    s = icov_ell[0] + tau
    h[s > 0] = 1 / s[s > 0]
    mexp = h * (icov_ell[0] + tau ** 2 * inv) * h
    ell, _ = layout.lm_arrays(lmax)
    np.testing.assert_allclose(prec(a), mexp[ell] * a, rtol=1e-9, atol=1e-13)
