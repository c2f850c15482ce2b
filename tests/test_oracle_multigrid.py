'''
Pins of the oracle's multigrid (oracle/wiener.py Level, get_levels,
restrict, v_cycle: the restatement of optweight's multigrid.py:13-447) that
do not go through the oracle's own formulas:

  * addition theorem: for spin 0, G = M_u Y d_l Y^T M_u (Y synthesis, Y^T
    its exact adjoint) has the entries
        G_pq = sum_l d_l (2l + 1) / (4 pi) P_l(cos gamma_pq)
    for any pixel set (gamma the angle between pixels p and q), so the
    dense G of a level, its diagonal probe and the restriction operators
    are brute-forced from scipy's Legendre polynomials on the level's own
    pixels (the equal-area reduced rings included);
  * the V-cycle restated with those dense matrices (Jacobi smoother
    0.2 / G_pp, coarsen C = M_c K_r W_f M_f, interpolation P = C^T, dense
    pseudo-inverse on the coarsest level) equals v_cycle;
  * coarsen and interpolate are exact transposes under the plain pixel dot
    product (multigrid.py:296-351);
  * one level: v_cycle is the pseudo-inverse of G;
  * the reference tests' properties: test_multigrid.py:105-204 (coarsen /
    interpolate against a full-size transform chain), :206-250 (the V-cycle
    approximately inverts G: |median| < 0.1, MAD < 0.3 with n_jacobi = 3),
    on seeded alm of this repository instead of curvedsky.rand_alm;
  * planted mistakes fail these checks: damping 0.5 instead of 0.2, the
    adjoint flags of restrict swapped, the coarsening beam taken at the fine
    band limit.
'''
import numpy as np
import pytest
from scipy import linalg
from scipy.special import eval_legendre

from oracle import geometry as og, wiener as ow, ops, sht as osht, layout


def pixel_dirs(geom):
    theta = np.repeat(geom.theta, geom.nphi)
    phi = np.concatenate([geom.phi0[r] + 2 * np.pi * np.arange(geom.nphi[r]) / geom.nphi[r]
                          for r in range(len(geom.theta))])
    return np.stack([np.sin(theta) * np.cos(phi), np.sin(theta) * np.sin(phi), np.cos(theta)],
                    axis=-1)


def pixel_weights(geom):
    return np.repeat(geom.weight, geom.nphi)


def kernel(c_ell, dirs_a, dirs_b):
    '''K_pq = sum_l c_l (2l + 1) / (4 pi) P_l(n_p . n_q).'''
    cosg = np.clip(dirs_a @ dirs_b.T, -1.0, 1.0)
    out = np.zeros(cosg.shape)
    for l, c in enumerate(c_ell):
        if c != 0:
            out += c * (2 * l + 1) / (4 * np.pi) * eval_legendre(l, cosg)
    return out


def t_setup(lmax=24, band_deg=15.):
    g = og.gauss_legendre(lmax + 1, 2 * lmax + 1)
    lat = np.pi / 2 - np.repeat(g.theta, g.nphi)
    mask = np.abs(lat) >= np.radians(band_deg)
    ell = np.arange(lmax + 1)
    cl = np.zeros(lmax + 1)
    cl[2:] = 1e3 * (ell[2:] + 10.) ** -2.5
    icov = np.zeros((1, lmax + 1))
    icov[0, 2:] = 1 / cl[2:]
    return g, mask, icov


@pytest.fixture(scope='module')
def two_levels():
    g, mask, icov = t_setup()
    levels = ow.get_levels(mask, g, icov, 0, min_pix=200, lmax_r_ell=24)
    assert len(levels) == 2 and levels[1].pinv_g is not None
    return levels


def dense_g(level):
    '''Brute-force G of a spin-0 level on its unobserved pixels.'''
    u = level.mask_unobs[0]
    d = pixel_dirs(level.geom)[u]
    return kernel(level.d_ell[0, 0], d, d)


def test_dense_g_addition_theorem(two_levels):
    for lev in two_levels:
        u = lev.mask_unobs[0]
        got = lev.dense_g() if lev.pinv_g is not None else None
        ref = dense_g(lev)
        if got is None:
            # the sparse level: G applied to a few unit vectors
            idx = np.nonzero(u)[0][:: max(1, u.sum() // 7)]
            for p in idx:
                e = np.zeros((1, lev.geom.npix))
                e[0, p] = 1
                col = lev.g_op(e)[0]
                np.testing.assert_allclose(col[u], ref[:, np.searchsorted(np.nonzero(u)[0], p)],
                                           rtol=0, atol=1e-12 * np.abs(ref).max())
                assert np.all(col[~u] == 0)
        else:
            assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max()


def probe_closed_form(d_ell, omega=0.2):
    '''omega / G_pp with G_pp = sum_l d_l (2l + 1) / (4 pi); the reciprocal
    is taken in float32 (the reference probes with a float32 transform,
    multigrid.py:82-113) and scaled by omega in float64.'''
    gpp = np.sum(d_ell * (2 * np.arange(d_ell.size) + 1) / (4 * np.pi))
    return omega * float(np.float32(1) / np.float32(gpp))


def test_probe_diag_closed_form(two_levels):
    lev = two_levels[0]
    got = lev.m_prec[0]
    assert got == pytest.approx(probe_closed_form(lev.d_ell[0, 0]), rel=1e-14)
    # planted mistake: a damping of 0.5 is caught
    assert abs(probe_closed_form(lev.d_ell[0, 0], omega=0.5) - got) > 0.5 * got


def restriction_matrices(lin, lout):
    '''Dense coarsen (lin -> lout, adjoint False) and interpolate (lout ->
    lin, adjoint True) for spin 0, from the addition theorem: coarsen =
    M_out K_r W_in M_in with r_l the Gaussian of FWHM 2 pi / lmax_out up to
    lmax_out; interpolate = M_in W_in K_r M_out with the same r_l (the coarse
    level's band limit in both directions).'''
    lc = og.minfo2lmax(lout.geom)
    r = ops.gauss_beam(2 * np.pi / lc, lc)
    k = kernel(r, pixel_dirs(lout.geom), pixel_dirs(lin.geom))
    mi = lin.mask_unobs[0].astype(float)
    mo = lout.mask_unobs[0].astype(float)
    coarsen = mo[:, None] * k * (pixel_weights(lin.geom) * mi)[None, :]
    interp = (mi * pixel_weights(lin.geom))[:, None] * k.T * mo[None, :]
    return coarsen, interp


def test_restrict_matches_brute_force(two_levels):
    fine, coarse = two_levels
    cmat, imat = restriction_matrices(fine, coarse)
    rng = np.random.default_rng(3)
    x = rng.normal(size=(1, fine.geom.npix))
    y = rng.normal(size=(1, coarse.geom.npix))
    got_c = ow.restrict(x, fine, coarse, 0, adjoint=False)[0]
    got_i = ow.restrict(y, coarse, fine, 0, adjoint=True)[0]
    assert np.linalg.norm(got_c - cmat @ x[0]) <= 1e-12 * np.linalg.norm(got_c)
    assert np.linalg.norm(got_i - imat @ y[0]) <= 1e-12 * np.linalg.norm(got_i)
    # planted mistake: both adjoint flags swapped (still a transposed pair,
    # with the weights on the other grid) misses the brute force
    bad_c = ow.restrict(x, fine, coarse, 0, adjoint=True)[0]
    assert np.linalg.norm(bad_c - cmat @ x[0]) > 1e-2 * np.linalg.norm(got_c)


def adjoint_gap(restrict, fine, coarse, spin, npol, seed=5):
    rng = np.random.default_rng(seed)
    x = rng.normal(size=(npol, fine.geom.npix))
    y = rng.normal(size=(npol, coarse.geom.npix))
    cx = restrict(x, fine, coarse, spin, adjoint=False)
    iy = restrict(y, coarse, fine, spin, adjoint=True)
    a, b = np.sum(cx * y), np.sum(x * iy)
    return abs(a - b) / abs(a)


@pytest.fixture(scope='module')
def tqu_levels():
    lmax = 22
    g = og.gauss_legendre(lmax + 1, 2 * lmax + 1)
    ells = np.arange(lmax + 1)
    icov = np.ones((3, 3, lmax + 1)) * np.asarray(
        [[1, 0.1, 0.1], [0.1, 1, 0.1], [0.1, 0.1, 1]])[:, :, None]
    icov *= 1e-4 * (ells + 1) ** 2
    icov[:, :, :2] = 0
    mask = np.ones(g.npix, dtype=bool)
    m2 = og.view_2d(mask, g)
    m2[8:12, :] = False
    levels = ow.get_levels(mask, g, icov, [0, 2], min_pix=10, lmax_r_ell=lmax)
    return g, icov, levels


def test_coarsen_interpolate_are_transposes(two_levels, tqu_levels):
    fine, coarse = two_levels
    assert adjoint_gap(ow.restrict, fine, coarse, 0, 1) < 1e-13
    _, _, levels = tqu_levels
    assert adjoint_gap(ow.restrict, levels[0], levels[1], [0, 2], 3) < 1e-13

    # planted mistakes: the coarsening with the interpolation's adjoint flag
    # (weights on the wrong side), the coarsening beam at the fine lmax
    def swapped(imap, lin, lout, spin, adjoint=False):
        return ow.restrict(imap, lin, lout, spin, adjoint=True)

    assert adjoint_gap(swapped, fine, coarse, 0, 1) > 1e-3

    def wrong_beam(imap, lin, lout, spin, adjoint=False):
        lmax_t = min(lin.lmax, lout.lmax)
        imap = imap * lin.mask_unobs
        alm = np.zeros((imap.shape[0], layout.nelem(lmax_t)), dtype=np.complex128)
        osht.map2alm(imap, alm, lin.geom, lmax_t, spin, adjoint=adjoint)
        r = ops.gauss_beam(2 * np.pi / lin.lmax, lmax_t)        # mistake: input lmax
        alm = ops.lmul(alm, r, lmax_t)
        omap = np.zeros((imap.shape[0], lout.geom.npix))
        osht.alm2map(alm, omap, lmax_t, lout.geom, spin, adjoint=adjoint)
        return omap * lout.mask_unobs

    assert adjoint_gap(wrong_beam, fine, coarse, 0, 1) > 1e-3


def dense_v_cycle(levels, r, n_jacobi, omega=0.2):
    '''The V-cycle of two spin-0 levels restated with brute-force dense
    matrices on each level's full pixel set.'''
    fine, coarse = levels
    u0 = fine.mask_unobs[0]
    d0 = pixel_dirs(fine.geom)
    g0 = np.zeros((fine.geom.npix,) * 2)
    g0[np.ix_(u0, u0)] = kernel(fine.d_ell[0, 0], d0[u0], d0[u0])
    s0 = probe_closed_form(fine.d_ell[0, 0], omega)
    u1 = coarse.mask_unobs[0]
    pinv1 = linalg.pinvh(dense_g(coarse))
    cmat, imat = restriction_matrices(fine, coarse)
    x = s0 * r
    for _ in range(n_jacobi - 1):
        x = x + s0 * (r - g0 @ x)
    rh = cmat @ (r - g0 @ x)
    ch = np.zeros(coarse.geom.npix)
    ch[u1] = pinv1 @ rh[u1]
    x = x + imat @ ch
    for _ in range(n_jacobi):
        x = x + s0 * (r - g0 @ x)
    return x


@pytest.mark.parametrize('n_jacobi', [1, 2])
def test_v_cycle_matches_dense_restatement(two_levels, n_jacobi):
    levels = two_levels
    rng = np.random.default_rng(7)
    r = rng.normal(size=levels[0].geom.npix) * levels[0].mask_unobs[0]
    got = ow.v_cycle(levels, r[None], 0, n_jacobi=n_jacobi)[0]
    ref = dense_v_cycle(levels, r, n_jacobi)
    assert np.linalg.norm(got - ref) <= 1e-9 * np.linalg.norm(ref)
    # planted mistake: damping 0.5 changes the result well beyond that
    bad = dense_v_cycle(levels, r, n_jacobi, omega=0.5)
    assert np.linalg.norm(got - bad) > 1e-2 * np.linalg.norm(ref)


def test_single_level_is_pinv(two_levels):
    '''One dense level (the coarse level's operator, whose spectrum has a
    clean gap between its null space and 1.7e-4, so the pseudo-inverse is
    well defined): v_cycle is pinvh of the brute-force G.'''
    coarse = two_levels[1]
    lev = ow.Level(~coarse.mask_unobs, coarse.geom, coarse.d_ell, 0, dense=True)
    levels = [lev]
    u = lev.mask_unobs[0]
    rng = np.random.default_rng(8)
    r = rng.normal(size=lev.geom.npix) * u
    got = ow.v_cycle(levels, r[None], 0)[0]
    ref = np.zeros_like(r)
    ref[u] = linalg.pinvh(dense_g(lev)) @ r[u]
    assert np.linalg.norm(got - ref) <= 1e-9 * np.linalg.norm(ref)


def test_coarsen_full_size_chain(tqu_levels):
    '''test_multigrid.py:105-153: coarsen equals the chain with the alm kept
    at the fine band limit and the beam zeroed above the coarse one.'''
    g, icov, levels = tqu_levels
    fine, coarse = levels[0], levels[1]
    lf, lc = fine.lmax, coarse.lmax
    cov = ops.matpow(icov, -1)
    alm = ops.rand_alm(cov, lf, 11)
    imap = np.zeros((3, fine.geom.npix))
    osht.alm2map(alm, imap, lf, fine.geom, [0, 2])
    got = ow.restrict(imap, fine, coarse, [0, 2], adjoint=False)
    tmp = np.zeros((3, layout.nelem(lf)), dtype=np.complex128)
    osht.map2alm(imap * fine.mask_unobs, tmp, fine.geom, lf, [0, 2])
    r = ops.gauss_beam(2 * np.pi / lc, lf)
    r[lc + 1:] = 0
    tmp = ops.lmul(tmp, r, lf)
    exp = np.zeros((3, coarse.geom.npix))
    osht.alm2map(tmp, exp, lf, coarse.geom, [0, 2])
    exp *= coarse.mask_unobs
    np.testing.assert_allclose(got, exp, rtol=0, atol=1e-12 * np.abs(exp).max())


def test_v_cycle_approximates_inverse(tqu_levels):
    '''test_multigrid.py:206-250: G^-1 recovered on large scales.'''
    g, icov, levels = tqu_levels
    lmax = levels[0].lmax
    cov = ops.matpow(icov, -1)
    cov[:, :, 5:] = 0
    alm = ops.rand_alm(cov, lmax, 12)
    imap = np.zeros((3, levels[0].geom.npix))
    osht.alm2map(alm, imap, lmax, levels[0].geom, [0, 2])
    omap = ow.v_cycle(levels, levels[0].g_op(imap), [0, 2], n_jacobi=3)
    u = levels[0].mask_unobs
    diff = (imap[u] - omap[u]) / np.abs(imap[u])
    assert abs(np.median(diff)) < 0.1
    assert np.median(np.abs(diff - np.median(diff))) < 0.3
