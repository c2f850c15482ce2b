'''
CAR (Clenshaw-Curtis / Fejer-1) geometry of the two-phase filter, on the
CPU: the quadrature weights (textbook values and polynomial exactness, a
planted mistake shown to fail), the reference's round trips on full-sky CAR
grids (test_map_utils.py:1258-1313), the CAR pixel-to-ring mapping against a
brute-force synthesis at every pixel's (dec, ra) read straight from the WCS
(all four cdelt sign variants, a cut-sky band, both grids), threshold_icov
against test_map_utils.py:636-678, the product's host geometry against the
oracle's, and the oracle's two-phase filter against a dense solve of A x = b.
'''
import numpy as np
import pytest

from oracle import geometry as og, sht as osht, layout, ops, wiener as ow, legendre as L
from optweight_amd import map_utils


def car_wcs(ny, nx, mtype, flip_x=False, flip_y=False, dec_band=None):
    '''WCS parameters of pixell's full-sky CAR convention (RA decreasing
    with x, dec increasing with y, pixel 0 at dec -90 for CC and at
    -90 + half a pixel for fejer1); dec_band = (row0, row1) keeps rows.'''
    dy = 180 / (ny - 1) if mtype == 'CC' else 180 / ny
    cdelt = [-360 / nx, dy]
    crpix = [nx / 2 + 1, (ny + 1) / 2]
    if flip_x:
        cdelt[0] *= -1
    if flip_y:
        cdelt[1] *= -1
    if dec_band is not None:
        crpix[1] -= dec_band[0] if not flip_y else 0
    return map_utils.CarWCS(cdelt, crpix, [0., 0.])


def rand_alm(lmax, npol, seed):
    rng = np.random.default_rng(seed)
    a = rng.normal(size=(npol, layout.nelem(lmax))) + 1j * rng.normal(size=(npol, layout.nelem(lmax)))
    ell, em = layout.lm_arrays(lmax)
    a[:, em == 0] = a[:, em == 0].real
    if npol == 3:
        a[1:, ell < 2] = 0
    return a


# ------------------------------------------------------------ weights

@pytest.mark.parametrize('kind', ['CC', 'F1'])
def test_gridweights_exact_polynomials(kind):
    '''CC with n + 1 rings integrates cos(theta)^k exactly for k <= n,
    Fejer-1 with n rings for k <= n - 1 (over the sphere: 2 pi * 2/(k+1) for
    even k, 0 for odd k); the weights sum to 4 pi.'''
    for nt in (3, 4, 7, 12, 37):
        w = og.gridweights(kind, nt)
        th = (np.linspace(0, np.pi, nt) if kind == 'CC'
              else (np.arange(nt) + 0.5) * np.pi / nt)
        deg = nt - 1 if kind == 'CC' else nt - 1
        for k in range(deg + 1):
            want = 0.0 if k % 2 else 2 * np.pi * 2 / (k + 1)
            assert abs(np.sum(w * np.cos(th) ** k) - want) < 1e-13
        np.testing.assert_allclose(w, map_utils.get_gridweights(kind, nt), rtol=1e-15)


def test_gridweights_textbook_and_planted_mistake():
    '''Simpson's rule for 3 CC rings, equal weights for 2 Fejer-1 rings; a CC
    rule with b_j = 2 at j = n/2 loses exactness at degree n.'''
    np.testing.assert_allclose(og.gridweights('CC', 3) / (2 * np.pi), [1 / 3, 4 / 3, 1 / 3],
                               rtol=1e-15)
    np.testing.assert_allclose(og.gridweights('F1', 2) / (2 * np.pi), [1, 1], rtol=1e-15)
    n = 6
    k = np.arange(n + 1)
    w = np.ones(n + 1)
    for j in range(1, n // 2 + 1):
        w -= 2.0 / (4.0 * j * j - 1) * np.cos(2 * j * k * np.pi / n)
    c = np.full(n + 1, 2.0)
    c[0] = c[-1] = 1.0
    bad = 2 * np.pi * c / n * w
    th = np.linspace(0, np.pi, n + 1)
    assert abs(np.sum(bad * np.cos(th) ** n) - 2 * np.pi * 2 / (n + 1)) > 1e-3


# ------------------------------------------------------ round trips

@pytest.mark.parametrize('mtype', ['CC', 'fejer1'])
@pytest.mark.parametrize('spin,npol', [(0, 1), ([0, 2], 3)])
def test_car_roundtrip(mtype, spin, npol):
    '''test_map_utils.py:1258-1313: on a full-sky CAR grid of 2 lmax + 1
    rows, map2alm(alm2map(a)) = a.'''
    lmax = 12
    ny, nx = 2 * lmax + 1, 2 * lmax + 2
    _, g, index = og.match_car((ny, nx), car_wcs(ny, nx, mtype), mtype=mtype)
    a = rand_alm(lmax, npol, 1)
    m = np.zeros((npol, g.npix))
    osht.alm2map(a, m, lmax, g, spin)
    b = np.zeros_like(a)
    osht.map2alm(m, b, g, lmax, spin)
    assert np.linalg.norm(b - a) / np.linalg.norm(a) < 1e-13


def _brute_car(a, lmax, wcs, ny, nx):
    '''sum_lm a_lm sY_lm at every pixel (y, x) of the CAR map, (dec, ra) from
    the WCS; T, Q, U for three components.'''
    dec, _ = og.car_pix2sky(wcs, np.arange(ny), np.zeros(ny))
    _, ra = og.car_pix2sky(wcs, np.zeros(nx), np.arange(nx))
    theta = np.pi / 2 - dec
    ell, em = layout.lm_arrays(lmax)
    out = np.zeros((a.shape[0], ny, nx))
    for i, (l, m) in enumerate(zip(ell, em)):
        e = np.exp(1j * m * ra)[None, :]
        fac = 1 if m == 0 else 2
        l0 = np.array([float(L.lam_exact(l, m, 0, t)) for t in theta])[:, None]
        out[0] += fac * (a[0, i] * l0 * e).real
        if a.shape[0] == 3 and l >= 2:
            lp = np.array([float(L.lam_exact(l, m, 2, t)) for t in theta])[:, None]
            ln = np.array([float(L.lam_exact(l, m, -2, t)) for t in theta])[:, None]
            F1, F2 = (lp + ln) / 2, (lp - ln) / 2
            out[1] += fac * (-(a[1, i] * F1 + 1j * a[2, i] * F2) * e).real
            out[2] += fac * (-(a[2, i] * F1 - 1j * a[1, i] * F2) * e).real
    return out


@pytest.mark.parametrize('mtype,flip_x,flip_y,band', [
    ('CC', False, False, None), ('CC', True, False, None), ('CC', False, True, None),
    ('CC', True, True, None), ('fejer1', False, False, None), ('CC', False, False, (3, 9)),
    ('fejer1', True, False, (2, 7))])
def test_car_mapping_brute_force(mtype, flip_x, flip_y, band):
    '''The oracle's CAR geometry (rings in row order, phi0, stride, the
    gather index) synthesises the same map as a direct sum at each pixel's
    (dec, ra); the product's match_enmap_minfo equals the oracle's.'''
    lmax = 4
    ny, nx = 2 * lmax + 1, 2 * lmax + 2
    wcs = car_wcs(ny, nx, mtype, flip_x=flip_x, flip_y=flip_y)
    shape = (ny, nx)
    if band is not None:
        wcs.wcs.crpix[1] -= band[0]
        shape = (band[1] - band[0], nx)
    ref_par, g, index = og.match_car(shape, wcs, mtype=mtype)
    a = rand_alm(lmax, 3, 2)
    m = np.zeros((3, g.npix))
    osht.alm2map(a, m, lmax, g, [0, 2])
    flat = np.zeros((3, shape[0] * nx))
    flat[:, index] = m
    want = _brute_car(a, lmax, wcs, shape[0], nx)
    assert np.abs(flat.reshape(want.shape) - want).max() < 1e-12
    mi = map_utils.match_enmap_minfo(shape, wcs, mtype=mtype)
    th, wt, nphi, phi0, offsets, stride = ref_par
    np.testing.assert_allclose(mi.theta, th, rtol=0, atol=1e-15)
    np.testing.assert_allclose(mi.weight, wt, rtol=1e-15)
    np.testing.assert_allclose(mi.phi0, phi0, rtol=0, atol=1e-15)
    np.testing.assert_array_equal(mi.offsets, offsets)
    np.testing.assert_array_equal(mi.stride, stride)
    _, idx = map_utils.unit_stride(mi)
    np.testing.assert_array_equal(idx, index)


def test_match_enmap_errors():
    wcs = car_wcs(9, 10, 'CC')
    with pytest.raises(ValueError):
        map_utils.match_enmap_minfo((9, 12), wcs)
    with pytest.raises(ValueError):
        map_utils.match_enmap_minfo((9, 10), wcs, mtype='fejer1')
    with pytest.raises(NotImplementedError):
        map_utils.match_enmap_minfo((9, 10), wcs, mtype='GL')


# ------------------------------------------------------- threshold

@pytest.mark.parametrize('fn', [ops.threshold_icov, map_utils.threshold_icov])
def test_threshold_icov_golden(fn):
    '''test_map_utils.py:636-678 (literal answers).'''
    icov = np.ones((3, 3, 4)) * np.eye(3)[:, :, None]
    out = fn(icov)
    assert not np.shares_memory(out, icov)
    np.testing.assert_allclose(out, icov)
    icov[0, 0, 0] = 0.1
    icov[0, 1, 0] = icov[1, 0, 0] = 0.01
    v = np.quantile([0.1, 1, 1, 1], 0.2)
    exp = icov.copy()
    exp[0, 0, 0] = v
    exp[0, 1, 0] = exp[1, 0, 0] = 0.01 * np.sqrt(v / 0.1)
    np.testing.assert_allclose(fn(icov, q_low=[0.2]), exp)
    icov = np.ones((3, 3, 4)) * np.eye(3)[:, :, None]
    icov[0, 0, 0] = 100
    icov[0, 1, 0] = icov[1, 0, 0] = 10
    v = np.quantile([100, 1, 1, 1], 0.8)
    exp = icov.copy()
    exp[0, 0, 0] = v
    exp[0, 1, 0] = exp[1, 0, 0] = 10 * np.sqrt(v / 100)
    np.testing.assert_allclose(fn(icov, q_high=[0.8]), exp)
    d = np.asarray([[0.0, 1, 2, 3, 100]])
    np.testing.assert_allclose(fn(d, q_high=0.5), [[0, 1, 2, 2.5, 2.5]])


# ------------------------------------------------------------ filter

def filter_problem(ncomp, mtype, seed=0):
    '''A small CAR problem: full-sky map of ny = 17 rows, a polar cap
    unobserved, noise depth varying with dec, theory spectra and a beam.'''
    lmax = 8
    ny, nx = 2 * lmax + 1, 2 * lmax + 2
    wcs = car_wcs(ny, nx, mtype)
    dec, _ = og.car_pix2sky(wcs, np.arange(ny), np.zeros(ny))
    rng = np.random.default_rng(seed)
    ell = np.arange(lmax + 1)
    cl = 1e2 / (ell + 5.0) ** 2
    cls = {'TT': cl, 'EE': 0.3 * cl, 'BB': 0.05 * cl, 'TE': 0.2 * cl}
    b_ell = np.exp(-0.5 * ell * (ell + 1) * 0.05 ** 2)
    depth = (1.0 + 0.5 * np.cos(dec))[:, None] * np.ones((ny, nx))
    icov = np.stack([depth, 0.5 * depth, 0.5 * depth])[:ncomp]
    mask = np.ones((ny, nx), dtype=bool)
    mask[np.degrees(dec) > 60] = False
    icov = icov * mask
    imap = rng.normal(size=(ncomp, ny, nx))
    return cls, b_ell, icov, mask, imap, wcs, lmax


def _dense_solution(sol, npol, lmax):
    '''A^-1 b of an oracle solver by a dense solve in real coordinates (real
    and imaginary parts of the m > 0 coefficients, real part at m = 0).'''
    n = layout.nelem(lmax)
    _, em = layout.lm_arrays(lmax)
    slots = [(c, i, p) for c in range(npol) for i in range(n)
             for p in ((1.0,) if em[i] == 0 else (1.0, 1j))]

    def to_real(v):
        return np.array([(v[c, i] * np.conj(p)).real for c, i, p in slots])

    cols = []
    for c, i, p in slots:
        e = np.zeros((npol, n), complex)
        e[c, i] = p
        cols.append(to_real(sol.a_matrix(e)))
    A = np.stack(cols, axis=1)
    assert np.abs(A - A.T).max() < 1e-12 * np.abs(A).max()
    return np.linalg.solve(A, to_real(sol.b0)), to_real


def test_oracle_filter_converges_to_dense_solution():
    '''The oracle filter's phase switch keeps the system (x carried over, b
    reset): after enough phase-2 steps its Wiener alm equal the dense
    solution of A x = b on the same arrays.'''
    cls, b_ell, icov, mask, imap, wcs, lmax = filter_problem(3, 'CC')
    f = ow.CGPixFilter(cls, b_ell, icov, mask, wcs, no_masked_prec_mg=True)
    out = f.filter(imap, niter=150, niter_masked_cg=2, err_tol=1e-30)
    sol = ow.WienerMap.from_arrays(imap.reshape(3, -1)[..., f.index], f.geom, lmax, f.icov_ell,
                                   f.icov_pix, b_ell=f.b_ell, mask_pix=f.mask_bool,
                                   spin=[0, 2])
    x, to_real = _dense_solution(sol, 3, lmax)
    assert np.linalg.norm(to_real(out['walm']) - x) / np.linalg.norm(x) < 1e-12
    assert np.isnan(out['errors'][0]) and 3 < len(out['errors']) <= 153
