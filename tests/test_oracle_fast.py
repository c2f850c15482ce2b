'''
The oracle's float64 OpenMP transform backend (oracle/fastsht.py over
oracle/legendre_omp.c) against the oracle's long-double numpy transforms
(oracle/sht.py, itself pinned by exact Wigner-d sums and brute force in
test_oracle.py): spin 0, spin 2 and TQU; Gauss-Legendre grids with even and
odd ring counts, phi0 != 0, rings too short for every m (aliasing), and
equal-area reduced rings; synthesis and adjoint to 1e-13. At lmax 255 (beyond
what the numpy path checks quickly) the backend's own invariants: the
dot-product adjoint identity and Y^T W Y = I on the GL grid.
'''
import numpy as np
import pytest

from oracle import geometry as og, sht as osht, layout, fastsht

pytestmark = pytest.mark.skipif(not fastsht.available(), reason='make -C oracle not run')


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


GEOMS = {
    'gl_even': lambda: (16, og.gauss_legendre(18, 33)),
    'gl_odd_phi0': lambda: (30, og.gauss_legendre(31, 63, phi0=0.3)),
    'gl_short_rings': lambda: (20, og.gauss_legendre(21, 31)),
    'equal_area': lambda: (40, og.equal_area(og.gauss_legendre(41, 81), orig_band=0.4,
                                             ratio_pow=1)),
    'gl_64': lambda: (64, og.gauss_legendre(65, 129)),
}


def rand_alm(lmax, npol, rng):
    a = rng.normal(size=(npol, layout.nelem(lmax))) + 1j * rng.normal(size=(npol,
                                                                          layout.nelem(lmax)))
    a[:, :lmax + 1] = a[:, :lmax + 1].real
    return a


@pytest.mark.parametrize('gname', sorted(GEOMS))
@pytest.mark.parametrize('spin,npol', [(0, 1), (2, 2), ([0, 2], 3)])
def test_fast_matches_numpy_oracle(gname, spin, npol):
    lmax, g = GEOMS[gname]()
    rng = np.random.default_rng(1)
    a = rand_alm(lmax, npol, rng)
    m_ref = np.zeros((npol, g.npix))
    m_got = np.zeros((npol, g.npix))
    osht.alm2map(a, m_ref, lmax, g, spin)
    fastsht.alm2map(a, m_got, lmax, g, spin)
    assert rel(m_got, m_ref) < 1e-13
    mp = rng.normal(size=(npol, g.npix))
    for adjoint in (False, True):
        a_ref = np.zeros_like(a)
        a_got = np.zeros_like(a)
        osht.map2alm(mp, a_ref, g, lmax, spin, adjoint=adjoint)
        fastsht.map2alm(mp, a_got, g, lmax, spin, adjoint=adjoint)
        assert rel(a_got, a_ref) < 1e-13


def test_fast_invariants_lmax255():
    lmax = 255
    g = og.gauss_legendre(lmax + 1, 2 * lmax + 1)
    rng = np.random.default_rng(2)
    a = rand_alm(lmax, 3, rng)
    ell, _ = layout.lm_arrays(lmax)
    a[1:, ell < 2] = 0          # E / B have no l < 2 modes
    m = np.zeros((3, g.npix))
    fastsht.alm2map(a, m, lmax, g, [0, 2])
    back = np.zeros_like(a)
    fastsht.map2alm(m, back, g, lmax, [0, 2])          # Y^T W Y = I
    assert rel(back, a) < 5e-12                         # float64 round-off: 1.5e-12
    f = rng.normal(size=(3, g.npix))
    yt = np.zeros_like(a)
    fastsht.map2alm(f, yt, g, lmax, [0, 2], adjoint=True)
    from oracle import ops
    lhs = np.sum(m * f)
    rhs = ops.contract_almxblm(a, yt)
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)


def test_install_routes_oracle_operators():
    lmax, g = GEOMS['gl_odd_phi0']()
    rng = np.random.default_rng(3)
    a = rand_alm(lmax, 1, rng)
    ref = np.zeros((1, g.npix))
    osht.alm2map(a, ref, lmax, g, 0)
    fastsht.install()
    try:
        assert osht.alm2map is fastsht.alm2map
        got = np.zeros((1, g.npix))
        osht.alm2map(a, got, lmax, g, 0)
    finally:
        fastsht.uninstall()
    assert osht.alm2map is not fastsht.alm2map
    assert rel(got, ref) < 1e-13
