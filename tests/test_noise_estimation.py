'''
Pins of the oracle's noise-model estimation (CPU): the normalisation of
band-limited covariance estimates by a closed form (a constant map keeps its
value through the spin-0 smoothing, then gets divided by the pixel area and
by sum_l (2l + 1) / 4 pi; a planted mistake in that sum fails), the
statistics of unit-spectrum noise (per-pixel variance times pixel area
averages to 1, independent components average to zero cross covariance),
symmetry and positive semi-definiteness of the full estimate, and the same
statistics per wavelet scale for estimate_cov_wav.
'''
import numpy as np

from oracle import geometry as og, sht as osht, layout, noise as on


def unit_alm(lmax, shape, seed):
    rng = np.random.default_rng(seed)
    a = (rng.normal(size=shape + (layout.nelem(lmax),))
         + 1j * rng.normal(size=shape + (layout.nelem(lmax),))) / np.sqrt(2)
    a[..., :lmax + 1] = a[..., :lmax + 1].real * np.sqrt(2)
    return a


def test_norm_cov_est_constant_map():
    lmax = 16
    g = og.gauss_legendre(lmax + 1, 2 * lmax + 1)
    c = 1.7
    m = np.full((1, 1, g.npix), c)
    cov = on.estimate_cov_pix(m, g, diag=True, lmax=lmax)
    s = (lmax + 1) ** 2 / (4 * np.pi)
    want = c ** 2 / np.repeat(g.weight, g.nphi) / s
    np.testing.assert_allclose(cov[0, 0, 0], want, rtol=1e-12)
    bad = c ** 2 / np.repeat(g.weight, g.nphi) / (s * 4 * np.pi)
    assert not np.allclose(cov[0, 0, 0], bad, rtol=1e-3)


def test_estimate_cov_pix_statistics():
    '''8 components of unit-spectrum noise, band limit 32: the normalised
    variance times the pixel area averages to 1 (sky average of 8 maps,
    1.5 % standard deviation), cross terms of the full estimate to 0; the
    full estimate is symmetric and positive semi-definite at every pixel.'''
    lmax = 32
    g = og.gauss_legendre(lmax + 1, 2 * lmax + 1)
    a = unit_alm(lmax, (8,), 3)
    m = np.zeros((8, g.npix))
    osht.alm2map(a, m, lmax, g, 0)
    cov = on.estimate_cov_pix(m[:, None, :], g, diag=True, lmax=lmax)
    w = np.repeat(g.weight, g.nphi)
    d = np.array([np.mean(cov[i, i, 0] * w) for i in range(8)])
    assert abs(np.mean(d) - 1) < 0.05
    # cross terms from the full estimate (with diag the reference clips every
    # negative entry, cross terms included)
    full = on.estimate_cov_pix(m[:3].reshape(3, 1, -1), g, diag=False, lmax=lmax)
    f = full.reshape(3, 3, -1)
    off = np.mean([np.mean(f[i, j] * w) for i in range(3) for j in range(3) if i != j])
    assert abs(off) < 0.1
    np.testing.assert_allclose(f, f.transpose(1, 0, 2), rtol=0, atol=1e-12 * np.abs(f).max())
    ev = np.linalg.eigvalsh(np.moveaxis(f, -1, 0))
    assert ev.min() >= -1e-10 * ev.max()


def test_estimate_cov_wav_statistics():
    '''Per wavelet scale j of unit-spectrum noise: the normalised variance
    times the pixel area averages to 1; the monopole is left out and the
    input alm are unchanged.'''
    lmax = 48
    ell = np.arange(lmax + 1)
    # three cosine-like bands
    w_ell = np.zeros((3, lmax + 1))
    edges = [0, 12, 24, 48]
    for j in range(3):
        lo, hi = edges[j], edges[j + 1]
        w_ell[j, lo:hi + 1] = np.sin(np.pi * (ell[lo:hi + 1] - lo) / (hi - lo)) + (j == 0) * 0.5
    w_ell = np.clip(w_ell, 0, None)
    a = unit_alm(lmax, (8, 1), 5)
    a0 = a.copy()
    out = on.estimate_cov_wav(a, lmax, w_ell, 0, diag=True)
    assert np.array_equal(a, a0)
    for j, (g, cov) in enumerate(out):
        w = np.repeat(g.weight, g.nphi)
        d = np.mean([np.mean(cov[i, i, 0] * w) for i in range(8)])
        assert abs(d - 1) < 0.1, (j, d)
