'''
Noise-model estimation (oracle; test infrastructure only), restating
optweight's noise_utils.estimate_cov_wav (noise_utils.py:10-77),
estimate_cov_pix (144-226), norm_cov_est (228-253), the spin-0 smoothing of
map_utils.lmul_pix (map_utils.py:1365-1408), the wavelet maps of
wavtrans.alm2wav (Gauss-Legendre grid of 2 lmax_j + 1 pixels per ring for
scale j) and mat_utils.get_near_psd (matrix power 1 through the eigenvalue
routine, negative eigenvalues dropped).
'''
import numpy as np

from oracle import ops, sht as osht, geometry, wavelet, layout


def band_limit_gauss_beam(lmax, fwhm=None):
    if fwhm is None:
        fwhm = 2 * np.pi / lmax
    return ops.gauss_beam(fwhm, lmax)


def lmul_pix(imap, lmat, geom):
    '''Y lmat Y^T W of spin-0 maps (k, npix).'''
    lmax = np.asarray(lmat).shape[-1] - 1
    imap = np.atleast_2d(imap)
    alm = np.zeros((imap.shape[0], layout.nelem(lmax)), dtype=np.complex128)
    osht.map2alm(imap, alm, geom, lmax, 0)
    alm = ops.lmul(alm, lmat, lmax)
    out = np.zeros_like(imap, dtype=np.float64)
    osht.alm2map(alm, out, lmax, geom, 0)
    return out


def norm_cov_est(cov, geom, kernel_ell):
    kernel_ell = np.asarray(kernel_ell, dtype=np.float64)
    ells = np.arange(kernel_ell.size)
    norm = np.sum(kernel_ell ** 2 * (2 * ells + 1)) / 4 / np.pi
    return cov / np.repeat(geom.weight, geom.nphi) / norm


def estimate_cov_pix(imap, geom, diag=False, fwhm=None, lmax=None, kernel_ell=None):
    imap = np.asarray(imap, dtype=np.float64)
    while imap.ndim < 3:
        imap = imap[None]
    ncomp, npol, npix = imap.shape
    if diag:
        cov = np.einsum('abc, dbc -> adbc', imap, imap)
    else:
        cov = np.einsum('abc, dec -> abdec', imap, imap).reshape(ncomp * npol, ncomp * npol,
                                                                 npix)
    lmax_g = geometry.minfo2lmax(geom)
    if fwhm is None:
        fwhm = 2 * np.pi / lmax_g
    b_ell = band_limit_gauss_beam(lmax_g, fwhm=fwhm)
    for idx in np.ndindex(cov.shape[:-1]):
        if idx[0] <= idx[1]:
            if fwhm > 0:
                cov[idx] = lmul_pix(cov[idx], b_ell, geom)[0]
        else:
            cov[idx] = cov[(idx[1], idx[0]) + idx[2:]]
    if lmax or kernel_ell is not None:
        if kernel_ell is None:
            kernel_ell = np.ones(lmax + 1)
        cov = norm_cov_est(cov, geom, kernel_ell)
    if not diag:
        cov = ops.eigpow(cov, 1, lim=1e-9).reshape(ncomp, npol, ncomp, npol, npix)
    else:
        cov[cov < 0] = 0
    return cov


def estimate_cov_wav(alm, lmax, w_ell, spin, diag=False, fwhm_fact=2):
    '''[(geom_j, cov_pix_j)] per wavelet scale j from noise alm
    (ncomp, npol, nelem); the monopole is left out.'''
    w_ell = np.atleast_2d(w_ell)
    a = np.array(alm, copy=True)
    a[..., 0, 0] = 0
    a3 = a.reshape((-1,) + a.shape[-2:])
    out = []
    for j in range(w_ell.shape[0]):
        lj = wavelet.wav_lmax(w_ell, j)
        g = geometry.gauss_legendre(lj + 1, 2 * lj + 1)
        maps = np.zeros(a3.shape[:-1] + (g.npix,))
        for c in range(a3.shape[0]):
            wlm = ops.lmul(ops.trunc_alm(a3[c], lmax, lj), w_ell[j, :lj + 1], lj)
            m = np.zeros((a3.shape[1], g.npix))
            osht.alm2map(wlm, m, lj, g, spin)
            maps[c] = m
        lmax_g = geometry.minfo2lmax(g)
        fact = fwhm_fact(lmax_g) if callable(fwhm_fact) else fwhm_fact
        maps = maps.reshape(a.shape[:-1] + (g.npix,))
        out.append((g, estimate_cov_pix(maps, g, diag=diag, fwhm=fact * np.pi / lmax_g,
                                        kernel_ell=w_ell[j])))
    return out
