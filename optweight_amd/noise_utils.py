'''
Noise helpers the Wiener path needs, mirroring optweight's noise_utils:
_band_limit_gauss_beam (noise_utils.py:401-421, healpy.gauss_beam closed
form exp(-l(l+1) sigma^2 / 2), sigma = fwhm / sqrt(8 ln 2)) and
get_white_icov_pix (592-641).
'''
import numpy as np


def gauss_beam(fwhm, lmax):
    sigma = fwhm / np.sqrt(8 * np.log(2))
    ell = np.arange(lmax + 1)
    return np.exp(-0.5 * ell * (ell + 1) * sigma ** 2)


def _band_limit_gauss_beam(lmax, fwhm=None):
    if fwhm is None:
        fwhm = 2 * np.pi / lmax
    return gauss_beam(fwhm, lmax)


def muKarcmin_to_n_ell(noise_level):
    return (noise_level * (np.pi / 60 / 180)) ** 2


def get_white_icov_pix(noise_level, minfo, out=None, dtype=np.float64):
    '''Inverse noise covariance map W / sigma^2 for a noise level in
    uK arcmin (host numpy).'''
    noise_level = np.atleast_1d(np.asarray(noise_level))
    if noise_level.ndim != 1:
        raise ValueError(f'{noise_level.shape=} not supported.')
    oshape = noise_level.shape + (minfo.npix,)
    if out is None:
        out = np.zeros(oshape, dtype=dtype)
    if out.shape != oshape:
        raise ValueError(f'{out.shape=} != {oshape=}')
    if out.dtype != dtype:
        raise ValueError(f'{out.dtype=} != {dtype=}')
    from optweight_amd import map_utils
    for idx in np.ndindex(noise_level.shape):
        out[idx] = 1 / muKarcmin_to_n_ell(noise_level[idx])
        map_utils._ringweight_np(out[idx], minfo, minfo.weight)
    return out
