'''
Noise helpers the Wiener path needs, mirroring optweight's noise_utils:
_band_limit_gauss_beam (noise_utils.py:401-421, healpy.gauss_beam closed
form exp(-l(l+1) sigma^2 / 2), sigma = fwhm / sqrt(8 ln 2)),
get_white_icov_pix (592-641), and the noise-model estimation
estimate_cov_wav (10-77), estimate_cov_fwav (79-142), estimate_cov_pix
(144-226), norm_cov_est (228-253) with map_utils.lmul_pix's spin-0
smoothing (map_utils.py:1365-1408) or the flat-sky one of fmul_pix
(1751-1793) on device maps, and unit_var_wav (423-458).
'''
import numpy as np
import torch


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


def norm_cov_est(cov_est, minfo, kernel_ell, inplace=False):
    '''Scale a covariance map estimated from a band-limited noise map
    (support kernel_ell in multipole) so that draws filtered by the same
    kernel get the right power: divide by the pixel area w and by
    sum_l kernel_ell^2 (2l + 1) / (4 pi). Device tensor (..., npix).'''
    kernel_ell = np.asarray(kernel_ell, dtype=np.float64)
    ells = np.arange(kernel_ell.size)
    norm = np.sum(kernel_ell ** 2 * (2 * ells + 1)) / 4 / np.pi
    out = cov_est if inplace else cov_est.clone()
    cache = minfo.__dict__.setdefault('_inv_area', {})
    key = (str(out.device), out.dtype)
    if key not in cache:
        cache[key] = torch.as_tensor(np.repeat(1.0 / minfo.weight, minfo.nphi), dtype=out.dtype,
                                     device=out.device)
    out *= cache[key]
    out /= norm
    return out


def _smooth_spin0(maps, b_ell, minfo):
    '''lmul_pix with spin 0 on a stack of maps (k, npix): Y b_ell Y^T W.'''
    from optweight_amd import sht, alm_c_utils, alm_utils
    lmax = b_ell.size - 1
    ainfo = alm_utils.AlmInfo(lmax)
    cdt = torch.complex128 if maps.dtype == torch.float64 else torch.complex64
    m3 = maps.reshape(-1, 1, maps.shape[-1]).contiguous()
    alm = torch.empty((m3.shape[0], 1, ainfo.nelem), dtype=cdt, device=maps.device)
    sht.map2alm(m3, alm, minfo, ainfo, 0)
    alm_c_utils.lmul(alm, b_ell, ainfo, inplace=True)
    out = torch.empty_like(m3)
    sht.alm2map(alm, out, ainfo, minfo, 0)
    return out.reshape(maps.shape)


def _smooth_flat(maps, fb, minfo):
    '''map_utils.fmul_pix with a shared 2-D Fourier diagonal fb on a stack
    of maps (k, npix) of a rectangular grid.'''
    from optweight_amd import map_utils
    m2 = maps.reshape(-1, maps.shape[-1]).contiguous()
    return map_utils.fmul_pix(m2, minfo, fmat2d=fb).reshape(maps.shape)


def estimate_cov_pix(imap, minfo, diag=False, fwhm=None, lmax=None, kernel_ell=None,
                     flatsky=False, modlmap=None):
    '''Pixel covariance of a noise map (ncomp, npol, npix) (device tensor or
    host array, result on the same side as a device tensor): the outer
    product per pixel, (ncomp, ncomp, npol, npix) with diag (no
    polarisation cross terms) else (ncomp, npol, ncomp, npol, npix),
    smoothed with a Gaussian of this fwhm (default 2 pi / lmax of the grid)
    as spin-0 maps (upper triangle, the lower one mirrored), normalised
    with norm_cov_est when the noise is band limited (lmax or kernel_ell),
    and made positive semi-definite per pixel (negative entries clipped to
    zero with diag). flatsky smooths in the 2-D Fourier plane instead, with
    the beam sampled on modlmap (|l| of the map's rfft coefficients).'''
    if not isinstance(imap, torch.Tensor):
        imap = torch.as_tensor(np.asarray(imap), device=torch.device(
            'cuda', torch.cuda.current_device()))
    while imap.ndim < 3:
        imap = imap.unsqueeze(0)
    if imap.ndim != 3:
        raise ValueError(f'imap is not (ncomp, npol, npix), but {tuple(imap.shape)}')
    ncomp, npol, npix = imap.shape
    if diag:
        cov = torch.einsum('abc, dbc -> adbc', imap, imap)
    else:
        cov = torch.einsum('abc, dec -> abdec', imap, imap).reshape(
            ncomp * npol, ncomp * npol, npix)
    if fwhm is None:
        fwhm = 2 * np.pi / map_utils_minfo2lmax(minfo)
    b_ell = _band_limit_gauss_beam(map_utils_minfo2lmax(minfo), fwhm=fwhm)
    if fwhm > 0:
        n0, n1 = cov.shape[:2]
        iu = [(i, j) for i in range(n0) for j in range(n1) if i <= j]
        ii = torch.as_tensor([i for i, _ in iu], device=cov.device)
        jj = torch.as_tensor([j for _, j in iu], device=cov.device)
        if flatsky:
            from optweight_amd import dft
            fb = dft.cl2flat(b_ell, np.arange(b_ell.shape[-1]), modlmap)
            sm = _smooth_flat(cov[ii, jj], fb, minfo)
        else:
            sm = _smooth_spin0(cov[ii, jj], b_ell, minfo)
        cov[ii, jj] = sm
        cov[jj, ii] = sm
    if lmax or kernel_ell is not None:
        if kernel_ell is None:
            kernel_ell = np.ones(lmax + 1)
        norm_cov_est(cov, minfo, kernel_ell, inplace=True)
    if not diag:
        # nearest positive semi-definite matrix per pixel (negative
        # eigenvalues dropped), on the host as the reference's set-up does
        from optweight_amd import mat_utils
        host = mat_utils.matpow(cov.cpu().numpy(), 1, skip_unit_pow=False)
        cov = torch.as_tensor(host, device=cov.device).reshape(ncomp, npol, ncomp, npol, npix)
    else:
        cov.clamp_(min=0)
    return cov


def map_utils_minfo2lmax(minfo):
    return int(0.5 * (np.max(minfo.nphi) - 1))


def estimate_cov_wav(alm, ainfo, w_ell, spin, diag=False, wav_template=None, fwhm_fact=2):
    '''Wavelet-based noise covariance from noise alm (ncomp, npol, nelem):
    the monopole is left out, the alm are split into wavelet maps
    (wavtrans.alm2wav) and each scale's pixel covariance is estimated with
    the smoothing FWHM = fwhm_fact * pi / lmax_j (fwhm_fact a number or a
    function of lmax_j) and the scale's kernel as its support; returns an
    (nwav, nwav) diagonal Wav of covariance maps.'''
    from optweight_amd import wavtrans
    w_ell = np.asarray(w_ell)
    if w_ell.ndim == 1:
        w_ell = w_ell[np.newaxis, :]
    mono = alm[..., 0, 0].clone()
    alm[..., 0, 0] = 0
    try:
        noise_wav = wavtrans.alm2wav(alm, ainfo, spin, w_ell, wav=wav_template)
    finally:
        alm[..., 0, 0] = mono
    rdt = np.float64 if alm.dtype == torch.complex128 else np.float32
    cov_wav = wavtrans.Wav(2, dtype=rdt)
    for jidx in range(w_ell.shape[0]):
        minfo = noise_wav.minfos[jidx]
        lmax_j = map_utils_minfo2lmax(minfo)
        fact = fwhm_fact(lmax_j) if callable(fwhm_fact) else fwhm_fact
        cov_pix = estimate_cov_pix(noise_wav.maps[jidx], minfo, kernel_ell=w_ell[jidx],
                                   diag=diag, fwhm=fact * np.pi / lmax_j)
        cov_wav.add((jidx, jidx), cov_pix, minfo)
    return cov_wav


def unit_var_wav(minfos, preshape, dtype, seed=None):
    """Block vector of unit-variance Gaussian maps, one per map info, drawn
    in order from numpy's default_rng(seed) (host arrays)."""
    from optweight_amd import wavtrans
    rng = np.random.default_rng(seed)
    wav = wavtrans.Wav(1, preshape=preshape, dtype=dtype)
    for widx, minfo in enumerate(minfos):
        wav.add(widx, rng.normal(size=tuple(preshape) + (minfo.npix,)).astype(dtype), minfo)
    return wav


def estimate_cov_fwav(fmap, fkernelset, wav_template, diag=False, fwhm_fact=2):
    """Fourier-wavelet noise covariance (nwav, nwav) block diagonal from
    noise 2-D coefficients fmap (device): the monopole is left out (and put
    back afterwards), fmap is cut into wavelet maps (wavtrans.f2wav into
    wav_template), and each scale's pixel covariance is estimated with
    flat-sky smoothing of FWHM fwhm_fact pi / lmax_j and divided by the
    mean squared kernel (noise_utils.py:79-142)."""
    from optweight_amd import wavtrans
    mono = fmap[..., 0, 0].clone()
    fmap[..., 0, 0] = 0
    noise_wav = wavtrans.f2wav(fmap, wav_template, fkernelset)
    fmap[..., 0, 0] = mono
    rdt = np.float64 if fmap.dtype == torch.complex128 else np.float32
    cov_wav = wavtrans.Wav(2, dtype=rdt)
    for jidx, fk in fkernelset:
        minfo = noise_wav.minfos[jidx]
        lmax_j = fk.lmax
        fact = fwhm_fact(lmax_j) if callable(fwhm_fact) else fwhm_fact
        cov = estimate_cov_pix(noise_wav.maps[jidx], minfo, diag=diag, fwhm=fact * np.pi / lmax_j,
                               flatsky=True, modlmap=fk.modlmap())
        cov /= float(np.mean(np.abs(np.asarray(fk.fkernel, dtype=np.float64)) ** 2))
        cov_wav.add((jidx, jidx), cov, minfo)
    return cov_wav
