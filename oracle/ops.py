'''
Streaming and setup operations (oracle), restating:
  lmul / lmul_diag / trunc_alm / wlm2alm   alm_c_utils.c:12-247 (+ .pyx:157-250)
  apply_ringweight                        map_c_utils.c:7-37 (oracle.sht.ringweight)
  full_matrix / matpow / _eigpow / matvec mat_utils.py:125-155, 220-365, 475-532
  eigpow                                  mat_c_utils.c:9-157 (V diag(lam^p) V^T)
  contract_almxblm                        alm_utils.py:9-44
  unit_var_alm / rand_alm                 alm_utils.py:247-338
  rand_map_pix                            map_utils.py:1084-1122
  inv_qweight_map / get_isotropic_ivar    map_utils.py:1124-1218
  gauss_beam (healpy)                     noise_utils.py:401-421
  lowpass_filter                          multigrid.py:449-471
  threshold_icov                          map_utils.py:1297-1363
'''
import numpy as np

from oracle import layout, sht as osht


def full_matrix(mat):
    mat = np.asarray(mat)
    if mat.ndim not in (2, 3):
        raise ValueError('matrix must be 2 or 3 dimensional')
    if mat.ndim == 2:
        npol = mat.shape[0]
        mat = np.eye(npol, dtype=mat.dtype)[:, :, None] * mat
    return mat


def lmul(alm, lmat, lmax, alm_out=None, inplace=False):
    '''out[..., i, lm] = sum_j lmat[i, j, l(lm)] alm[..., j, lm].
    lmat is (npol, npol, nell), (npol, nell) (diagonal) or (nell,).
    lmat is cast to the alm precision first (alm_c_utils.pyx:224-238).'''
    if alm_out is not None and inplace:
        raise ValueError('Cannot set both inplace and alm_out')
    alm = np.asarray(alm)
    if alm.dtype not in (np.complex64, np.complex128):
        raise ValueError('alm dtype must be complex64/128')
    rtype = np.float32 if alm.dtype == np.complex64 else np.float64
    npol = alm.shape[-2] if alm.ndim >= 2 else 1
    lmat = np.asarray(lmat)
    if lmat.ndim == 1:
        lmat = lmat * np.ones((npol, lmat.size), dtype=lmat.dtype)
    lmat = lmat.astype(rtype)
    ell, _ = layout.lm_arrays(lmax)
    a = alm.reshape((-1, npol, alm.shape[-1]))
    if lmat.ndim == 3:
        res = np.einsum('ijn,tjn->tin', lmat[:, :, ell], a)
    else:
        res = lmat[None, :, ell] * a
    res = res.astype(alm.dtype).reshape(alm.shape)
    if inplace:
        alm[...] = res
        return alm
    if alm_out is not None:
        alm_out[...] = res
        return alm_out
    return res


def trunc_alm(alm, lmax, lmax_out):
    '''Keep l <= lmax_out, m <= lmax_out (alm_c_utils.c:219-247).'''
    out = np.zeros(alm.shape[:-1] + (layout.nelem(lmax_out),), dtype=alm.dtype)
    for m in range(lmax_out + 1):
        i0 = layout.mstart(m, lmax)
        j0 = layout.mstart(m, lmax_out)
        out[..., j0 + m:j0 + lmax_out + 1] = alm[..., i0 + m:i0 + lmax_out + 1]
    return out


def add_to_alm(alm, blm, lmax_a, lmax_b, overwrite=False):
    '''In place: alm[..., lm_a(l, m)] (+)= blm[..., lm_b(l, m)] for l, m
    <= lmax_b (alm_utils.py:428-480; mmax = lmax on both sides).'''
    if lmax_b > lmax_a:
        raise ValueError(f'lmax blm exceeds that of alm : {lmax_b} > {lmax_a}')
    for m in range(lmax_b + 1):
        i0 = layout.mstart(m, lmax_a) + m
        j0 = layout.mstart(m, lmax_b) + m
        n = lmax_b + 1 - m
        if overwrite:
            alm[..., i0:i0 + n] = blm[..., j0:j0 + n]
        else:
            alm[..., i0:i0 + n] += blm[..., j0:j0 + n]
    return alm


def wlm2alm(w_ell, wlm, alm, lmax_w, lmax_a):
    '''alm[..., lm] += w_ell[l] wlm[..., lm'] for l, m <= lmax_w
    (alm_c_utils.c:113-144).'''
    for m in range(lmax_w + 1):
        i0 = layout.mstart(m, lmax_a)
        j0 = layout.mstart(m, lmax_w)
        ells = np.arange(m, lmax_w + 1)
        alm[..., i0 + ells] += (w_ell[ells].astype(alm.real.dtype)
                                * wlm[..., j0 + ells])
    return alm


def contract_almxblm(alm, blm):
    '''2 Re sum a conj(b) - Re sum_{m=0} a conj(b)  (alm_utils.py:9-44).'''
    if alm.shape != blm.shape:
        raise ValueError('shape mismatch')
    lmax = layout.getlmax(alm.shape[-1])
    bc = np.conj(blm)
    csum = complex(np.tensordot(alm, bc, axes=alm.ndim))
    out = 2 * np.real(csum)
    out -= np.real(np.sum(alm[..., :lmax + 1] * bc[..., :lmax + 1]))
    return out


def eigpow(mat, power, lim=None, lim0=None):
    '''(n, n, N) symmetric PSD matrices -> V diag(lam^p) V^T with
    lam < lim * lam_max set to 0 and all-zero output when lam_max < lim0
    (mat_c_utils.c:9-157). Computed in float64 with numpy.linalg.eigh.'''
    if lim is None:
        lim = 1e-6
    if lim0 is None:
        lim0 = np.finfo(np.float64).tiny ** 0.5
    m = np.moveaxis(np.asarray(mat, dtype=np.float64), -1, 0)
    w, v = np.linalg.eigh(m)
    wmax = np.max(w, axis=-1, keepdims=True)
    mask = (w >= lim * wmax) & (wmax >= lim0)
    wp = np.zeros_like(w)
    wp[mask] = w[mask] ** power
    out = np.einsum('nij,nj,nkj->nik', v, wp, v)
    return np.moveaxis(out, 0, -1)


def matpow(mat, power, return_diag=False, lim=1e-9, lim0=None):
    '''mat_utils.matpow for the (npol,npol,N) / (npol,N) cases.'''
    mat = np.asarray(mat)
    if mat.ndim == 1:
        mat = mat[None, :]
    if power != 1:
        if mat.ndim == 2:
            out = mat.copy()
            mask = out <= 0
            out[mask] = 0
            np.power(out, power, out=out, where=~mask)
            mat = out
        else:
            mat = eigpow(mat, power, lim=lim, lim0=lim0).astype(mat.dtype)
    if not return_diag:
        mat = full_matrix(mat)
    return mat.copy()


def matvec(mat, vec):
    '''mat_utils.matvec: elementwise if shapes match, else per-pixel
    (n, n, N) x (n, N).'''
    if mat.shape == vec.shape:
        return mat * vec
    return np.einsum('ijk,jk->ik', mat, vec)


def unit_var_alm(lmax, npol, seed, dtype=np.complex128):
    '''alm_utils.unit_var_alm: normal(scale=1/sqrt2) over interleaved re/im,
    m=0 imag = 0 and real * sqrt2.'''
    n = layout.nelem(lmax)
    rng = np.random.default_rng(seed)
    rtype = np.float32 if dtype == np.complex64 else np.float64
    out = np.zeros((npol, n), dtype=dtype)
    flat = out.reshape(-1).view(rtype)
    flat[:] = rng.normal(scale=1 / np.sqrt(2), size=flat.size).astype(rtype, copy=False)
    out[..., :lmax + 1].imag = 0
    out[..., :lmax + 1].real *= np.sqrt(2)
    return out


def rand_alm(cov_ell, lmax, seed, dtype=np.complex128):
    cov_ell = np.asarray(cov_ell)
    if cov_ell.ndim == 1:
        cov_ell = cov_ell[None]
    a = unit_var_alm(lmax, cov_ell.shape[0], seed, dtype=dtype)
    sq = matpow(full_matrix(cov_ell), 0.5)
    return lmul(a, sq, lmax)


def rand_map_pix(cov_pix, seed):
    npol, npix = cov_pix.shape[-2:]
    rng = np.random.default_rng(seed)
    uv = rng.standard_normal(npol * npix).astype(cov_pix.dtype, copy=False)
    uv = uv.reshape((npol, npix))
    if cov_pix.ndim == 2:
        return uv * np.sqrt(cov_pix)
    return np.einsum('ijk,jk->ik', matpow(cov_pix, 0.5), uv)


def inv_qweight_map(imap, geom):
    out = np.array(imap, copy=True)
    w = 1 / geom.weight
    for idx in np.ndindex(out.shape[:-1]):
        osht.ringweight(out[idx], geom, weight=w)
    return out


def get_isotropic_ivar(icov_pix, geom, mask=None):
    '''tau_ii = sum (icov_ii/w)^2 / sum (icov_ii/w) (diagonal only).'''
    wcov = inv_qweight_map(icov_pix, geom)
    if mask is not None:
        if wcov.ndim == 2:
            wcov = wcov * mask
        else:
            sm = np.sqrt(np.atleast_2d(mask))
            wcov = wcov * np.einsum('ap,bp->abp', sm, sm)
    num = np.sum(wcov ** 2, axis=-1)
    den = np.sum(wcov, axis=-1)
    if icov_pix.ndim == 2:
        return np.diag(num / den)
    return np.diag(np.diag(num) / np.diag(den))


def gauss_beam(fwhm, lmax):
    '''healpy.gauss_beam (temperature): exp(-l(l+1) sigma^2 / 2).'''
    sigma = fwhm / np.sqrt(8 * np.log(2))
    ell = np.arange(lmax + 1)
    return np.exp(-0.5 * ell * (ell + 1) * sigma ** 2)


def band_limit_gauss_beam(lmax, fwhm=None):
    if fwhm is None:
        fwhm = 2 * np.pi / lmax
    return gauss_beam(fwhm, lmax)


def lowpass_filter(lmax_r_ell, lmax=None):
    if lmax is None:
        lmax = lmax_r_ell
    beta = -np.log(np.sqrt(0.05)) / ((lmax_r_ell / 2) * (lmax_r_ell / 2 + 1)) ** 2
    ells = np.arange(lmax + 1)
    return np.exp(-beta * (ells * (ells + 1)) ** 2.)


def threshold_icov(icov_pix, q_low=0, q_high=1):
    '''Clip the diagonal of icov_pix (npol, npix) or (npol, npol, npix) per
    component to its [q_low, q_high] quantiles over the nonzero pixels; a 3d
    matrix is rescaled symmetrically by sqrt of the per-pixel factors.'''
    icov_pix = np.asarray(icov_pix)
    if icov_pix.ndim not in (2, 3):
        raise ValueError(f'{icov_pix.ndim=}, expected 2 or 3')
    npol = icov_pix.shape[0]
    q_low = q_low * np.ones(npol)
    q_high = q_high * np.ones(npol)
    scale = np.ones((npol, icov_pix.shape[-1]))
    for p in range(npol):
        d = icov_pix[p, p] if icov_pix.ndim == 3 else icov_pix[p]
        nz = d > 0
        if not np.any(nz):
            continue
        lo = np.quantile(d[nz], q_low[p])
        hi = np.quantile(d[nz], q_high[p])
        ml = (d < lo) & nz
        mh = (d > hi) & nz
        scale[p][ml] = lo / d[ml]
        scale[p][mh] = hi / d[mh]
    if icov_pix.ndim == 2:
        out = icov_pix * scale
    else:
        r = np.sqrt(scale)
        out = r[:, None, :] * icov_pix * r[None, :, :]
    return out.astype(icov_pix.dtype, copy=False)
