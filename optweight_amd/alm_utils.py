'''
alm layout and algebra, mirroring optweight's alm_utils: contract_almxblm
(alm_utils.py:9-44), alm2wlm_axisym (46-103), wlm2alm_axisym (105-155),
trunc_alm (180-241), unit_var_alm (247-307), rand_alm (309-338),
add_to_alm (428-480); plus the
pixell alm_info layout optweight relies on (triangular, mmax = lmax,
m-major, stride 1).
'''
import numpy as np
import torch

from optweight_amd import _lib, mat_utils


class AlmInfo:
    '''Triangular alm layout with mmax = lmax (the fields of pixell's
    alm_info that optweight uses: lmax, mmax, nelem, stride, mstart).'''

    def __init__(self, lmax):
        self.lmax = int(lmax)
        self.mmax = self.lmax
        self.stride = 1
        self.nelem = (self.lmax ** 2 + 3 * self.lmax) // 2 + 1
        m = np.arange(self.mmax + 1)
        self.mstart = (m * (2 * self.lmax + 1 - m) // 2).astype(np.int64)

    def lm_arrays(self):
        ells, ms = [], []
        for m in range(self.mmax + 1):
            ells.append(np.arange(m, self.lmax + 1))
            ms.append(np.full(self.lmax + 1 - m, m))
        return np.concatenate(ells), np.concatenate(ms)


def alm_info(lmax):
    return AlmInfo(lmax)


def getlmax(nelem):
    lmax = int((-3 + np.sqrt(1 + 8 * nelem)) / 2)
    if (lmax * lmax + 3 * lmax) // 2 + 1 != nelem:
        raise ValueError(f'{nelem} is not a triangular alm size')
    return lmax


def trunc_alm(alm, ainfo, lmax):
    '''Device alm truncated to l, m <= lmax: (alm_trunc, ainfo_trunc).'''
    from optweight_amd import alm_c_utils
    if lmax > ainfo.lmax:
        raise ValueError(f'New lmax {lmax} exceeds old lmax {ainfo.lmax}')
    winfo = AlmInfo(lmax)
    out = torch.empty(tuple(alm.shape[:-1]) + (winfo.nelem,), dtype=alm.dtype,
                      device=alm.device)
    alm_c_utils.trunc_alm(alm.contiguous(), out, ainfo.lmax, lmax)
    return out, winfo


_ADD_INDEX = {}


def add_to_alm(alm, blm, ainfo, binfo, overwrite=False):
    '''In place on device: alm[..., lm_a(l, m)] += blm[..., lm_b(l, m)]
    (or =, with overwrite) for l, m <= binfo.lmax; returns alm. Raises
    ValueError when blm's lmax or mmax exceeds alm's.'''
    from optweight_amd import alm_c_utils
    if binfo.lmax > ainfo.lmax:
        raise ValueError('lmax blm exceeds that of alm : {} > {}'.format(binfo.lmax, ainfo.lmax))
    if binfo.mmax > ainfo.mmax:
        raise ValueError('mmax blm exceeds that of alm : {} > {}'.format(binfo.mmax, ainfo.mmax))
    _lib.require_cuda(alm, blm)
    if alm.dtype != blm.dtype:
        raise ValueError(f'dtype mismatch: {alm.dtype} != {blm.dtype}')
    if tuple(alm.shape[:-1]) != tuple(blm.shape[:-1]):
        raise ValueError('leading dimensions of alm and blm do not match')
    if overwrite:
        key = (str(alm.device), ainfo.lmax, binfo.lmax)
        idx = _ADD_INDEX.get(key)
        if idx is None:
            ells, ms = binfo.lm_arrays()
            pos = ainfo.mstart[ms] + ells
            idx = torch.as_tensor(pos, dtype=torch.long, device=alm.device)
            _ADD_INDEX[key] = idx
        alm[..., idx] = blm
        return alm
    ones = np.ones(binfo.lmax + 1)
    return alm_c_utils.wlm2alm(ones, blm.contiguous(), alm, binfo.lmax, ainfo.lmax)


def alm2wlm_axisym(alm, ainfo, w_ell, lmaxs=None):
    '''Wavelet coefficients w^j_lm = w_ell[j, l] a_lm, l <= lmax_j (lmax_j:
    last multipole with w_ell[j] > 0, or lmaxs[j]): (wlms, winfos).'''
    from optweight_amd import alm_c_utils
    w_ell = np.asarray(w_ell)
    if w_ell.ndim == 1:
        w_ell = w_ell[np.newaxis, :]
    lmax = ainfo.lmax
    if w_ell.shape[1] != lmax + 1:
        raise ValueError(f'lmax alm : {lmax} and lmax of w_ell : {w_ell.shape[1] - 1} '
                         'do not match')
    wlms, winfos = [], []
    for idx in range(w_ell.shape[0]):
        lmax_w = lmaxs[idx] if lmaxs is not None else lmax - np.argmax(w_ell[idx, ::-1] > 0)
        wlm, winfo = trunc_alm(alm, ainfo, int(lmax_w))
        alm_c_utils.lmul(wlm, w_ell[idx, :winfo.lmax + 1], winfo, inplace=True)
        wlms.append(wlm)
        winfos.append(winfo)
    return wlms, winfos


def wlm2alm_axisym(wlms, winfos, w_ell, alm=None, ainfo=None):
    '''alm += sum_j w_ell[j, l] w^j_lm (added to alm when given): (alm, ainfo).'''
    from optweight_amd import alm_c_utils
    w_ell = np.asarray(w_ell)
    if w_ell.ndim == 1:
        w_ell = w_ell[np.newaxis, :]
    lmax_w = max(w.lmax for w in winfos)
    if ainfo is None:
        ainfo = AlmInfo(lmax_w)
    elif ainfo.lmax < lmax_w:
        raise ValueError(f'lmax alm {ainfo.lmax} < lmax wavelets {lmax_w}')
    if alm is None:
        alm = torch.zeros(tuple(wlms[0].shape[:-1]) + (ainfo.nelem,), dtype=wlms[0].dtype,
                          device=wlms[0].device)
    for idx, (wlm, winfo) in enumerate(zip(wlms, winfos)):
        alm_c_utils.wlm2alm(w_ell[idx, :winfo.lmax + 1], wlm.contiguous(), alm, winfo.lmax,
                            ainfo.lmax)
    return alm, ainfo


def contract_almxblm_dev(alm, blm, out=None):
    '''Real-field inner product sum_{m=0} Re(a b*) + 2 sum_{m>0} Re(a b*) as
    a 0-d float64 device tensor (no host synchronisation).'''
    if tuple(blm.shape) != tuple(alm.shape):
        raise ValueError(f'Shape alm ({tuple(alm.shape)}) != shape blm ({tuple(blm.shape)})')
    _lib.require_cuda(alm, blm)
    if alm.dtype != blm.dtype:
        raise ValueError('dtype mismatch')
    lmax = getlmax(alm.shape[-1])
    if out is None:
        out = torch.empty((), dtype=torch.float64, device=alm.device)
    nrows = alm.numel() // alm.shape[-1]
    _lib.call('owmi_alm_dot', _lib.dtype_code(alm), _lib.ptr(alm), _lib.ptr(blm), lmax, nrows,
              _lib.ptr(out), _lib.stream_ptr(alm.device))
    return out


def contract_almxblm_batch(alm, blm, out=None):
    '''Per-right-hand-side real-field inner products of alm, blm with shape
    (nbatch, ..., nelem): float64 device tensor (nbatch,), no host
    synchronisation.'''
    if tuple(blm.shape) != tuple(alm.shape):
        raise ValueError(f'Shape alm ({tuple(alm.shape)}) != shape blm ({tuple(blm.shape)})')
    _lib.require_cuda(alm, blm)
    if alm.dtype != blm.dtype:
        raise ValueError('dtype mismatch')
    lmax = getlmax(alm.shape[-1])
    nb = alm.shape[0]
    if out is None:
        out = torch.empty((nb,), dtype=torch.float64, device=alm.device)
    nrows = alm.numel() // (alm.shape[-1] * nb)
    _lib.call('owmi_alm_dot_batch', _lib.dtype_code(alm), _lib.ptr(alm), _lib.ptr(blm), lmax,
              nrows, nb, _lib.ptr(out), _lib.stream_ptr(alm.device))
    return out


def contract_almxblm(alm, blm):
    '''Python float of the real-field inner product (synchronises).'''
    if not isinstance(alm, torch.Tensor):
        alm = np.asarray(alm)
        blm = np.asarray(blm)
        if blm.shape != alm.shape:
            raise ValueError(f'Shape alm ({alm.shape}) != shape blm ({blm.shape})')
        lmax = getlmax(alm.shape[-1])
        bc = np.conj(blm)
        had = 2 * np.real(complex(np.tensordot(alm, bc, axes=alm.ndim)))
        return had - np.real(np.sum(alm[..., :lmax + 1] * bc[..., :lmax + 1]))
    return float(contract_almxblm_dev(alm, blm).item())


def unit_var_alm(ainfo, preshape, seed, out=None, dtype=np.complex128):
    '''Unit-variance complex Gaussian alm (host numpy): normal(scale=1/sqrt2)
    over interleaved (re, im); m = 0 imaginary parts zero, real parts
    scaled by sqrt(2).'''
    if dtype not in (np.complex64, np.complex128):
        raise ValueError(f'{dtype=} not supported')
    oshape = tuple(preshape) + (ainfo.nelem,)
    if out is None:
        out = np.zeros(oshape, dtype=dtype)
    elif out.shape != oshape or out.dtype != dtype:
        raise ValueError(f'{oshape=}, {dtype=} != {out.shape=}, {out.dtype=}')
    out = mat_utils.atleast_nd(out, 2)
    shape_nd = out.shape
    rng = np.random.default_rng(seed)
    rtype = np.float32 if dtype == np.complex64 else np.float64
    flat = out.reshape(-1).view(rtype)
    flat[:] = rng.normal(scale=1 / np.sqrt(2), size=flat.size).astype(rtype, copy=False)
    out = flat.view(dtype).reshape(shape_nd)
    out[..., :ainfo.lmax + 1].imag = 0
    out[..., :ainfo.lmax + 1].real *= np.sqrt(2)
    return out.reshape(oshape)


def rand_alm(cov_ell, ainfo, seed, out=None, dtype=np.complex128):
    '''Draw alm from a per-l covariance: cov^(1/2) applied to unit-variance
    alm (host numpy).'''
    cov_ell = mat_utils.atleast_nd(np.asarray(cov_ell), 2)
    a = unit_var_alm(ainfo, (cov_ell.shape[0],), seed, out=out, dtype=dtype)
    sq = mat_utils.matpow(mat_utils.full_matrix(cov_ell), 0.5)
    ell, _ = ainfo.lm_arrays()
    rtype = np.float32 if dtype == np.complex64 else np.float64
    res = np.einsum('ijn,jn->in', sq[:, :, ell].astype(rtype), a).astype(dtype)
    a[...] = res
    return a


def rand_alm_pix(cov_pix, ainfo, minfo, spin, seed, adjoint=False, device=None):
    """Y^T W N^(1/2) r (Y^T N^(1/2) r if adjoint), r ~ N(0, 1) drawn on the
    host (alm_utils.py:340-373); device complex alm (npol, nelem)."""
    from optweight_amd import map_utils, sht
    cov_pix = np.asarray(cov_pix)
    if device is None:
        device = torch.device('cuda', torch.cuda.current_device())
    noise = torch.as_tensor(map_utils.rand_map_pix(cov_pix, seed), device=device)
    cdt = torch.complex64 if noise.dtype == torch.float32 else torch.complex128
    alm = torch.zeros((noise.shape[0], ainfo.nelem), dtype=cdt, device=device)
    sht.map2alm(noise, alm, minfo, ainfo, spin, adjoint=adjoint)
    return alm


def rand_alm_wav(cov_wav, ainfo, w_ell, spin, seed, adjoint=False, device=None):
    """sum_j K_j^T Y^T [W] N_j^(1/2) r_j with one draw per wavelet scale
    from a block-diagonal wavelet covariance (alm_utils.py:375-419)."""
    from optweight_amd import map_utils, sht, wavtrans
    if device is None:
        device = torch.device('cuda', torch.cuda.current_device())
    npol = wavtrans.preshape2npol(cov_wav.preshape)
    rng = np.random.default_rng(seed)
    w_ell = np.asarray(w_ell)
    out = None
    for jidx in range(cov_wav.shape[0]):
        cov = cov_wav.maps[(jidx, jidx)]
        cov = cov.detach().cpu().numpy() if isinstance(cov, torch.Tensor) else np.asarray(cov)
        minfo = cov_wav.minfos[(jidx, jidx)]
        rmap = torch.as_tensor(map_utils.rand_map_pix(cov, rng), device=device)
        cdt = torch.complex64 if rmap.dtype == torch.float32 else torch.complex128
        if out is None:
            out = torch.zeros((npol, ainfo.nelem), dtype=cdt, device=device)
        winfo = AlmInfo((int(minfo.nphi[0]) - 1) // 2)
        wlm = torch.zeros((npol, winfo.nelem), dtype=cdt, device=device)
        sht.map2alm(rmap, wlm, minfo, winfo, spin, adjoint=adjoint)
        wlm2alm_axisym([wlm], [winfo], w_ell[jidx:jidx + 1], alm=out, ainfo=ainfo)
    return out
