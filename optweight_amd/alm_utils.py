'''
alm layout and algebra, mirroring optweight's alm_utils: contract_almxblm
(alm_utils.py:9-44), unit_var_alm (247-307), rand_alm (309-338); plus the
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
