'''
Per-multipole operations on device alm, the interface of optweight's
alm_c_utils Cython module (cython/alm_c_utils.pyx): lmul (157-250),
trunc_alm (84-137) and wlm2alm (4-62), running on the HIP kernels
owmi_lmul / owmi_trunc_alm / owmi_wlm2alm.
'''
import collections
import hashlib

import numpy as np
import torch

from optweight_amd import _lib


def _lmax_of(ainfo):
    return ainfo if isinstance(ainfo, (int, np.integer)) else ainfo.lmax


def _nelem(lmax):
    return (lmax * lmax + 3 * lmax) // 2 + 1


# Device copies of host multipole weights, by content: callers pass freshly
# computed numpy arrays on every call (filters, beams), and a host-to-device
# copy per call would stall the stream (and cannot be recorded into a HIP
# graph). Bounded LRU.
_LMAT_CACHE = collections.OrderedDict()
_LMAT_CACHE_MAX = 512


def _lmat_device(lmat, npol, rdtype, device):
    '''lmat cast to the alm precision (alm_c_utils.pyx:224-238), on device.'''
    if isinstance(lmat, torch.Tensor) and lmat.device == torch.device(device):
        t = lmat.to(dtype=rdtype)
    else:
        a = lmat.detach().cpu().numpy() if isinstance(lmat, torch.Tensor) else lmat
        a = np.ascontiguousarray(a)
        key = (str(device), rdtype, a.dtype.str, a.shape,
               hashlib.blake2b(a.tobytes(), digest_size=16).digest())
        t = _LMAT_CACHE.get(key)
        if t is None:
            t = torch.as_tensor(a, device=device).to(rdtype)
            _LMAT_CACHE[key] = t
            if len(_LMAT_CACHE) > _LMAT_CACHE_MAX:
                _LMAT_CACHE.popitem(last=False)
        else:
            _LMAT_CACHE.move_to_end(key)
    if t.ndim == 1:
        t = t.expand(npol, t.shape[0])
    return t.contiguous()


def lmul(alm, lmat, ainfo, alm_out=None, inplace=False):
    '''a'[..., i, lm] = m[i, j, l] a[..., j, lm]; lmat (npol, npol, nell),
    (npol, nell) (diagonal) or (nell,).'''
    if not isinstance(alm, torch.Tensor):
        raise ValueError('alm must be a device tensor')
    _lib.require_cuda(alm)
    if alm.dtype not in (torch.complex64, torch.complex128):
        raise ValueError(f'{alm.dtype=} should be complex64/128')
    lmax = _lmax_of(ainfo)
    nelem = _nelem(lmax)
    shape_in = alm.shape
    a3 = alm.reshape((-1,) + tuple(alm.shape[-2:])) if alm.ndim >= 2 else alm.reshape(1, 1, -1)
    npol = a3.shape[-2]
    if a3.shape[-1] != nelem:
        raise ValueError(f'nelem ({nelem}) != alm.shape[-1] ({a3.shape[-1]}); '
                         'only triangular alm storage supported')
    if alm_out is not None and inplace:
        raise ValueError('Cannot set both inplace and alm_out')
    if inplace:
        out = alm
    elif alm_out is None:
        out = torch.empty_like(alm)
    else:
        if alm_out.numel() != alm.numel():
            raise ValueError(f'alm.size {alm.numel()} != alm_out.size {alm_out.numel()}')
        if alm_out.dtype != alm.dtype:
            raise ValueError(f'alm.dtype {alm.dtype} != alm_out.dtype {alm_out.dtype}')
        _lib.require_cuda(alm_out)
        out = alm_out
    rdtype = torch.float64 if alm.dtype == torch.complex128 else torch.float32
    L = _lmat_device(lmat, npol, rdtype, alm.device)
    if L.ndim == 3:
        if tuple(L.shape) != (npol, npol, lmax + 1):
            raise ValueError(f'lmat shape {tuple(L.shape)} != {(npol, npol, lmax + 1)}')
        diag = 0
    elif L.ndim == 2:
        if tuple(L.shape) != (npol, lmax + 1):
            raise ValueError(f'lmat shape {tuple(L.shape)} != {(npol, lmax + 1)}')
        diag = 1
    else:
        raise ValueError(f'lmat ndim {L.ndim} not supported')
    _lib.call('owmi_lmul', _lib.dtype_code(alm), _lib.ptr(L), diag, _lib.ptr(alm),
              _lib.ptr(out), lmax, npol, a3.shape[0], _lib.stream_ptr(alm.device))
    return out.reshape(shape_in)


def lmul_add(a, lmat_a, b, lmat_b, ainfo, out, dot_out=None):
    '''out = lmat_a a + lmat_b b (lmat_a None: a itself) in one pass, full
    (npol, npol, nell) device matrices of the alm precision; a or b may be
    out. dot_out: float64 device tensor (ntrans,) that receives the
    real-field <b, out> of every transform in the same pass (b may then not
    be out).'''
    _lib.require_cuda(a, b, out)
    if not (a.shape == b.shape == out.shape) or not (a.dtype == b.dtype == out.dtype):
        raise ValueError('a, b and out must have one shape and dtype')
    lmax = _lmax_of(ainfo)
    nelem = _nelem(lmax)
    if a.shape[-1] != nelem:
        raise ValueError(f'nelem ({nelem}) != alm.shape[-1] ({a.shape[-1]})')
    npol = a.shape[-2] if a.ndim >= 2 else 1
    rdtype = torch.float64 if a.dtype == torch.complex128 else torch.float32
    mats = []
    for L in (lmat_a, lmat_b):
        if L is None:
            mats.append(None)
            continue
        L = _lmat_device(L, npol, rdtype, a.device)
        if tuple(L.shape) != (npol, npol, lmax + 1):
            raise ValueError(f'lmat shape {tuple(L.shape)} != {(npol, npol, lmax + 1)}')
        mats.append(L)
    if mats[1] is None:
        raise ValueError('lmat_b is required')
    ntrans = a.numel() // (npol * nelem)
    if dot_out is not None:
        _lib.require_cuda(dot_out)
        if dot_out.dtype != torch.float64 or dot_out.numel() != ntrans:
            raise ValueError(f'dot_out must be float64 with {ntrans} elements')
        _lib.call('owmi_lmul_add_dot', _lib.dtype_code(a),
                  None if mats[0] is None else _lib.ptr(mats[0]), _lib.ptr(a),
                  _lib.ptr(mats[1]), _lib.ptr(b), _lib.ptr(out), lmax, npol, ntrans,
                  _lib.ptr(dot_out), _lib.stream_ptr(a.device))
        return out
    _lib.call('owmi_lmul_add', _lib.dtype_code(a),
              None if mats[0] is None else _lib.ptr(mats[0]), _lib.ptr(a), _lib.ptr(mats[1]),
              _lib.ptr(b), _lib.ptr(out), lmax, npol, ntrans, _lib.stream_ptr(a.device))
    return out


def trunc_alm(alm, alm_out, lmax_in, lmax_out):
    '''Copy l, m <= lmax_out into alm_out (truncation to a smaller lmax).'''
    _lib.require_cuda(alm, alm_out)
    if alm.dtype != alm_out.dtype:
        raise ValueError('dtype mismatch')
    if alm.shape[-1] != _nelem(lmax_in) or alm_out.shape[-1] != _nelem(lmax_out):
        raise ValueError('alm sizes do not match lmax')
    if alm.shape[:-1] != alm_out.shape[:-1]:
        raise ValueError('leading dimensions do not match')
    if lmax_out > lmax_in:
        raise ValueError(f'lmax_out {lmax_out} > lmax_in {lmax_in}')
    nrows = alm.numel() // alm.shape[-1]
    _lib.call('owmi_trunc_alm', _lib.dtype_code(alm), _lib.ptr(alm), _lib.ptr(alm_out),
              lmax_in, lmax_out, nrows, _lib.stream_ptr(alm.device))
    return alm_out


def wlm2alm(w_ell, wlm, alm, lmax_w, lmax_a):
    '''alm[..., lm] += w_ell[l] wlm[..., lm] for l, m <= lmax_w.'''
    _lib.require_cuda(wlm, alm)
    if wlm.dtype != alm.dtype:
        raise ValueError('dtype mismatch')
    if wlm.shape[:-1] != alm.shape[:-1]:
        raise ValueError('leading dimensions do not match')
    if wlm.shape[-1] != _nelem(lmax_w) or alm.shape[-1] != _nelem(lmax_a):
        raise ValueError('alm sizes do not match lmax')
    rdtype = torch.float64 if alm.dtype == torch.complex128 else torch.float32
    w = _lmat_device(w_ell, 1, rdtype, alm.device).reshape(-1)
    if w.shape[0] < lmax_w + 1:
        raise ValueError('w_ell too short')
    nrows = alm.numel() // alm.shape[-1]
    _lib.call('owmi_wlm2alm', _lib.dtype_code(alm), _lib.ptr(w), _lib.ptr(wlm), _lib.ptr(alm),
              lmax_w, lmax_a, nrows, _lib.stream_ptr(alm.device))
    return alm
