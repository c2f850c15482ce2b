'''
In-place ring weights on device maps, the interface of optweight's
map_c_utils.apply_ringweight (cython/map_c_utils.pyx:6-51) running on the
HIP kernel owmi_apply_ringweight.
'''
import numpy as np
import torch

from optweight_amd import _lib


def apply_ringweight(imap, minfo, custom_weight=None):
    '''imap[..., ring pixels] *= weight[ring] in place; imap (npix,) or
    (ncomp, npix) device tensor.'''
    _lib.require_cuda(imap)
    if imap.shape[-1] != minfo.npix:
        raise ValueError(f'input {tuple(imap.shape)=} disagrees with {minfo.npix=}')
    if not (0 < imap.ndim < 3):
        raise ValueError(f'input map has to be 1 or 2d, got {imap.ndim=}')
    if imap.dtype not in (torch.float32, torch.float64):
        raise ValueError(f'unsupported dtype {imap.dtype}')
    w, nphi, offsets, stride = minfo.device_arrays(imap.device)
    if custom_weight is not None:
        cw = np.asarray(custom_weight)
        if cw.shape != minfo.weight.shape:
            raise ValueError(f'{cw.shape=}, expected {minfo.weight.shape=}')
        w = torch.as_tensor(cw, dtype=torch.float64, device=imap.device)
    nmaps = 1 if imap.ndim == 1 else imap.shape[0]
    _lib.call('owmi_apply_ringweight', _lib.dtype_code(imap), _lib.ptr(imap), _lib.ptr(w),
              _lib.ptr(nphi), _lib.ptr(offsets), _lib.ptr(stride), minfo.nrow, nmaps,
              imap.shape[-1], _lib.stream_ptr(imap.device))
    return imap


def map_fused(a, b=None, x=None, dcomp=None, mask=None, out=None):
    """out = x + dcomp[c] mask[c] (a - b) on maps (..., ncomp, npix) in one
    pass (owmi_map_fused); b, x, dcomp (ncomp,), mask (ncomp, npix) optional,
    out (default: a new array) may be any of the inputs."""
    _lib.require_cuda(a)
    ncomp, npix = a.shape[-2], a.shape[-1]
    for t in (b, x):
        if t is not None:
            _lib.require_cuda(t)
            if t.shape != a.shape or t.dtype != a.dtype:
                raise ValueError('map_fused: arrays must share shape and dtype')
    if dcomp is not None:
        _lib.require_cuda(dcomp)
        if dcomp.numel() != ncomp or dcomp.dtype != a.dtype:
            raise ValueError('map_fused: dcomp must be (ncomp,) of the map dtype')
    if mask is not None:
        _lib.require_cuda(mask)
        if tuple(mask.shape) != (ncomp, npix) or mask.dtype != a.dtype:
            raise ValueError('map_fused: mask must be (ncomp, npix) of the map dtype')
    if out is None:
        out = torch.empty_like(a)
    _lib.require_cuda(out)
    if out.shape != a.shape or out.dtype != a.dtype:
        raise ValueError('map_fused: out must match a')
    opt = lambda t: None if t is None else _lib.ptr(t)
    _lib.call('owmi_map_fused', _lib.dtype_code(a), _lib.ptr(out), opt(x), _lib.ptr(a), opt(b),
              opt(dcomp), opt(mask), ncomp, npix, a.numel() // (ncomp * npix),
              _lib.stream_ptr(a.device))
    return out
