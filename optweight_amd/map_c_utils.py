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
