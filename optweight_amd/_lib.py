'''
ctypes binding of libowmi.so (the C ABI declared in include/owmi.h).

The library is required: importing a transform or operator without it
raises, there is no CPU fallback.
'''
import ctypes
import os

import numpy as np
import torch

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, '_lib', 'libowmi.so')

F32 = 0
F64 = 1

_c_ll = ctypes.c_longlong
_c_int = ctypes.c_int
_c_dbl = ctypes.c_double
_vp = ctypes.c_void_p

_SIGNATURES = {
    'owmi_last_error': (ctypes.c_char_p, []),
    'owmi_abi_version': (_c_int, []),
    'owmi_lmul': (_c_int, [_c_int, _vp, _c_int, _vp, _vp, _c_int, _c_int, _c_ll, _vp]),
    'owmi_trunc_alm': (_c_int, [_c_int, _vp, _vp, _c_int, _c_int, _c_ll, _vp]),
    'owmi_lmul_add': (_c_int, [_c_int, _vp, _vp, _vp, _vp, _vp, _c_int, _c_int, _c_ll, _vp]),
    'owmi_lmul_add_dot': (_c_int, [_c_int, _vp, _vp, _vp, _vp, _vp, _c_int, _c_int, _c_ll, _vp,
                                   _vp]),
    'owmi_map_fused': (_c_int, [_c_int, _vp, _vp, _vp, _vp, _vp, _vp, _c_int, _c_ll, _c_ll,
                                _vp]),
    'owmi_wlm2alm': (_c_int, [_c_int, _vp, _vp, _vp, _c_int, _c_int, _c_ll, _vp]),
    'owmi_alm_dot': (_c_int, [_c_int, _vp, _vp, _c_int, _c_ll, _vp, _vp]),
    'owmi_apply_ringweight': (_c_int, [_c_int, _vp, _vp, _vp, _vp, _vp, _c_ll, _c_ll, _c_ll, _vp]),
    'owmi_pix_matvec': (_c_int, [_c_int, _vp, _c_int, _vp, _vp, _c_int, _c_ll, _c_ll, _vp]),
    'owmi_cg_xr_mixed': (_c_int, [_c_int, _vp, _c_int, _vp, _vp, _vp, _vp, _vp, _c_ll, _vp]),
    'owmi_cg_xr': (_c_int, [_c_int, _vp, _vp, _vp, _vp, _vp, _vp, _c_ll, _vp]),
    'owmi_cg_p': (_c_int, [_c_int, _vp, _vp, _vp, _vp, _c_ll, _vp]),
    'owmi_alm_dot_batch': (_c_int, [_c_int, _vp, _vp, _c_int, _c_ll, _c_ll, _vp, _vp]),
    'owmi_alm_dot_batch_noalloc': (_c_int, [_c_int, _vp, _vp, _c_int, _c_ll, _c_ll, _vp, _vp]),
    'owmi_cg_xr_batch': (_c_int, [_c_int, _vp, _c_int, _vp, _vp, _vp, _vp, _vp, _c_ll, _c_ll,
                                  _vp]),
    'owmi_cg_xr_rr_batch': (_c_int, [_c_int, _vp, _c_int, _vp, _vp, _vp, _vp, _vp, _c_int,
                                     _c_ll, _c_ll, _vp, _vp]),
    'owmi_cg_p_batch': (_c_int, [_c_int, _vp, _vp, _vp, _vp, _c_ll, _c_ll, _vp]),
    'owmi_axpby': (_c_int, [_c_int, _c_dbl, _vp, _c_dbl, _vp, _c_ll, _vp]),
    'owmi_convert': (_c_int, [_c_int, _vp, _c_int, _vp, _c_ll, _vp]),
    'owmi_fmul': (_c_int, [_c_int, _vp, _c_int, _vp, _vp, _c_int, _c_ll, _c_ll, _c_ll, _vp]),
    'owmi_fmul_generic': (_c_int, [_c_int, _vp, _vp, _vp, _vp, _c_int, _c_ll, _c_ll, _vp]),
    'owmi_fwav_gather': (_c_int, [_c_int, _vp, _c_ll, _c_ll, _c_ll, _c_int, _vp, _vp, _vp, _c_ll,
                                  _vp, _vp]),
    'owmi_fwav_scatter': (_c_int, [_c_int, _vp, _c_ll, _c_ll, _c_ll, _c_int, _vp, _vp, _vp, _vp,
                                   _vp]),
    'owmi_fourier_fix': (_c_int, [_c_int, _vp, _c_ll, _c_ll, _c_ll, _c_int, _c_int, _vp]),
    'owmi_fourier_dot': (_c_int, [_c_int, _vp, _vp, _c_ll, _c_ll, _c_ll, _c_int, _vp, _vp]),
    'owmi_sht_plan_create': (_c_int, [ctypes.POINTER(_vp), _c_ll, _vp, _vp, _vp, _vp, _vp, _c_int]),
    'owmi_sht_plan_create_rings': (_c_int, [ctypes.POINTER(_vp), _c_ll, _vp, _vp, _vp, _vp, _vp,
                                            _c_int, _c_ll]),
    'owmi_sht_plan_destroy': (_c_int, [_vp]),
    'owmi_sht_plan_npix': (_c_ll, [_vp]),
    'owmi_sht_plan_tune': (_c_int, [_vp, _c_ll, _c_int, _c_int]),
    'owmi_sht_plan_tune_reset': (_c_int, [_vp]),
    'owmi_sht_synthesis': (_c_int, [_vp, _c_int, _c_int, _vp, _c_ll, _vp, _c_ll, _c_ll, _c_int, _vp]),
    'owmi_sht_adjoint': (_c_int, [_vp, _c_int, _c_int, _vp, _c_ll, _vp, _c_ll, _c_ll, _c_int, _vp]),
    'owmi_sht_pixel_sandwich': (_c_int, [_vp, _c_int, _c_int, _vp, _c_ll, _vp, _c_ll, _vp, _c_ll,
                                         _c_ll, _vp]),
    'owmi_sht_legendre': (_c_int, [_vp, _c_int, _c_int, _c_int, _vp, _c_ll, _c_ll, _vp]),
    'owmi_sht_sandwich_length': (_c_ll, [_vp]),
    'owmi_sht_sandwich_multipliers': (_c_int, [_vp, _c_int, _vp, _c_int, _vp, _vp]),
    'owmi_sht_pixel_sandwich_spectral': (_c_int, [_vp, _c_int, _c_int, _vp, _c_ll, _vp, _c_ll,
                                                  _vp, _c_ll, _vp]),
}

_lib = None


def symbols():
    '''Names of every function the C ABI exports.'''
    return sorted(_SIGNATURES)


def load(path=None):
    '''Load libowmi.so (once) and declare the signatures. Raises OSError if
    the library has not been built.'''
    global _lib
    if _lib is not None and path is None:
        return _lib
    path = path or LIB_PATH
    if not os.path.exists(path):
        raise OSError(f'owmi HIP library not found at {path}; build it with '
                      '`python -m optweight_amd.build`')
    lib = ctypes.CDLL(path)
    for name, (res, args) in _SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path == LIB_PATH:
        _lib = lib
    return lib


def check(status):
    if status != 0:
        msg = load().owmi_last_error().decode()
        raise RuntimeError(f'owmi error {status}: {msg}')


def call(name, *args):
    check(getattr(load(), name)(*args))


def stream_ptr(device=None):
    return ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)


def ptr(t):
    return ctypes.c_void_p(t.data_ptr())


def dtype_code(t):
    if t.dtype in (torch.float64, torch.complex128):
        return F64
    if t.dtype in (torch.float32, torch.complex64):
        return F32
    raise ValueError(f'unsupported dtype {t.dtype}')


def real_dtype(dtype):
    '''The real dtype of a (complex) alm dtype.'''
    return {torch.complex64: torch.float32, torch.complex128: torch.float64}.get(dtype, dtype)


def require_cuda(*ts):
    for t in ts:
        if not isinstance(t, torch.Tensor) or not t.is_cuda:
            raise ValueError('owmi kernels need CUDA (HIP) tensors')
        if not t.is_contiguous():
            raise ValueError('owmi kernels need contiguous tensors')


def np_ptr(a):
    return ctypes.c_void_p(a.ctypes.data)


def as_np(a, dtype):
    return np.ascontiguousarray(np.asarray(a, dtype=dtype))
