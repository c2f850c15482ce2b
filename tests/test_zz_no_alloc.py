'''
owmi_alm_dot_batch_noalloc (the allocation-free batched real-field dot,
partial sums in the library's fixed device array) against
owmi_alm_dot_batch, through the C ABI: equal bit for bit, on two calls in a
row (the second on a smaller input, so stale partials of the first call
would show), in both precisions; nbatch above the array's 1024 partial
sums is rejected before any device work.
'''
import ctypes

import numpy as np
import pytest
import torch

from optweight_amd import _lib

pytestmark = pytest.mark.gpu


def rand_alm(shape, dtype, seed, dev):
    rng = np.random.default_rng(seed)
    a = rng.normal(size=shape) + 1j * rng.normal(size=shape)
    return torch.as_tensor(a, device=dev).to(dtype).contiguous()


def dots(name, a, b, lmax, nbatch):
    out = torch.full((nbatch,), float('nan'), dtype=torch.float64, device=a.device)
    nrows = a.numel() // (a.shape[-1] * nbatch)
    _lib.call(name, _lib.dtype_code(a), _lib.ptr(a), _lib.ptr(b), lmax, nrows, nbatch,
              _lib.ptr(out), _lib.stream_ptr(a.device))
    return out


@pytest.mark.parametrize('dtype', [torch.complex128, torch.complex64])
def test_noalloc_dot_equals_allocating_dot(dev, dtype):
    cases = [(96, 24, 3), (40, 5, 1)]      # (lmax, nbatch, ncomp): second call smaller
    for k, (lmax, nbatch, ncomp) in enumerate(cases):
        nelem = (lmax + 1) * (lmax + 2) // 2
        a = rand_alm((nbatch, ncomp, nelem), dtype, 10 + k, dev)
        b = rand_alm((nbatch, ncomp, nelem), dtype, 20 + k, dev)
        ref = dots('owmi_alm_dot_batch', a, b, lmax, nbatch)
        got = dots('owmi_alm_dot_batch_noalloc', a, b, lmax, nbatch)
        torch.cuda.synchronize()
        assert torch.equal(got, ref), (k, got, ref)
        # and both are the real-field product (m = 0 once, m > 0 twice)
        an = a.cpu().numpy().astype(np.complex128)
        bn = b.cpu().numpy().astype(np.complex128)
        prod = np.real(an * np.conj(bn))
        exp = 2 * prod.sum(axis=(1, 2)) - prod[..., :lmax + 1].sum(axis=(1, 2))
        tol = 1e-12 if dtype == torch.complex128 else 1e-5
        np.testing.assert_allclose(got.cpu().numpy(), exp, rtol=tol, atol=0)


def test_noalloc_dot_rejects_large_batch(dev):
    lib = _lib.load()
    out = torch.zeros(2048, dtype=torch.float64, device=dev)
    a = torch.zeros((2048, 1, 3), dtype=torch.complex128, device=dev)
    st = _lib.stream_ptr(dev)
    rc = lib.owmi_alm_dot_batch_noalloc(_lib.F64, _lib.ptr(a), _lib.ptr(a), 1, 1, 2048,
                                        _lib.ptr(out), st)
    assert rc != 0
    assert b'1024' in lib.owmi_last_error()
    # the failed call launched nothing; a valid call still works after it
    rc = lib.owmi_alm_dot_batch_noalloc(_lib.F64, _lib.ptr(a), _lib.ptr(a), 1, 1, 1024,
                                        _lib.ptr(out), st)
    assert rc == 0
    torch.cuda.synchronize()
    assert torch.count_nonzero(out[:1024]).item() == 0
