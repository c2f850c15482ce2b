'''
Spherical harmonic transforms on device tensors, the interface of
optweight's sht module: map2alm (sht.py:10-93), alm2map (sht.py:95-178),
default_spin (sht.py:180-215). The transforms run on the HIP Legendre and
ring-FFT kernels through owmi_sht_* (one plan per geometry and lmax, cached
on the MapInfo object).

    alm2map(alm, omap, ainfo, minfo, spin, adjoint=False)   Y   (W Y if adjoint)
    map2alm(imap, alm, minfo, ainfo, spin, adjoint=False)   Y^T W   (Y^T if adjoint)

Both take an optional ``rings`` (sorted ring indices of minfo): the
transform then runs over those rings only. Synthesis leaves the pixels of
other rings untouched; the adjoint reads only the given rings, which is the
full adjoint for a map that vanishes on all other rings (masked operators).
'''
import ctypes
import os
import weakref

import numpy as np
import torch

from optweight_amd import _lib


# Useful Legendre flops issued through this module, by dtype:
# F = N_lm * N_theta * (2 n0 + 8 n2) per transform direction (N_theta: rings
# of the plan or of its ring subset; n0 spin-0 components, n2 spin-2 pairs),
# counted on the host when a transform is launched.
LEGENDRE_FLOPS = {'float32': 0.0, 'float64': 0.0}


def reset_legendre_flops():
    for k in LEGENDRE_FLOPS:
        LEGENDRE_FLOPS[k] = 0.0


def _count(plan, spin, ncomp, ntrans, dtype, directions=1):
    per = 2 * ncomp if spin == 0 else 4 * ncomp
    key = 'float64' if dtype in (torch.float64, torch.complex128) else 'float32'
    LEGENDRE_FLOPS[key] += float(_nelem(plan.lmax)) * plan.nrings * per * ntrans * directions


class Plan:
    '''Owns one owmi_sht_plan (geometry + lmax on one device).'''

    def __init__(self, minfo, lmax, device, rings=None):
        self.lmax = lmax
        self.npix = minfo.npix
        self.nrings = int(minfo.nrow if rings is None else np.asarray(rings).size)
        self.device = device
        self._lib = _lib.load()
        self.handle = None
        h = ctypes.c_void_p()
        sel = slice(None) if rings is None else np.asarray(rings, dtype=np.int64)
        theta = _lib.as_np(np.asarray(minfo.theta)[sel], np.float64)
        weight = _lib.as_np(np.asarray(minfo.weight)[sel], np.float64)
        nphi = _lib.as_np(np.asarray(minfo.nphi)[sel], np.int64)
        offsets = _lib.as_np(np.asarray(minfo.offsets)[sel], np.int64)
        phi0 = _lib.as_np(np.asarray(minfo.phi0)[sel], np.float64)
        if np.any(minfo.stride != 1):
            raise ValueError('only unit pixel stride is supported')
        if theta.size == 0:
            raise ValueError('empty ring selection')
        with torch.cuda.device(device):
            if rings is None:
                _lib.check(self._lib.owmi_sht_plan_create(
                    ctypes.byref(h), theta.size, _lib.np_ptr(theta), _lib.np_ptr(weight),
                    _lib.np_ptr(nphi), _lib.np_ptr(offsets), _lib.np_ptr(phi0), lmax))
            else:
                _lib.check(self._lib.owmi_sht_plan_create_rings(
                    ctypes.byref(h), theta.size, _lib.np_ptr(theta), _lib.np_ptr(weight),
                    _lib.np_ptr(nphi), _lib.np_ptr(offsets), _lib.np_ptr(phi0), lmax,
                    int(minfo.npix)))
        self.handle = h

    def __del__(self):
        try:
            if self.handle:
                if torch.cuda.is_current_stream_capturing():
                    # freeing device memory would invalidate a HIP-graph
                    # recording in progress: destroy after it
                    _DEFERRED.append((self._lib, self.handle))
                else:
                    self._lib.owmi_sht_plan_destroy(self.handle)
                self.handle = None
        except Exception:
            pass


# plans released during a HIP-graph recording, destroyed by the next
# get_plan outside one
_DEFERRED = []


def release_deferred():
    '''Destroy plans whose release was deferred by a graph recording.'''
    if _DEFERRED and not torch.cuda.is_current_stream_capturing():
        while _DEFERRED:
            lib, handle = _DEFERRED.pop()
            lib.owmi_sht_plan_destroy(handle)


def get_plan(minfo, lmax, device, rings=None):
    '''Cached plan of (minfo, lmax, device), optionally restricted to the
    sorted ring indices ``rings``.'''
    device = torch.device(device)
    if device.index is None:
        device = torch.device('cuda', torch.cuda.current_device())
    release_deferred()
    cache = minfo.__dict__.setdefault('_plans', {})
    if rings is not None:
        rings = np.asarray(rings, dtype=np.int64)
        if rings.size == minfo.nrow:
            rings = None
    key = (str(device), int(lmax), None if rings is None else rings.tobytes())
    if key not in cache:
        plan = Plan(minfo, int(lmax), device, rings=rings)
        # launch choices an autotuner cached for this geometry
        from optweight_amd import autotune
        autotune.apply_cached(plan, minfo, int(lmax), rings)
        cache[key] = plan
    return cache[key]


def default_spin(shape):
    '''npol 1 -> 0, 2 -> 2, 3 -> [0, 2].'''
    if len(shape) > 2:
        raise ValueError(f'Cannot determine default spin for alm/map with more than '
                         f'one leading dimension. Got shape : {shape}')
    if len(shape) == 1 or shape[0] == 1:
        return 0
    if shape[0] == 2:
        return 2
    if shape[0] == 3:
        return [0, 2]
    raise ValueError(f'Cannot determine default spin value for shape : {shape}')


def spin_groups(spin, npol):
    '''(s, i1, i2): spin s acts on components i1:i2 (spin 0 takes one
    component, spin 2 two), walking the spin list over npol components.'''
    spins = np.atleast_1d(spin).tolist()
    out = []
    i = 0
    k = 0
    while i < npol:
        s = int(spins[min(k, len(spins) - 1)])
        if s not in (0, 2):
            raise ValueError(f'spin {s} not supported (0 or 2)')
        n = 1 if s == 0 else 2
        if i + n > npol:
            raise ValueError(f'spin {spin} incompatible with npol {npol}')
        out.append((s, i, i + n))
        i += n
        k += 1
    return out


def _lmax_of(ainfo):
    return ainfo if isinstance(ainfo, (int, np.integer)) else ainfo.lmax


def _nelem(lmax):
    return (lmax * lmax + 3 * lmax) // 2 + 1


def _check(alm, imap, ainfo, minfo, spin):
    lmax = _lmax_of(ainfo)
    if np.asarray(spin).max() > lmax:
        raise ValueError('Spin exceeds lmax')
    if imap.shape[-1] != minfo.npix:
        raise ValueError(f'Wrong map size, got {imap.shape[-1]} expected {minfo.npix}')
    if alm.shape[-1] != _nelem(lmax):
        raise ValueError(f'Wrong alm size, got {alm.shape[-1]} expected {_nelem(lmax)}')
    _lib.require_cuda(alm, imap)
    a3 = alm.reshape((-1,) + tuple(alm.shape[-2:])) if alm.ndim >= 2 else alm.reshape(1, 1, -1)
    m3 = imap.reshape((-1,) + tuple(imap.shape[-2:])) if imap.ndim >= 2 else imap.reshape(1, 1, -1)
    if a3.shape[:-1] != m3.shape[:-1]:
        raise ValueError(f'Mismatch leading dimension of map and alm :'
                         f'{tuple(imap.shape[:-1])} and {tuple(alm.shape[:-1])}')
    real = torch.float64 if alm.dtype == torch.complex128 else torch.float32
    if alm.dtype not in (torch.complex64, torch.complex128) or imap.dtype != real:
        raise ValueError(f'dtype mismatch: alm {alm.dtype}, map {imap.dtype}')
    return lmax, a3, m3


def alm2map(alm, omap, ainfo, minfo, spin, adjoint=False, rings=None):
    '''Synthesis Y (or W Y if adjoint) of alm (..., npol, nelem) into omap
    (..., npol, npix), overwritten (on ``rings`` only, if given).'''
    lmax, a3, m3 = _check(alm, omap, ainfo, minfo, spin)
    plan = get_plan(minfo, lmax, alm.device, rings)
    npol = m3.shape[-2]
    code = _lib.dtype_code(omap)
    st = _lib.stream_ptr(alm.device)
    for s, i1, i2 in spin_groups(spin, npol):
        _lib.check(plan._lib.owmi_sht_synthesis(
            plan.handle, code, s, _lib.ptr(a3[:, i1]), npol * a3.shape[-1], _lib.ptr(m3[:, i1]),
            npol * m3.shape[-1], m3.shape[0], int(bool(adjoint)), st))
        _count(plan, s, i2 - i1, m3.shape[0], alm.dtype)
    return omap


def map2alm(imap, alm, minfo, ainfo, spin, adjoint=False, rings=None):
    '''Adjoint synthesis Y^T W (or Y^T if adjoint) of imap (..., npol, npix)
    into alm (..., npol, nelem), overwritten (reading ``rings`` only, if
    given).'''
    lmax, a3, m3 = _check(alm, imap, ainfo, minfo, spin)
    plan = get_plan(minfo, lmax, alm.device, rings)
    npol = m3.shape[-2]
    code = _lib.dtype_code(imap)
    st = _lib.stream_ptr(alm.device)
    for s, i1, i2 in spin_groups(spin, npol):
        _lib.check(plan._lib.owmi_sht_adjoint(
            plan.handle, code, s, _lib.ptr(m3[:, i1]), npol * m3.shape[-1], _lib.ptr(a3[:, i1]),
            npol * a3.shape[-1], m3.shape[0], int(not adjoint), st))
        _count(plan, s, i2 - i1, m3.shape[0], alm.dtype)
    return alm


def spectral_enabled():
    '''The spectral sandwich is used unless OWMI_SANDWICH=bluestein.'''
    return os.environ.get('OWMI_SANDWICH', 'spectral') != 'bluestein'


def sandwich_multipliers(pixdiag, minfo, lmax, rings=None):
    '''Per-ring multipliers (npol, nrings, M) of the spectral sandwich for a
    pixel diagonal (npol, npix) shared by every transform, or None where the
    plan's ring FFTs do not cover the length M (owmi_sht_sandwich_length);
    rings: the plan restricted to these rings (nrings of them). They are
    computed once (float64 FFTs, stored in the diagonal's dtype) and kept on
    the diagonal tensor with its version counter, so an in-place change of
    the diagonal recomputes them.'''
    plan = get_plan(minfo, lmax, pixdiag.device, rings=rings)
    M = int(plan._lib.owmi_sht_sandwich_length(plan.handle))
    if M == 0:
        return None
    nr = minfo.nrow if rings is None else len(rings)
    rkey = None if rings is None or nr == minfo.nrow else np.asarray(rings, np.int64).tobytes()
    key = (int(lmax), M, pixdiag.dtype, str(pixdiag.device), rkey)
    cached = getattr(pixdiag, '_owmi_mult', None)
    if (cached is not None and cached[0] == pixdiag._version and cached[1] == key
            and cached[2]() is minfo):
        return cached[3]
    d = pixdiag.reshape((-1, minfo.npix))
    if not d.is_contiguous():
        raise ValueError('pixdiag must be contiguous')
    mult = torch.empty((d.shape[0], nr, M), dtype=d.dtype, device=d.device)
    _lib.check(plan._lib.owmi_sht_sandwich_multipliers(
        plan.handle, _lib.dtype_code(d), _lib.ptr(d), d.shape[0], _lib.ptr(mult),
        _lib.stream_ptr(d.device)))
    pixdiag._owmi_mult = (pixdiag._version, key, weakref.ref(minfo), mult)
    return mult


def pixel_sandwich(alm, oalm, ainfo, minfo, spin, pixdiag, spectral=None, rings=None):
    '''oalm = Y^T D Y alm with D = pixdiag (npol, npix) per component; the
    map stays on chip between the two transforms. A diagonal shared by all
    transforms runs as the spectral sandwich (per-ring multipliers from
    sandwich_multipliers, two FFTs of length M per ring pair) unless
    spectral is False or OWMI_SANDWICH=bluestein; a per-transform diagonal,
    or a plan without the spectral length, runs the Bluestein sandwich.
    rings: sorted ring indices outside which D vanishes (ring_support of
    the diagonal); the others contribute nothing and are skipped.'''
    lmax = _lmax_of(ainfo)
    _lib.require_cuda(alm, oalm, pixdiag)
    if tuple(alm.shape) != tuple(oalm.shape) or alm.dtype != oalm.dtype:
        raise ValueError('alm and oalm must match')
    a3 = alm.reshape((-1,) + tuple(alm.shape[-2:]))
    o3 = oalm.reshape((-1,) + tuple(oalm.shape[-2:]))
    npol = a3.shape[-2]
    if tuple(pixdiag.shape[-2:]) != (npol, minfo.npix):
        raise ValueError(f'pixdiag shape {tuple(pixdiag.shape)} != {(npol, minfo.npix)}')
    p3 = pixdiag.reshape((-1, npol, minfo.npix))
    pix_ts = 0 if p3.shape[0] == 1 else npol * minfo.npix
    plan = get_plan(minfo, lmax, alm.device, rings=rings)
    code = _lib.dtype_code(alm)
    st = _lib.stream_ptr(alm.device)
    if spectral is None:
        spectral = spectral_enabled()
    mult = None
    if spectral and pix_ts == 0 and pixdiag.dtype == _lib.real_dtype(alm.dtype):
        mult = sandwich_multipliers(pixdiag, minfo, lmax, rings=rings)
    if mult is not None:
        for s, i1, i2 in spin_groups(spin, npol):
            _lib.check(plan._lib.owmi_sht_pixel_sandwich_spectral(
                plan.handle, code, s, _lib.ptr(a3[:, i1]), npol * a3.shape[-1],
                _lib.ptr(o3[:, i1]), npol * o3.shape[-1], _lib.ptr(mult[i1]), a3.shape[0], st))
            _count(plan, s, i2 - i1, a3.shape[0], alm.dtype, directions=2)
        return oalm
    for s, i1, i2 in spin_groups(spin, npol):
        _lib.check(plan._lib.owmi_sht_pixel_sandwich(
            plan.handle, code, s, _lib.ptr(a3[:, i1]), npol * a3.shape[-1], _lib.ptr(o3[:, i1]),
            npol * o3.shape[-1], _lib.ptr(p3[:, i1]), pix_ts, a3.shape[0], st))
        _count(plan, s, i2 - i1, a3.shape[0], alm.dtype, directions=2)
    return oalm
