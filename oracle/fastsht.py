'''
Oracle (test infrastructure): the spherical harmonic transforms of
``oracle.sht`` (optweight sht.py:10-178: ducc0 synthesis / adjoint_synthesis
on a ring geometry) with the Legendre stage in float64 C with OpenMP
(``oracle/legendre_omp.c``, built into ``oracle/_lib/liboracle_sht.so`` by
``oracle/Makefile``) and the ring Fourier sums batched per ring length with
scipy.fft. Same signatures and results as ``oracle.sht.alm2map / map2alm``
(pinned against them by tests/test_oracle_fast.py); it makes full-size
problems (lmax 1023 and above) run in seconds for the CPU baseline and the
golden fixtures. ``install()`` makes ``oracle.sht`` itself use it.
'''
import ctypes
import os

import numpy as np
import scipy.fft

from oracle import geometry, layout, legendre, sht as osht

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, '_lib', 'liboracle_sht.so')

_lib = None
_seed_cache = {}


def threads():
    n = os.environ.get('ORACLE_THREADS')
    return int(n) if n else len(os.sched_getaffinity(0))


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise OSError(f'{LIB_PATH} not built (make -C oracle)')
        lib = ctypes.CDLL(LIB_PATH)
        vp, ci = ctypes.c_void_p, ctypes.c_int
        for name in ('ol_legendre_synth', 'ol_legendre_adjoint'):
            fn = getattr(lib, name)
            fn.restype = None
            fn.argtypes = [ci, ci, ci, vp, vp, vp, vp, vp, vp, ci, vp, ci]
        _lib = lib
    return _lib


def available():
    return os.path.exists(LIB_PATH)


def _p(a):
    return ctypes.c_void_p(a.ctypes.data)


class _Pairs:
    '''Mirror pairs of a geometry with their float64 x = cos(theta_north) and
    the seed tables per spin (cached per geometry and lmax).'''

    def __init__(self, geom, lmax):
        north, south = geometry.mirror_pairs(geom)
        th_eff, x = geometry.ring_x(geom)
        self.north = np.ascontiguousarray(north, dtype=np.int32)
        self.south = np.ascontiguousarray(south, dtype=np.int32)
        self.x = np.ascontiguousarray(np.asarray(x[north], dtype=np.float64))
        self.theta = np.asarray(th_eff[north], dtype=np.longdouble)
        self.lmax = lmax
        self.seeds = {}

    def seed(self, s):
        if s not in self.seeds:
            mant, expo = legendre.seed_tables(self.lmax, self.theta, s)
            self.seeds[s] = (np.ascontiguousarray(mant), np.ascontiguousarray(expo))
        return self.seeds[s]


def _pairs(geom, lmax):
    key = (lmax, geom.theta.tobytes(), np.asarray(geom.nphi).tobytes())
    p = _seed_cache.get(key)
    if p is None:
        if len(_seed_cache) > 32:
            _seed_cache.clear()
        p = _seed_cache[key] = _Pairs(geom, lmax)
    return p


def _ring_groups(geom):
    '''Rings grouped by length: {nphi: ring indices}.'''
    groups = {}
    for r, n in enumerate(np.asarray(geom.nphi)):
        groups.setdefault(int(n), []).append(r)
    return {n: np.asarray(rs) for n, rs in groups.items()}


def _ring_synth(F, geom, mmax):
    '''F (ncomp, nrow, mmax+1) complex -> (ncomp, npix), as
    oracle.sht._ring_synth (frequency m folded onto bin m mod nphi).'''
    ncomp = F.shape[0]
    out = np.zeros((ncomp, geom.npix))
    ms = np.arange(mmax + 1)
    for n, rs in _ring_groups(geom).items():
        C = F[:, rs, :]
        if np.any(geom.phi0[rs] != 0):
            C = C * np.exp(1j * ms[None, :] * geom.phi0[rs][:, None])
        if n > 2 * mmax:
            # every m has its own bin: a real inverse FFT (the imaginary
            # part of bin 0 is ignored, as Re C_0 above)
            H = np.zeros((ncomp, rs.size, n // 2 + 1), dtype=np.complex128)
            H[..., :mmax + 1] = C
            vals = scipy.fft.irfft(H, n=n, axis=-1, workers=threads()) * n
        else:
            # fold: bin j gets C_m for m = j mod n (m >= 1), conj(C_m) for
            # -m = j mod n, and Re C_0 at j = 0
            k = mmax // n + 1
            buf = np.zeros((ncomp, rs.size, k * n), dtype=np.complex128)
            buf[..., 1:mmax + 1] = C[..., 1:]
            P = buf.reshape(ncomp, rs.size, k, n).sum(axis=2)
            Z = P + np.conj(P[..., (-np.arange(n)) % n])
            Z[..., 0] += C[..., 0].real
            vals = scipy.fft.ifft(Z, axis=-1, workers=threads()).real * n
        if rs.size == rs[-1] - rs[0] + 1 and np.all(np.diff(geom.offsets[rs]) == n):
            o = geom.offsets[rs[0]]
            out[:, o:o + rs.size * n] = vals.reshape(ncomp, -1)
        else:
            idx = geom.offsets[rs][:, None] + np.arange(n)[None, :]
            out[:, idx] = vals
    return out


def _ring_adjoint(imap, geom, mmax):
    ncomp = imap.shape[0]
    G = np.zeros((ncomp, geom.nrow, mmax + 1), dtype=np.complex128)
    ms = np.arange(mmax + 1)
    for n, rs in _ring_groups(geom).items():
        if rs.size == rs[-1] - rs[0] + 1 and np.all(np.diff(geom.offsets[rs]) == n):
            o = geom.offsets[rs[0]]
            vals = np.asarray(imap[:, o:o + rs.size * n], dtype=np.float64).reshape(
                ncomp, rs.size, n)
        else:
            idx = geom.offsets[rs][:, None] + np.arange(n)[None, :]
            vals = np.asarray(imap[:, idx], dtype=np.float64)
        if n > 2 * mmax:
            X = scipy.fft.rfft(vals, axis=-1, workers=threads())[..., :mmax + 1]
        else:
            X = scipy.fft.fft(vals, axis=-1, workers=threads())[..., ms % n]
        if np.any(geom.phi0[rs] != 0):
            X = X * np.exp(-1j * ms[None, :] * geom.phi0[rs][:, None])
        G[:, rs, :] = X
    return G


def _synth_group(alm, geom, lmax, s):
    pr = _pairs(geom, lmax)
    mant, expo = pr.seed(s)
    a = np.ascontiguousarray(alm, dtype=np.complex128)
    F = np.zeros((a.shape[0], geom.nrow, lmax + 1), dtype=np.complex128)
    load().ol_legendre_synth(lmax, s, pr.north.size, _p(pr.north), _p(pr.south), _p(pr.x),
                             _p(mant), _p(expo), _p(a), geom.nrow, _p(F), threads())
    return _ring_synth(F, geom, lmax)


def _adjoint_group(imap, geom, lmax, s):
    pr = _pairs(geom, lmax)
    mant, expo = pr.seed(s)
    G = np.ascontiguousarray(_ring_adjoint(imap, geom, lmax))
    out = np.zeros((imap.shape[0], layout.nelem(lmax)), dtype=np.complex128)
    load().ol_legendre_adjoint(lmax, s, pr.north.size, _p(pr.north), _p(pr.south), _p(pr.x),
                               _p(mant), _p(expo), _p(G), geom.nrow, _p(out), threads())
    return out


def alm2map(alm, omap, lmax, geom, spin, adjoint=False):
    '''Y (adjoint=False) or W Y (adjoint=True); writes omap in place.'''
    osht._check(alm.shape[-1], omap.shape[-1], lmax, geom, spin)
    alm3 = alm.reshape((-1,) + alm.shape[-2:]) if alm.ndim >= 2 else alm.reshape(1, 1, -1)
    om3 = omap.reshape((-1,) + omap.shape[-2:]) if omap.ndim >= 2 else omap.reshape(1, 1, -1)
    if alm3.shape[:-1] != om3.shape[:-1]:
        raise ValueError('Mismatch leading dimension of map and alm')
    npol = om3.shape[-2]
    for t in range(om3.shape[0]):
        for s, i1, i2 in osht.spin_groups(spin, npol):
            m = _synth_group(alm3[t, i1:i2], geom, lmax, s)
            if adjoint:
                osht.ringweight(m, geom)
            om3[t, i1:i2] = m
    return omap


def map2alm(imap, alm, geom, lmax, spin, adjoint=False):
    '''Y^T W (adjoint=False) or Y^T (adjoint=True); writes alm in place.'''
    osht._check(alm.shape[-1], imap.shape[-1], lmax, geom, spin)
    im3 = imap.reshape((-1,) + imap.shape[-2:]) if imap.ndim >= 2 else imap.reshape(1, 1, -1)
    alm3 = alm.reshape((-1,) + alm.shape[-2:]) if alm.ndim >= 2 else alm.reshape(1, 1, -1)
    if alm3.shape[:-1] != im3.shape[:-1]:
        raise ValueError('Mismatch leading dimension of map and alm')
    npol = im3.shape[-2]
    for t in range(im3.shape[0]):
        for s, i1, i2 in osht.spin_groups(spin, npol):
            m = np.array(im3[t, i1:i2], dtype=np.float64)
            if not adjoint:
                osht.ringweight(m, geom)
            alm3[t, i1:i2] = _adjoint_group(m, geom, lmax, s)
    return alm


_saved = None


def install():
    '''Route oracle.sht.alm2map / map2alm (and so every oracle operator)
    through this backend; ``uninstall()`` restores the numpy path.'''
    global _saved
    load()
    if _saved is None:
        _saved = (osht.alm2map, osht.map2alm)
        osht.alm2map, osht.map2alm = alm2map, map2alm


def uninstall():
    global _saved
    if _saved is not None:
        osht.alm2map, osht.map2alm = _saved
        _saved = None
