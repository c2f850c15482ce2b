'''
GPU parity of the spectral pixel sandwich (owmi_sht_pixel_sandwich_spectral:
the ring stage as a convolution over m with per-ring multipliers of the
pixel diagonal) against the Bluestein sandwich of the same plan and against
the oracle's Y^T D Y, through the C ABI.

Tolerances: float64 relative L2 1e-12 (oracle) and 1e-13 (Bluestein path);
float32 2e-5 (oracle) and 1e-5 (Bluestein path, both float32).
'''
import numpy as np
import pytest
import torch

from oracle import geometry as og, sht as osht, layout
from optweight_amd import sht, map_utils, alm_utils

pytestmark = pytest.mark.gpu


def rand_alm(lmax, npol, seed, ntrans=None):
    rng = np.random.default_rng(seed)
    n = layout.nelem(lmax)
    shape = (npol, n) if ntrans is None else (ntrans, npol, n)
    a = rng.normal(size=shape) + 1j * rng.normal(size=shape)
    ell, em = layout.lm_arrays(lmax)
    a[..., em == 0] = a[..., em == 0].real
    return a


def rel(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b))


def both(a, ai, m, spin, d):
    '''(spectral, Bluestein) sandwiches of device alm a with diagonal d.'''
    out_s = torch.zeros_like(a)
    out_b = torch.zeros_like(a)
    sht.pixel_sandwich(a, out_s, ai, m, spin, d, spectral=True)
    sht.pixel_sandwich(a, out_b, ai, m, spin, d, spectral=False)
    return out_s.cpu().numpy(), out_b.cpu().numpy()


def geometry(kind, lmax):
    '''(oracle geometry, MapInfo) of a test grid.'''
    if kind == 'gl':
        g = og.gauss_legendre(lmax + 1, 2 * lmax + 1)
        return g, map_utils.MapInfo(g.theta, g.weight, nphi=g.nphi)
    if kind == 'gl_odd':
        # odd ring count: the equator ring has no mirror (single-ring job)
        g = og.gauss_legendre(lmax + 1 if lmax % 2 == 0 else lmax + 2, 2 * lmax + 1)
        return g, map_utils.MapInfo(g.theta, g.weight, nphi=g.nphi)
    if kind == 'wide':
        # rings longer than 2 lmax + 1 (the Bluestein length exceeds M)
        g = og.gauss_legendre(lmax + 1, 2 * lmax + 6)
        return g, map_utils.MapInfo(g.theta, g.weight, nphi=g.nphi)
    if kind == 'shifted':
        g0 = og.gauss_legendre(lmax + 1, 2 * lmax + 1)
        phi0 = np.where(np.arange(g0.nrow) % 3 == 1, 0.02 + 0.013 * np.arange(g0.nrow), 0.0)
        g = og.Geom(g0.theta, g0.weight, g0.nphi, phi0=phi0)
        return g, map_utils.MapInfo(g0.theta, g0.weight, nphi=g0.nphi, phi0=phi0)
    if kind == 'reduced':
        # rings shorter than 2 lmax + 1 towards the poles (their m alias)
        g0 = og.gauss_legendre(lmax + 1, 2 * lmax + 1)
        nphi = np.maximum(8, np.round((2 * lmax + 1) * np.sin(g0.theta))).astype(np.int64)
        nphi[g0.nrow // 2] = 2 * lmax + 1
        g = og.Geom(g0.theta, g0.weight, nphi)
        return g, map_utils.MapInfo(g0.theta, g0.weight, nphi=nphi)
    raise ValueError(kind)


@pytest.mark.parametrize('kind', ['gl', 'gl_odd', 'wide', 'shifted', 'reduced'])
@pytest.mark.parametrize('dtype', ['float64', 'float32'])
def test_spectral_vs_oracle(dev, kind, dtype):
    '''Y^T D Y of a TQU alm on each grid kind against the oracle.'''
    lmax = 24
    print(f'spectral sandwich vs oracle: {kind} {dtype}', flush=True)
    g, m = geometry(kind, lmax)
    ai = alm_utils.AlmInfo(lmax)
    cdt = torch.complex128 if dtype == 'float64' else torch.complex64
    rdt = torch.float64 if dtype == 'float64' else torch.float32
    tol = 1e-12 if dtype == 'float64' else 2e-5
    plan = sht.get_plan(m, lmax, dev)
    assert plan._lib.owmi_sht_sandwich_length(plan.handle) == 128
    a = rand_alm(lmax, 3, 3)
    rng = np.random.default_rng(4)
    d = rng.uniform(0.5, 2, size=(3, g.npix))
    mp = np.zeros((3, g.npix))
    osht.alm2map(a, mp, lmax, g, [0, 2])
    ref = np.zeros_like(a)
    osht.map2alm(mp * d, ref, g, lmax, [0, 2], adjoint=True)
    at = torch.as_tensor(a, device=dev).to(cdt)
    out = torch.zeros_like(at)
    sht.pixel_sandwich(at, out, ai, m, [0, 2], torch.as_tensor(d, device=dev).to(rdt),
                       spectral=True)
    assert rel(out.cpu().numpy(), ref) < tol


@pytest.mark.parametrize('lmax', [17, 64, 200])
@pytest.mark.parametrize('ntrans', [1, 3, 8, 13])
@pytest.mark.parametrize('dtype', ['float64', 'float32'])
def test_spectral_vs_bluestein_batches(dev, lmax, ntrans, dtype):
    '''Batches around the 8 transforms per workgroup (partial last groups),
    spins 0, 2 and TQU: equal to the Bluestein sandwich of the same plan.'''
    print(f'spectral vs bluestein lmax={lmax} ntrans={ntrans} {dtype}', flush=True)
    g, m = geometry('gl_odd', lmax)
    ai = alm_utils.AlmInfo(lmax)
    cdt = torch.complex128 if dtype == 'float64' else torch.complex64
    rdt = torch.float64 if dtype == 'float64' else torch.float32
    tol = 1e-13 if dtype == 'float64' else 1e-5
    rng = np.random.default_rng(lmax + ntrans)
    for spin, npol in ((0, 1), (2, 2), ([0, 2], 3)):
        a = torch.as_tensor(rand_alm(lmax, npol, 5, ntrans), device=dev).to(cdt)
        d = torch.as_tensor(rng.uniform(0, 2, size=(npol, m.npix)), device=dev).to(rdt)
        s, b = both(a, ai, m, spin, d)
        assert rel(s, b) < tol, (spin, rel(s, b))


@pytest.mark.parametrize('lmax,nrings', [(2047, 24), (4095, 12)])
@pytest.mark.parametrize('dtype', ['float64', 'float32'])
def test_spectral_long_rings(dev, lmax, nrings, dtype):
    '''FFT lengths 8192 and 16384 (512- and 1024-thread workgroups; float64
    at 16384 keeps its FFT buffer in global memory) on a few Gauss-Legendre
    rings of the full band limit: equal to the Bluestein sandwich.'''
    print(f'spectral long rings lmax={lmax} {dtype}', flush=True)
    g = og.gauss_legendre(nrings, 2 * lmax + 1)
    m = map_utils.MapInfo(g.theta, g.weight, nphi=g.nphi)
    ai = alm_utils.AlmInfo(lmax)
    cdt = torch.complex128 if dtype == 'float64' else torch.complex64
    rdt = torch.float64 if dtype == 'float64' else torch.float32
    tol = 1e-12 if dtype == 'float64' else 2e-5
    rng = np.random.default_rng(9)
    a = torch.as_tensor(rand_alm(lmax, 3, 7, 2), device=dev).to(cdt)
    d = torch.as_tensor(rng.uniform(0.5, 2, size=(3, m.npix)), device=dev).to(rdt)
    s, b = both(a, ai, m, [0, 2], d)
    assert rel(s, b) < tol


def test_short_rings_fall_back(dev):
    '''When every ring is shorter than 2 lmax + 1, the plan's FFT length is
    below 4 lmax + 1: no spectral length, and the sandwich runs the
    Bluestein path (equal to the oracle).'''
    lmax = 24
    g0 = og.gauss_legendre(lmax + 1, lmax)
    m = map_utils.MapInfo(g0.theta, g0.weight, nphi=g0.nphi)
    plan = sht.get_plan(m, lmax, dev)
    assert plan._lib.owmi_sht_sandwich_length(plan.handle) == 0
    d = torch.ones((3, m.npix), dtype=torch.float64, device=dev)
    assert sht.sandwich_multipliers(d, m, lmax) is None
    a = rand_alm(lmax, 3, 2)
    mp = np.zeros((3, g0.npix))
    osht.alm2map(a, mp, lmax, g0, [0, 2])
    ref = np.zeros_like(a)
    osht.map2alm(mp, ref, g0, lmax, [0, 2], adjoint=True)
    out = torch.zeros_like(torch.as_tensor(a, device=dev))
    sht.pixel_sandwich(torch.as_tensor(a, device=dev), out, alm_utils.AlmInfo(lmax), m, [0, 2], d)
    assert rel(out.cpu().numpy(), ref) < 1e-12


def test_multipliers_follow_the_diagonal(dev):
    '''The multipliers are cached on the diagonal tensor: reused while it is
    unchanged, recomputed after an in-place change (version counter), and
    a per-transform diagonal runs the Bluestein path.'''
    lmax = 20
    g, m = geometry('gl', lmax)
    ai = alm_utils.AlmInfo(lmax)
    a = torch.as_tensor(rand_alm(lmax, 3, 1, 2), device=dev)
    d = torch.full((3, m.npix), 1.5, dtype=torch.float64, device=dev)
    m1 = sht.sandwich_multipliers(d, m, lmax)
    assert sht.sandwich_multipliers(d, m, lmax) is m1
    s1, b1 = both(a, ai, m, [0, 2], d)
    assert rel(s1, b1) < 1e-13
    d.mul_(torch.linspace(0.1, 2, m.npix, dtype=torch.float64, device=dev))
    m2 = sht.sandwich_multipliers(d, m, lmax)
    assert m2 is not m1
    s2, b2 = both(a, ai, m, [0, 2], d)
    assert rel(s2, b2) < 1e-13 and rel(s2, s1) > 1e-3
    # a constant diagonal c on rings of N = 2 lmax + 1 pixels: the kernel is
    # K(d) = c N at d = 0 and 0 for 0 < |d| <= 2 lmax, so every multiplier
    # is c N / M
    dc = torch.full((1, m.npix), 0.75, dtype=torch.float64, device=dev)
    mc = sht.sandwich_multipliers(dc, m, lmax).cpu().numpy()
    assert np.abs(mc - 0.75 * (2 * lmax + 1) / 128).max() < 1e-14
    # per-transform diagonal: Bluestein path, same values as two shared calls
    dd = torch.stack([d, 2 * d])
    out = torch.zeros_like(a)
    sht.pixel_sandwich(a, out, ai, m, [0, 2], dd)
    o0 = torch.zeros_like(a[:1])
    sht.pixel_sandwich(a[:1].contiguous(), o0, ai, m, [0, 2], d)
    assert rel(out[0].cpu().numpy(), o0[0].cpu().numpy()) < 1e-13


@pytest.mark.parametrize('dtype', ['float64', 'float32'])
@pytest.mark.parametrize('spectral', [True, False])
def test_sandwich_skips_vanishing_rings(dev, dtype, spectral):
    '''A diagonal that vanishes on a band of rings (both rings of those
    pairs) and on single rings of two other pairs: the sandwich restricted
    to the rings where it does not vanish (map_utils.ring_support) equals
    the sandwich over every ring (float64 1e-13, float32 1e-5), spectral and
    Bluestein, and the float64 result equals the oracle's Y^T D Y.'''
    lmax = 30
    print(f'sandwich ring skip {dtype} spectral={spectral}', flush=True)
    g, m = geometry('gl', lmax)
    ai = alm_utils.AlmInfo(lmax)
    cdt = torch.complex128 if dtype == 'float64' else torch.complex64
    rdt = torch.float64 if dtype == 'float64' else torch.float32
    tol = 1e-13 if dtype == 'float64' else 1e-5
    rng = np.random.default_rng(12)
    d = rng.uniform(0.5, 2, size=(3, g.npix))
    off = np.asarray(m.offsets, dtype=np.int64)
    nphi = np.asarray(m.nphi, dtype=np.int64)
    nr = m.nrow
    zero = list(range(nr // 2 - 3, nr // 2 + 3)) + [2, nr - 6]
    for r in zero:
        d[:, off[r]:off[r] + nphi[r]] = 0
    rings = map_utils.ring_support(d, m)
    assert rings.size == nr - len(zero)
    a = rand_alm(lmax, 3, 8, 4)
    at = torch.as_tensor(a, device=dev).to(cdt)
    dt = torch.as_tensor(d, device=dev).to(rdt)
    full = torch.zeros_like(at)
    part = torch.zeros_like(at)
    sht.pixel_sandwich(at, full, ai, m, [0, 2], dt, spectral=spectral)
    sht.pixel_sandwich(at, part, ai, m, [0, 2], dt, spectral=spectral, rings=rings)
    assert rel(part.cpu().numpy(), full.cpu().numpy()) < tol
    if dtype == 'float64':
        mp = np.zeros((3, g.npix))
        osht.alm2map(a[1], mp, lmax, g, [0, 2])
        ref = np.zeros_like(a[1])
        osht.map2alm(mp * d, ref, g, lmax, [0, 2], adjoint=True)
        assert rel(part[1].cpu().numpy(), ref) < 1e-12
