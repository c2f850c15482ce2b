'''
GPU parity of the HIP spherical harmonic transforms against the oracle
(float64 restatement, longdouble Legendre recurrence), through the C ABI.

Tolerances: float64 transforms agree with the oracle to relative L2 1e-12
(single application); float32 transforms to 2e-5.
'''
import numpy as np
import pytest
import torch

from oracle import geometry as og, sht as osht, layout
from optweight_amd import sht, map_utils, alm_utils

pytestmark = pytest.mark.gpu


def rand_alm(lmax, npol, seed, spin2_zero=True):
    rng = np.random.default_rng(seed)
    n = layout.nelem(lmax)
    a = rng.normal(size=(npol, n)) + 1j * rng.normal(size=(npol, n))
    ell, em = layout.lm_arrays(lmax)
    a[:, em == 0] = a[:, em == 0].real
    return a


def geoms(lmax):
    g = og.gauss_legendre(lmax + 1, 2 * lmax + 1)
    m = map_utils.MapInfo.map_info_gauss_legendre(lmax + 1, 2 * lmax + 1)
    return g, m


def rel(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b))


@pytest.mark.parametrize('lmax', [4, 17, 64])
@pytest.mark.parametrize('spin,npol', [(0, 1), (2, 2), ([0, 2], 3)])
def test_synthesis_fp64(dev, lmax, spin, npol):
    print(f'synthesis fp64 lmax={lmax} spin={spin}', flush=True)
    g, m = geoms(lmax)
    a = rand_alm(lmax, npol, 1)
    ref = np.zeros((npol, g.npix))
    osht.alm2map(a, ref, lmax, g, spin)
    out = torch.zeros((npol, m.npix), dtype=torch.float64, device=dev)
    sht.alm2map(torch.as_tensor(a, device=dev), out, alm_utils.AlmInfo(lmax), m, spin)
    assert rel(out.cpu().numpy(), ref) < 1e-12


@pytest.mark.parametrize('lmax', [4, 17, 64])
@pytest.mark.parametrize('spin,npol', [(0, 1), (2, 2), ([0, 2], 3)])
@pytest.mark.parametrize('adjoint', [False, True])
def test_adjoint_fp64(dev, lmax, spin, npol, adjoint):
    print(f'map2alm fp64 lmax={lmax} spin={spin} adjoint={adjoint}', flush=True)
    g, m = geoms(lmax)
    rng = np.random.default_rng(2)
    f = rng.normal(size=(npol, g.npix))
    ref = np.zeros((npol, layout.nelem(lmax)), dtype=np.complex128)
    osht.map2alm(f, ref, g, lmax, spin, adjoint=adjoint)
    out = torch.zeros((npol, layout.nelem(lmax)), dtype=torch.complex128, device=dev)
    sht.map2alm(torch.as_tensor(f, device=dev), out, m, alm_utils.AlmInfo(lmax), spin,
                adjoint=adjoint)
    assert rel(out.cpu().numpy(), ref) < 1e-12


def test_equal_area_roundtrip_parity(dev):
    '''Reduced rings (nphi < 2 lmax + 1 alias) on the multigrid geometry.'''
    lmax = 40
    gfull = og.gauss_legendre(lmax + 1, 2 * lmax + 1)
    g = og.equal_area(gfull, orig_band=0.4)
    m = map_utils.MapInfo(g.theta, g.weight, nphi=g.nphi)
    a = rand_alm(lmax, 1, 3)
    ref = np.zeros((1, g.npix))
    osht.alm2map(a, ref, lmax, g, 0)
    out = torch.zeros((1, m.npix), dtype=torch.float64, device=dev)
    sht.alm2map(torch.as_tensor(a, device=dev), out, alm_utils.AlmInfo(lmax), m, 0)
    assert rel(out.cpu().numpy(), ref) < 1e-12
    rng = np.random.default_rng(4)
    f = rng.normal(size=(1, g.npix))
    refa = np.zeros((1, layout.nelem(lmax)), dtype=np.complex128)
    osht.map2alm(f, refa, g, lmax, 0, adjoint=True)
    outa = torch.zeros((1, layout.nelem(lmax)), dtype=torch.complex128, device=dev)
    sht.map2alm(torch.as_tensor(f, device=dev), outa, m, alm_utils.AlmInfo(lmax), 0, adjoint=True)
    assert rel(outa.cpu().numpy(), refa) < 1e-12


@pytest.mark.parametrize('spin,npol', [(0, 1), ([0, 2], 3)])
def test_fp32_transforms(dev, spin, npol):
    lmax = 64
    g, m = geoms(lmax)
    a = rand_alm(lmax, npol, 5)
    ref = np.zeros((npol, g.npix))
    osht.alm2map(a, ref, lmax, g, spin)
    out = torch.zeros((npol, m.npix), dtype=torch.float32, device=dev)
    sht.alm2map(torch.as_tensor(a.astype(np.complex64), device=dev), out,
                alm_utils.AlmInfo(lmax), m, spin)
    assert rel(out.cpu().numpy(), ref) < 2e-5
    refa = np.zeros_like(a)
    osht.map2alm(ref, refa, g, lmax, spin)
    outa = torch.zeros((npol, layout.nelem(lmax)), dtype=torch.complex64, device=dev)
    sht.map2alm(torch.as_tensor(ref.astype(np.float32), device=dev), outa, m,
                alm_utils.AlmInfo(lmax), spin)
    assert rel(outa.cpu().numpy(), refa) < 2e-5


def test_batched_transforms(dev):
    lmax = 24
    g, m = geoms(lmax)
    a = np.stack([rand_alm(lmax, 3, s) for s in range(4)])
    out = torch.zeros((4, 3, m.npix), dtype=torch.float64, device=dev)
    sht.alm2map(torch.as_tensor(a, device=dev), out, alm_utils.AlmInfo(lmax), m, [0, 2])
    for t in range(4):
        ref = np.zeros((3, g.npix))
        osht.alm2map(a[t], ref, lmax, g, [0, 2])
        assert rel(out[t].cpu().numpy(), ref) < 1e-12


def test_pixel_sandwich(dev):
    lmax = 32
    g, m = geoms(lmax)
    a = rand_alm(lmax, 3, 6)
    rng = np.random.default_rng(7)
    d = rng.uniform(0.5, 2, size=(3, g.npix))
    mp = np.zeros((3, g.npix))
    osht.alm2map(a, mp, lmax, g, [0, 2])
    ref = np.zeros_like(a)
    osht.map2alm(mp * d, ref, g, lmax, [0, 2], adjoint=True)
    out = torch.zeros_like(torch.as_tensor(a, device=dev))
    sht.pixel_sandwich(torch.as_tensor(a, device=dev), out, alm_utils.AlmInfo(lmax), m, [0, 2],
                       torch.as_tensor(d, device=dev))
    assert rel(out.cpu().numpy(), ref) < 1e-12


@pytest.mark.parametrize('dtype', ['float64', 'float32'])
def test_shifted_rings(dev, dtype):
    '''Rings with a first-pixel longitude phi0 (rings 1, 5, 9, ...; the
    others at phi0 = 0), so that one plan has mirror-ring pairs with one
    shifted ring (general folding path) and pairs with none (unshifted fast
    path): synthesis, adjoint analysis and the pixel sandwich against the
    oracle.'''
    lmax = 24
    g0 = og.gauss_legendre(lmax + 1, 2 * lmax + 1)
    phi0 = np.where(np.arange(g0.nrow) % 4 == 1, 0.03 + 0.01 * np.arange(g0.nrow), 0.0)
    g = og.Geom(g0.theta, g0.weight, g0.nphi, phi0=phi0)
    m = map_utils.MapInfo(g0.theta, g0.weight, nphi=g0.nphi, phi0=phi0)
    ai = alm_utils.AlmInfo(lmax)
    cdt = torch.complex128 if dtype == 'float64' else torch.complex64
    rdt = torch.float64 if dtype == 'float64' else torch.float32
    tol = 1e-12 if dtype == 'float64' else 2e-5
    a = rand_alm(lmax, 3, 11)
    ref = np.zeros((3, g.npix))
    osht.alm2map(a, ref, lmax, g, [0, 2])
    out = torch.zeros((3, m.npix), dtype=rdt, device=dev)
    sht.alm2map(torch.as_tensor(a, device=dev).to(cdt), out, ai, m, [0, 2])
    assert rel(out.cpu().numpy(), ref) < tol
    rng = np.random.default_rng(12)
    f = rng.normal(size=(3, g.npix))
    refa = np.zeros((3, layout.nelem(lmax)), dtype=np.complex128)
    osht.map2alm(f, refa, g, lmax, [0, 2], adjoint=True)
    outa = torch.zeros((3, layout.nelem(lmax)), dtype=cdt, device=dev)
    sht.map2alm(torch.as_tensor(f, device=dev).to(rdt), outa, m, ai, [0, 2], adjoint=True)
    assert rel(outa.cpu().numpy(), refa) < tol
    d = rng.uniform(0.5, 2, size=(3, g.npix))
    refs = np.zeros_like(a)
    osht.map2alm(ref * d, refs, g, lmax, [0, 2], adjoint=True)
    outs = torch.zeros((3, layout.nelem(lmax)), dtype=cdt, device=dev)
    sht.pixel_sandwich(torch.as_tensor(a, device=dev).to(cdt), outs, ai, m, [0, 2],
                       torch.as_tensor(d, device=dev).to(rdt))
    assert rel(outs.cpu().numpy(), refs) < tol


@pytest.mark.parametrize('lmax', [255])
def test_large_lmax_roundtrip(dev, lmax):
    '''Y^T W Y = I on the GL grid (size-independent property). The oracle
    itself reaches 1.86e-12 on this case (quadrature rounding floor).'''
    m = map_utils.MapInfo.map_info_gauss_legendre(lmax + 1, 2 * lmax + 1)
    a = rand_alm(lmax, 3, 8)
    ell, _ = layout.lm_arrays(lmax)
    a[1:, ell < 2] = 0
    at = torch.as_tensor(a, device=dev)
    mp = torch.zeros((3, m.npix), dtype=torch.float64, device=dev)
    ai = alm_utils.AlmInfo(lmax)
    sht.alm2map(at, mp, ai, m, [0, 2])
    back = torch.zeros_like(at)
    sht.map2alm(mp, back, m, ai, [0, 2])
    assert rel(back.cpu().numpy(), a) < 1e-11


@pytest.mark.parametrize('spin,npol', [(0, 1), ([0, 2], 3)])
def test_ring_subset_transforms(dev, spin, npol):
    '''Transforms restricted to a ring subset (unpaired mirror rings
    included) against the oracle: synthesis equals the full synthesis on the
    subset and leaves other rings untouched; the adjoint of a map vanishing
    off the subset equals the full adjoint.'''
    lmax = 40
    print(f'ring subset spin={spin}', flush=True)
    g, m = geoms(lmax)
    rings = np.array([0, 3, 4, 5, 20, lmax - 5, lmax - 4, lmax], dtype=np.int64)
    a = rand_alm(lmax, npol, 7)
    ref = np.zeros((npol, g.npix))
    osht.alm2map(a, ref, lmax, g, spin)
    onring = np.zeros(m.npix, dtype=bool)
    for r in rings:
        onring[m.offsets[r]:m.offsets[r] + m.nphi[r]] = True
    sentinel = 12345.0
    out = torch.full((npol, m.npix), sentinel, dtype=torch.float64, device=dev)
    sht.alm2map(torch.as_tensor(a, device=dev), out, alm_utils.AlmInfo(lmax), m, spin,
                rings=rings)
    o = out.cpu().numpy()
    assert rel(o[:, onring], ref[:, onring]) < 1e-12
    assert np.all(o[:, ~onring] == sentinel)
    rng = np.random.default_rng(3)
    f = rng.normal(size=(npol, g.npix)) * onring
    aref = np.zeros((npol, layout.nelem(lmax)), dtype=np.complex128)
    osht.map2alm(f, aref, g, lmax, spin)
    got = torch.zeros((npol, layout.nelem(lmax)), dtype=torch.complex128, device=dev)
    sht.map2alm(torch.as_tensor(f, device=dev), got, m, alm_utils.AlmInfo(lmax), spin,
                rings=rings)
    assert rel(got.cpu().numpy(), aref) < 1e-12


@pytest.mark.parametrize('ntrans', [2, 3, 5, 9])
@pytest.mark.parametrize('dtype', [torch.complex128, torch.complex64])
def test_batched_legendre_tails(dev, ntrans, dtype):
    '''Batches that are not multiples of the transforms-per-block widths
    (4, 2) against one-at-a-time transforms: synthesis and adjoint, spin 0
    and 2, float64 (relative 1e-13) and float32 (2e-6).'''
    lmax = 70
    print(f'batched tails ntrans={ntrans} {dtype}', flush=True)
    g, m = geoms(lmax)
    ai = alm_utils.AlmInfo(lmax)
    rdt = torch.float64 if dtype == torch.complex128 else torch.float32
    tol = 1e-13 if dtype == torch.complex128 else 2e-6
    a = torch.as_tensor(np.stack([rand_alm(lmax, 3, 10 + s) for s in range(ntrans)]),
                        device=dev).to(dtype)
    out = torch.zeros((ntrans, 3, m.npix), dtype=rdt, device=dev)
    sht.alm2map(a, out, ai, m, [0, 2])
    back = torch.zeros_like(a)
    sht.map2alm(out, back, m, ai, [0, 2])
    for t in range(ntrans):
        o1 = torch.zeros((3, m.npix), dtype=rdt, device=dev)
        sht.alm2map(a[t].contiguous(), o1, ai, m, [0, 2])
        assert rel(out[t].cpu().numpy(), o1.cpu().numpy()) < tol
        b1 = torch.zeros_like(a[t])
        sht.map2alm(out[t].contiguous(), b1, m, ai, [0, 2])
        assert rel(back[t].cpu().numpy(), b1.cpu().numpy()) < tol
    # and the float64 batch against the oracle
    if dtype == torch.complex128:
        ref = np.zeros((3, g.npix))
        osht.alm2map(a[ntrans - 1].cpu().numpy(), ref, lmax, g, [0, 2])
        assert rel(out[ntrans - 1].cpu().numpy(), ref) < 1e-12


@pytest.mark.parametrize('lmax,nrings', [(2, 3), (3, 4), (5, 6), (70, 71), (200, 201),
                                         (60, 401), (60, 1101)])
@pytest.mark.parametrize('ntrans', [8, 13, 16, 24, 40])
def test_matrix_core_legendre(dev, lmax, nrings, ntrans):
    '''float32 batches of at least 8 transforms take the matrix-core Legendre
    kernels (16 transforms per block, partial last blocks at 13, 24 and
    40):
    synthesis and adjoint of every transform against the float64 transforms
    of the same input (relative 1e-5, or exactly zero where the float64
    result is), and the last transform's synthesis against the oracle
    (2e-5). The Gauss-Legendre grids cover band limits below one product
    group (lmax 2, 3, 5: spin 2 has 1-4 multipoles), an odd ring count (an
    equator ring without
    a mirror), one partial wave of ring pairs (36), two waves (101), four-wave
    adjoint blocks (201 pairs) and two blocks of eight waves with partial
    sums over ring-pair groups (551 pairs).'''
    print(f'matrix-core legendre lmax={lmax} nrings={nrings} ntrans={ntrans}', flush=True)
    nphi = 2 * lmax + 1
    g = og.gauss_legendre(nrings, nphi)
    m = map_utils.MapInfo.map_info_gauss_legendre(nrings, nphi)
    ai = alm_utils.AlmInfo(lmax)
    a64 = torch.as_tensor(np.stack([rand_alm(lmax, 3, 40 + s) for s in range(ntrans)]),
                          device=dev)
    a32 = a64.to(torch.complex64)
    out = torch.zeros((ntrans, 3, m.npix), dtype=torch.float32, device=dev)
    sht.alm2map(a32, out, ai, m, [0, 2])
    ref = torch.zeros((ntrans, 3, m.npix), dtype=torch.float64, device=dev)
    sht.alm2map(a32.to(torch.complex128), ref, ai, m, [0, 2])
    back = torch.zeros_like(a32)
    sht.map2alm(out, back, m, ai, [0, 2])
    bref = torch.zeros_like(a64)
    sht.map2alm(out.to(torch.float64), bref, m, ai, [0, 2])
    def close(a, b, tol):
        a, b = a.cpu().numpy(), b.cpu().numpy()
        if np.linalg.norm(b) == 0:
            return np.linalg.norm(a) == 0
        return rel(a, b) < tol

    for t in range(ntrans):
        for c in range(3):
            assert close(out[t, c], ref[t, c], 1e-5), (t, c)
            assert close(back[t, c], bref[t, c], 1e-5), (t, c)
    oref = np.zeros((3, g.npix))
    osht.alm2map(a32[ntrans - 1].cpu().numpy().astype(np.complex128), oref, lmax, g, [0, 2])
    assert close(out[ntrans - 1], torch.as_tensor(oref), 2e-5)


def test_long_rings_global_buffer(dev):
    '''Rings of 4201 pixels need a Bluestein length of 16384: the float64
    FFT buffer then lives in global memory and the kernel spectrum is
    computed on the host. Y^T W Y = I at lmax 2100 (spin 0) in float64, and
    float64 / float32 (LDS path) syntheses agree.'''
    lmax = 2100
    print(f'long rings lmax={lmax}', flush=True)
    m = map_utils.MapInfo.map_info_gauss_legendre(lmax + 1, 2 * lmax + 1)
    ai = alm_utils.AlmInfo(lmax)
    a = torch.as_tensor(rand_alm(lmax, 1, 12), device=dev)
    mp = torch.zeros((1, m.npix), dtype=torch.float64, device=dev)
    sht.alm2map(a, mp, ai, m, 0)
    back = torch.zeros_like(a)
    sht.map2alm(mp, back, m, ai, 0)
    assert rel(back.cpu().numpy(), a.cpu().numpy()) < 1e-10
    mp32 = torch.zeros((1, m.npix), dtype=torch.float32, device=dev)
    sht.alm2map(a.to(torch.complex64), mp32, ai, m, 0)
    assert rel(mp32.double().cpu().numpy(), mp.cpu().numpy()) < 1e-5


def test_legendre_flop_counter(dev):
    '''sht.LEGENDRE_FLOPS counts N_lm * N_theta * (2 n0 + 8 n2) per direction
    and transform: a TQU batch of 2 through synthesis, adjoint and the
    sandwich (two directions), and a spin-0 transform on a ring subset.'''
    lmax = 20
    g, m = geoms(lmax)
    ai = alm_utils.AlmInfo(lmax)
    a = torch.as_tensor(np.stack([rand_alm(lmax, 3, 1), rand_alm(lmax, 3, 2)]), device=dev)
    mp = torch.zeros((2, 3, m.npix), dtype=torch.float64, device=dev)
    sht.reset_legendre_flops()
    sht.alm2map(a, mp, ai, m, [0, 2])
    sht.map2alm(mp, torch.zeros_like(a), m, ai, [0, 2])
    d = torch.ones((3, m.npix), dtype=torch.float64, device=dev)
    sht.pixel_sandwich(a, torch.zeros_like(a), ai, m, [0, 2], d)
    per = layout.nelem(lmax) * m.nrow * (2 * 1 + 8 * 1) * 2
    assert sht.LEGENDRE_FLOPS['float64'] == 4 * per
    assert sht.LEGENDRE_FLOPS['float32'] == 0
    rings = np.array([0, 3, 7], dtype=np.int64)
    sht.reset_legendre_flops()
    sht.alm2map(a[0, :1].to(torch.complex64).contiguous(),
                torch.zeros((1, m.npix), dtype=torch.float32, device=dev), ai, m, 0, rings=rings)
    assert sht.LEGENDRE_FLOPS['float32'] == layout.nelem(lmax) * 3 * 2


@pytest.mark.parametrize('lmax', [17, 64])
@pytest.mark.parametrize('ntrans', [4, 5, 17, 33])
@pytest.mark.parametrize('adjoint', [False, True])
def test_adjoint_fp64_batched(dev, lmax, ntrans, adjoint):
    '''Batches of 4 or more float64 transforms take the adjoint on the
    float64 matrix cores (partial column tiles at 5, 17 and 33).'''
    print(f'map2alm fp64 batch lmax={lmax} ntrans={ntrans} adjoint={adjoint}', flush=True)
    g, m = geoms(lmax)
    rng = np.random.default_rng(3)
    f = rng.normal(size=(ntrans, 3, g.npix))
    ref = np.zeros((ntrans, 3, layout.nelem(lmax)), dtype=np.complex128)
    for t in range(ntrans):
        osht.map2alm(f[t], ref[t], g, lmax, [0, 2], adjoint=adjoint)
    out = torch.zeros(ref.shape, dtype=torch.complex128, device=dev)
    sht.map2alm(torch.as_tensor(f, device=dev), out, m, alm_utils.AlmInfo(lmax), [0, 2],
                adjoint=adjoint)
    assert rel(out.cpu().numpy(), ref) < 1e-12


@pytest.mark.parametrize('lmax,nrings', [(2, 3), (3, 4), (5, 6), (17, 18), (70, 71),
                                         (200, 201), (60, 401), (60, 1101)])
@pytest.mark.parametrize('ntrans', [8, 13, 16, 24, 33])
def test_synthesis_fp64_matrix_cores(dev, lmax, nrings, ntrans):
    '''float64 batches of at least 8 transforms take the synthesis on the
    float64 matrix cores (spin 0: 8 or 16 transforms per block, spin 2: 4 or
    8; partial column tiles at 13, 24 and 33): every transform of a TQU
    batch against the float64 vector synthesis of the same plan (relative
    1e-13), and the first, eighth and last transforms against the oracle
    (1e-12). Band limits below one group of 8 multipoles (lmax 2, 3, 5), an
    odd ring count (an equator ring without a mirror), one partial wave of
    ring pairs (9, 36), two waves (101), four-wave blocks (201 pairs) and
    several blocks (551 pairs).'''
    from optweight_amd import _lib, autotune
    print(f'fp64 matrix-core synthesis lmax={lmax} nrings={nrings} ntrans={ntrans}', flush=True)
    nphi = 2 * lmax + 1
    g = og.gauss_legendre(nrings, nphi)
    m = map_utils.MapInfo.map_info_gauss_legendre(nrings, nphi)
    ai = alm_utils.AlmInfo(lmax)
    a = np.stack([rand_alm(lmax, 3, 60 + s) for s in range(ntrans)])
    plan = sht.get_plan(m, lmax, dev)
    got = []
    for vector in (False, True):
        _lib.call('owmi_sht_plan_tune_reset', plan.handle)
        if vector:
            _lib.call('owmi_sht_plan_tune', plan.handle, 1, autotune.TUNE_MFMA64_MIN, 1 << 20)
        out = torch.zeros((ntrans, 3, m.npix), dtype=torch.float64, device=dev)
        sht.alm2map(torch.as_tensor(a, device=dev), out, ai, m, [0, 2])
        got.append(out.cpu().numpy())
    _lib.call('owmi_sht_plan_tune_reset', plan.handle)
    for t in range(ntrans):
        assert rel(got[0][t], got[1][t]) < 1e-13, t
    for t in (0, 7, ntrans - 1):
        ref = np.zeros((3, g.npix))
        osht.alm2map(a[t], ref, lmax, g, [0, 2])
        assert rel(got[0][t], ref) < 1e-12, t


@pytest.mark.parametrize('lmax,rings', [(150, None), (600, None), (600, 'subset')])
def test_fp64_matrix_cores_match_vector_kernels(dev, lmax, rings):
    '''The float64 matrix-core synthesis and adjoint equal the float64
    vector kernels (the plan's threshold raised out of reach) to relative
    1e-13 on geometries of 76 and 301 ring pairs (blocks of 2 to 8 waves,
    several ring-pair groups for spin 2) and on a ring subset.'''
    from optweight_amd import _lib, autotune
    print(f'fp64 matrix cores vs vector lmax={lmax} rings={rings}', flush=True)
    m = map_utils.MapInfo.map_info_gauss_legendre(lmax + 1, 2 * lmax + 1)
    ring_sel = None if rings is None else np.arange(5, m.nrow - 40, dtype=np.int64)
    ai = alm_utils.AlmInfo(lmax)
    rng = np.random.default_rng(4)
    f = torch.as_tensor(rng.normal(size=(16, 3, m.npix)), device=dev)
    plan = sht.get_plan(m, lmax, dev, rings=ring_sel)
    a = torch.as_tensor(np.stack([rand_alm(lmax, 3, 70 + s) for s in range(16)]), device=dev)
    outs, maps = [], []
    for vector in (False, True):
        _lib.call('owmi_sht_plan_tune_reset', plan.handle)
        if vector:
            _lib.call('owmi_sht_plan_tune', plan.handle, 1, autotune.TUNE_MFMA64_MIN, 1 << 20)
        out = torch.zeros((16, 3, ai.nelem), dtype=torch.complex128, device=dev)
        sht.map2alm(f, out, m, ai, [0, 2], rings=ring_sel)
        outs.append(out.cpu().numpy())
        mp = torch.zeros((16, 3, m.npix), dtype=torch.float64, device=dev)
        sht.alm2map(a, mp, ai, m, [0, 2], rings=ring_sel)
        maps.append(mp.cpu().numpy())
    _lib.call('owmi_sht_plan_tune_reset', plan.handle)
    for c in (slice(0, 1), slice(1, 3)):
        assert rel(outs[0][:, c], outs[1][:, c]) < 1e-13
        assert rel(maps[0][:, c], maps[1][:, c]) < 1e-13
