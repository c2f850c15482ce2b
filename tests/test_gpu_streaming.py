'''
GPU parity of the streaming kernels (per-l multiply, truncation, wavelet
accumulate, ring weights, pixel matvec, alm dot, CG updates) against the
oracle, through the C ABI. Integer/layout work is bit-exact; float work
agrees to 1e-15 relative (float64) / 1e-6 (float32).
'''
import numpy as np
import pytest
import torch

from oracle import layout, ops, sht as osht, geometry as og
from optweight_amd import alm_c_utils, map_c_utils, alm_utils, mat_utils, map_utils, _lib

pytestmark = pytest.mark.gpu


def rand_alm(lmax, npol, seed, dtype=np.complex128):
    rng = np.random.default_rng(seed)
    n = layout.nelem(lmax)
    return (rng.normal(size=(npol, n)) + 1j * rng.normal(size=(npol, n))).astype(dtype)


@pytest.mark.parametrize('dtype', [np.complex128, np.complex64])
@pytest.mark.parametrize('npol', [1, 3])
@pytest.mark.parametrize('diag', [False, True])
@pytest.mark.parametrize('inplace', [False, True])
def test_lmul(dev, dtype, npol, diag, inplace):
    lmax = 37
    a = rand_alm(lmax, npol, 1, dtype)
    rng = np.random.default_rng(2)
    lmat = rng.normal(size=(npol, lmax + 1)) if diag else rng.normal(size=(npol, npol, lmax + 1))
    ref = ops.lmul(a, lmat, lmax)
    at = torch.as_tensor(a, device=dev)
    if inplace:
        out = alm_c_utils.lmul(at, lmat, lmax, inplace=True)
        assert out.data_ptr() == at.data_ptr()
    else:
        out = alm_c_utils.lmul(at, lmat, lmax)
    tol = 1e-15 if dtype == np.complex128 else 1e-6
    np.testing.assert_allclose(out.cpu().numpy(), ref, rtol=tol, atol=tol * np.abs(ref).max())


def test_lmul_batched_and_1d(dev):
    lmax = 20
    a = np.stack([rand_alm(lmax, 3, s) for s in range(5)])
    lmat = np.random.default_rng(3).normal(size=(3, 3, lmax + 1))
    out = alm_c_utils.lmul(torch.as_tensor(a, device=dev), lmat, lmax).cpu().numpy()
    for t in range(5):
        np.testing.assert_allclose(out[t], ops.lmul(a[t], lmat, lmax), rtol=1e-15, atol=1e-14)
    w = np.arange(lmax + 1, dtype=float)
    out = alm_c_utils.lmul(torch.as_tensor(a, device=dev), w, lmax).cpu().numpy()
    np.testing.assert_allclose(out, ops.lmul(a, w, lmax), rtol=1e-15, atol=1e-14)


@pytest.mark.parametrize('dtype', [np.complex128, np.complex64])
@pytest.mark.parametrize('ntrans', [1, 16, 17, 40])
def test_lmul_transform_groups(dev, dtype, ntrans):
    '''Batches of transforms (full 3x3 and diagonal matrices, out of place
    and in place) against the oracle per transform.'''
    lmax = 45
    a = np.stack([rand_alm(lmax, 3, 10 + s, dtype) for s in range(ntrans)])
    rng = np.random.default_rng(4)
    full = rng.normal(size=(3, 3, lmax + 1))
    diag = rng.normal(size=(3, lmax + 1))
    tol = 1e-15 if dtype == np.complex128 else 1e-6
    for lmat in (full, diag):
        ref = np.stack([ops.lmul(a[t], lmat, lmax) for t in range(ntrans)])
        out = alm_c_utils.lmul(torch.as_tensor(a, device=dev), lmat, lmax).cpu().numpy()
        np.testing.assert_allclose(out, ref, rtol=tol, atol=tol * np.abs(ref).max())
        at = torch.as_tensor(a, device=dev)
        alm_c_utils.lmul(at, lmat, lmax, inplace=True)
        np.testing.assert_allclose(at.cpu().numpy(), ref, rtol=tol, atol=tol * np.abs(ref).max())


@pytest.mark.parametrize('src, dst', [(torch.complex128, torch.complex64),
                                      (torch.complex64, torch.complex128),
                                      (torch.complex128, torch.complex128),
                                      (torch.complex64, torch.complex64)])
def test_convert(dev, src, dst):
    '''owmi_convert of n complex elements: round to nearest, bit-exact
    against numpy's cast.'''
    rng = np.random.default_rng(3)
    n = 1001
    v = (rng.normal(size=n) + 1j * rng.normal(size=n)) * 10.0 ** rng.integers(-30, 30, n)
    x = torch.as_tensor(v, device=dev).to(src)
    y = torch.empty(n, dtype=dst, device=dev)
    _lib.call('owmi_convert', _lib.dtype_code(x), _lib.ptr(x), _lib.dtype_code(y), _lib.ptr(y),
              n, _lib.stream_ptr(dev))
    want = x.cpu().numpy().astype({torch.complex64: np.complex64,
                                   torch.complex128: np.complex128}[dst])
    np.testing.assert_array_equal(y.cpu().numpy(), want)


@pytest.mark.parametrize('npol', [1, 2, 4, 5, 7])
def test_lmul_components(dev, npol):
    '''Component counts on either side of the diagonal kernel's load groups
    of four and of the unrolled full-matrix kernels (1-4; 5 and 7 take the
    generic one), two transforms, against the oracle.'''
    lmax = 30
    a = np.stack([rand_alm(lmax, npol, 20 + s, np.complex128) for s in range(2)])
    rng = np.random.default_rng(npol)
    for lmat in (rng.normal(size=(npol, npol, lmax + 1)), rng.normal(size=(npol, lmax + 1))):
        ref = np.stack([ops.lmul(a[t], lmat, lmax) for t in range(2)])
        out = alm_c_utils.lmul(torch.as_tensor(a, device=dev), lmat, lmax).cpu().numpy()
        np.testing.assert_allclose(out, ref, rtol=1e-14, atol=1e-14 * np.abs(ref).max())


@pytest.mark.parametrize('dtype', [np.float64, np.float32])
@pytest.mark.parametrize('npix', [1000, 1001])
@pytest.mark.parametrize('ntrans', [1, 17, 40])
def test_pix_diag_batches(dev, dtype, npix, ntrans):
    '''Pixel-diagonal multiply of a batch of maps by one diagonal: the
    16-byte vector path (npix 1000) and the scalar path (3 x 1001 elements,
    not a multiple of the vector width), with map counts that leave partial
    groups of the maps a thread handles, out of place and in place;
    bit-exact against numpy.'''
    rng = np.random.default_rng(7)
    d = rng.normal(size=(3, npix)).astype(dtype)
    v = rng.normal(size=(ntrans, 3, npix)).astype(dtype)
    dt, vt = torch.as_tensor(d, device=dev), torch.as_tensor(v, device=dev)
    out = mat_utils.matvec(dt, vt)
    np.testing.assert_array_equal(out.cpu().numpy(), d * v)
    mat_utils.matvec(dt, vt, inplace=True)
    np.testing.assert_array_equal(vt.cpu().numpy(), d * v)


def test_trunc_alm_golden(dev):
    '''test_alm_c_utils.py:10-59 literal arrays, bit exact.'''
    alm = np.zeros((2, 6), dtype=np.complex128)
    alm[0] = [1, 2, 3, 4, 5, 6]
    alm[1] = np.asarray([1, 2, 3, 4, 5, 6]) * 2
    exp = np.zeros((2, 3), dtype=np.complex128)
    exp[0] = [1, 2, 4]
    exp[1] = np.asarray([1, 2, 4]) * 2
    out = torch.zeros((2, 3), dtype=torch.complex128, device=dev)
    alm_c_utils.trunc_alm(torch.as_tensor(alm, device=dev), out, 2, 1)
    np.testing.assert_array_equal(out.cpu().numpy(), exp)
    a = rand_alm(30, 3, 4)
    out = torch.zeros((3, layout.nelem(11)), dtype=torch.complex128, device=dev)
    alm_c_utils.trunc_alm(torch.as_tensor(a, device=dev), out, 30, 11)
    np.testing.assert_array_equal(out.cpu().numpy(), ops.trunc_alm(a, 30, 11))


@pytest.mark.parametrize('dtype', [np.complex128, np.complex64])
def test_wlm2alm_golden(dev, dtype):
    '''test_alm_c_utils.py:387-441 literal arrays.'''
    alm = np.ones((3, 10), dtype=dtype)
    alm[1] *= 2
    alm[2] *= 3
    wlm = np.ones((3, 6), dtype=dtype) * 1j
    wlm[1] *= 10
    wlm[2] *= 100
    wlm *= np.arange(6)
    w_ell = np.arange(4, dtype=float)
    w_ell[-1] = 0
    exp = np.zeros_like(alm)
    exp[0] = [1, 1 + 1j, 1 + 4j, 1, 1 + 3j, 1 + 8j, 1, 1 + 10j, 1, 1]
    exp[1] = [2, 2 + 10j, 2 + 40j, 2, 2 + 30j, 2 + 80j, 2, 2 + 100j, 2, 2]
    exp[2] = [3, 3 + 100j, 3 + 400j, 3, 3 + 300j, 3 + 800j, 3, 3 + 1000j, 3, 3]
    at = torch.as_tensor(alm, device=dev)
    alm_c_utils.wlm2alm(w_ell, torch.as_tensor(wlm, device=dev), at, 2, 3)
    np.testing.assert_allclose(at.cpu().numpy(), exp)


@pytest.mark.parametrize('dtype', [np.float64, np.float32])
def test_ringweight(dev, dtype):
    g = og.gauss_legendre(13, 25)
    m = map_utils.MapInfo(g.theta, g.weight, nphi=g.nphi)
    f = np.random.default_rng(5).normal(size=(3, g.npix)).astype(dtype)
    ref = osht.ringweight(f.copy(), g)
    ft = torch.as_tensor(f, device=dev)
    map_c_utils.apply_ringweight(ft, m)
    np.testing.assert_array_equal(ft.cpu().numpy(), ref)


def test_pix_matvec(dev):
    rng = np.random.default_rng(6)
    mat = rng.normal(size=(3, 3, 1000))
    vec = rng.normal(size=(3, 1000))
    out = mat_utils.matvec(torch.as_tensor(mat, device=dev), torch.as_tensor(vec, device=dev))
    np.testing.assert_allclose(out.cpu().numpy(), ops.matvec(mat, vec), rtol=1e-14, atol=1e-14)
    d = rng.normal(size=(3, 1000))
    out = mat_utils.matvec(torch.as_tensor(d, device=dev), torch.as_tensor(vec, device=dev))
    np.testing.assert_array_equal(out.cpu().numpy(), d * vec)


def test_contract_golden(dev):
    '''test_alm_utils.py:411-448: -40, -120, -240.'''
    a = np.asarray([1, 1, 1, 1, 2j, 2j, 2j, 3j, 3j, 4j])
    b = np.asarray([2, 2, 2, 2, 3, 3, 3, -4j, -4j, 5])
    A = torch.as_tensor(a, device=dev)
    B = torch.as_tensor(b, device=dev)
    assert alm_utils.contract_almxblm(A, B) == -40
    A3 = torch.as_tensor(np.ones((3, 10)) * a, device=dev)
    B3 = torch.as_tensor(np.ones((3, 10)) * b, device=dev)
    assert alm_utils.contract_almxblm(A3, B3) == -120
    A4 = torch.as_tensor(np.ones((2, 3, 10)) * a, device=dev)
    B4 = torch.as_tensor(np.ones((2, 3, 10)) * b, device=dev)
    assert alm_utils.contract_almxblm(A4, B4) == -240


def test_contract_large(dev):
    a = rand_alm(300, 3, 7)
    b = rand_alm(300, 3, 8)
    ref = ops.contract_almxblm(a, b)
    got = alm_utils.contract_almxblm(torch.as_tensor(a, device=dev), torch.as_tensor(b, device=dev))
    assert abs(got - ref) <= 1e-12 * abs(ref) + 1e-9


def test_cg_updates(dev):
    rng = np.random.default_rng(9)
    n = 5000
    x, r, p, q, z = [rng.normal(size=n) + 1j * rng.normal(size=n) for _ in range(5)]
    X, R, P, Q, Z = [torch.as_tensor(v, device=dev) for v in (x, r, p, q, z)]
    num = torch.tensor(2.5, dtype=torch.float64, device=dev)
    den = torch.tensor(-1.25, dtype=torch.float64, device=dev)
    _lib.call('owmi_cg_xr', _lib.F64, _lib.ptr(X), _lib.ptr(R), _lib.ptr(P), _lib.ptr(Q),
              _lib.ptr(num), _lib.ptr(den), n, _lib.stream_ptr(dev))
    np.testing.assert_allclose(X.cpu().numpy(), x - 2 * p, rtol=1e-15, atol=1e-15)
    np.testing.assert_allclose(R.cpu().numpy(), r + 2 * q, rtol=1e-15, atol=1e-15)
    _lib.call('owmi_cg_p', _lib.F64, _lib.ptr(P), _lib.ptr(Z), _lib.ptr(num), _lib.ptr(den), n,
              _lib.stream_ptr(dev))
    np.testing.assert_allclose(P.cpu().numpy(), z - 2 * p, rtol=1e-15, atol=1e-15)


@pytest.mark.parametrize('dtype', ['float64', 'float32'])
@pytest.mark.parametrize('npol', [1, 3, 4])
@pytest.mark.parametrize('case', ['identity_a', 'both', 'alias_a', 'alias_b'])
def test_lmul_add(dev, dtype, npol, case):
    '''out = La a + Lb b against the oracle's lmul, with out aliasing a or b.'''
    print(f'lmul_add {case} npol={npol} {dtype}', flush=True)
    from oracle import ops as oops
    from optweight_amd import alm_c_utils
    lmax = 37
    rng = np.random.default_rng(npol)
    n = layout.nelem(lmax)
    a = rng.normal(size=(2, npol, n)) + 1j * rng.normal(size=(2, npol, n))
    b = rng.normal(size=(2, npol, n)) + 1j * rng.normal(size=(2, npol, n))
    la = rng.normal(size=(npol, npol, lmax + 1))
    lb = rng.normal(size=(npol, npol, lmax + 1))
    cdt = torch.complex128 if dtype == 'float64' else torch.complex64
    at, bt = torch.as_tensor(a, device=dev).to(cdt), torch.as_tensor(b, device=dev).to(cdt)
    use_a = None if case == 'identity_a' else la
    ref = np.stack([(a[t] if use_a is None else oops.lmul(a[t], la, lmax)) + oops.lmul(b[t], lb, lmax)
                    for t in range(2)])
    out = {'alias_a': at, 'alias_b': bt}.get(case, torch.empty_like(at))
    got = alm_c_utils.lmul_add(at, use_a, bt, lb, alm_utils.AlmInfo(lmax), out=out)
    assert got.data_ptr() == out.data_ptr()
    tol = 1e-14 if dtype == 'float64' else 2e-6
    assert np.linalg.norm(got.cpu().numpy() - ref) / np.linalg.norm(ref) < tol


@pytest.mark.parametrize('dtype', ['float64', 'float32'])
@pytest.mark.parametrize('npol', [1, 3])
@pytest.mark.parametrize('ntrans', [1, 5, 128])
@pytest.mark.parametrize('alias_a', [False, True])
def test_lmul_add_dot(dev, dtype, npol, ntrans, alias_a):
    '''owmi_lmul_add_dot: out equal to owmi_lmul_add's bit for bit, and the
    real-field <b, out> of every transform equal to owmi_alm_dot_batch of
    b with out (relative 1e-13 in float64, 1e-12 in float32: the products
    are summed in float64 in another fixed order), with out aliasing a.'''
    print(f'lmul_add_dot npol={npol} ntrans={ntrans} {dtype} alias_a={alias_a}', flush=True)
    from optweight_amd import alm_c_utils
    lmax = 37
    rng = np.random.default_rng(10 * npol + ntrans)
    n = layout.nelem(lmax)
    a = rng.normal(size=(ntrans, npol, n)) + 1j * rng.normal(size=(ntrans, npol, n))
    b = rng.normal(size=(ntrans, npol, n)) + 1j * rng.normal(size=(ntrans, npol, n))
    a[..., :lmax + 1] = a[..., :lmax + 1].real
    b[..., :lmax + 1] = b[..., :lmax + 1].real
    la = rng.normal(size=(npol, npol, lmax + 1))
    lb = rng.normal(size=(npol, npol, lmax + 1))
    cdt = torch.complex128 if dtype == 'float64' else torch.complex64
    ai = alm_utils.AlmInfo(lmax)
    at, bt = torch.as_tensor(a, device=dev).to(cdt), torch.as_tensor(b, device=dev).to(cdt)
    ref = alm_c_utils.lmul_add(at, la, bt, lb, ai, out=torch.empty_like(at))
    dref = alm_utils.contract_almxblm_batch(bt, ref).cpu().numpy()
    a2 = at.clone()
    out = a2 if alias_a else torch.empty_like(at)
    dot = torch.empty(ntrans, dtype=torch.float64, device=dev)
    got = alm_c_utils.lmul_add(a2, la, bt, lb, ai, out=out, dot_out=dot)
    assert torch.equal(got, ref)
    tol = 1e-13 if dtype == 'float64' else 1e-12
    assert np.all(np.abs(dot.cpu().numpy() - dref) <= tol * np.abs(dref))
    with pytest.raises(Exception):
        alm_c_utils.lmul_add(at, la, bt, lb, ai, out=bt, dot_out=dot)


@pytest.mark.parametrize('dtype', ['float64', 'float32'])
@pytest.mark.parametrize('case', ['jacobi', 'masked_residual', 'masked_add', 'scale'])
@pytest.mark.parametrize('npix', [1000, 1001])
def test_map_fused(dev, dtype, case, npix):
    '''out = x + d_c mask (a - b) against the same sequence of torch operations
    (bit-exact in float64), in place on x or b; 16-byte vectors (npix 1000)
    and single elements (npix 1001).'''
    print(f'map_fused {case} {dtype} npix={npix}', flush=True)
    rng = np.random.default_rng(7)
    rdt = torch.float64 if dtype == 'float64' else torch.float32
    t = lambda a: torch.as_tensor(a, device=dev).to(rdt)
    a, b, x = (t(rng.normal(size=(4, 2, npix))) for _ in range(3))
    d = t(rng.uniform(0.1, 1, size=2))
    mask = t(rng.uniform(size=(2, npix)) > 0.3)
    if case == 'jacobi':
        ref = x + d[:, None] * (a - b)
        got = map_c_utils.map_fused(a, b=b, x=x, dcomp=d, out=x)
    elif case == 'masked_residual':
        ref = (a - b) * mask
        got = map_c_utils.map_fused(a, b=b, mask=mask, out=b)
    elif case == 'masked_add':
        ref = x + a * mask
        got = map_c_utils.map_fused(a, x=x, mask=mask, out=x)
    else:
        ref = d[:, None] * a
        got = map_c_utils.map_fused(a, dcomp=d)
    tol = 0 if dtype == 'float64' else 1e-6
    assert (got - ref).abs().max().item() <= tol * ref.abs().max().item()


def test_dots_on_two_streams(dev):
    '''Batched dots queued on two streams at once each use their own partial
    sums (one reduction buffer per stream): both equal the serial results.'''
    lmax, nb = 200, 8
    a = [torch.as_tensor(np.stack([rand_alm(lmax, 3, 40 + 10 * k + i) for i in range(nb)]),
                         device=dev) for k in range(2)]
    ref = [alm_utils.contract_almxblm_batch(x, x).cpu().numpy() for x in a]
    streams = [torch.cuda.Stream(dev) for _ in range(2)]
    outs = [torch.empty(nb, dtype=torch.float64, device=dev) for _ in range(2)]
    torch.cuda.synchronize()
    for rep in range(20):
        for k in range(2):
            with torch.cuda.stream(streams[k]):
                alm_utils.contract_almxblm_batch(a[k], a[k], out=outs[k])
        torch.cuda.synchronize()
        for k in range(2):
            assert np.array_equal(outs[k].cpu().numpy(), ref[k]), (rep, k)
