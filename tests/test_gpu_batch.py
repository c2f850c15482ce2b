'''
Batched right-hand sides and the mixed-precision solver on the GPU.

  * per-RHS dots (owmi_alm_dot_batch) against single dots: relative 1e-12;
  * batched CG updates against the update formula in float64, and the
    zero-denominator freeze (x, r, p of that RHS unchanged / cleared);
  * a batch of 3 RHS through CGWienerMap (pinv + multigrid on T) against 3
    single-RHS solvers: relative L2 of x <= 1e-12 at every step;
  * solve_fp64 / solve_mixed on a batch: every RHS reaches float64 true
    relative residual <= 1e-6, and both solutions agree with the oracle's
    converged float64 solution to relative L2 <= 1e-4 (a true residual of
    1e-6 bounds the solution error by cond(A) * 1e-6).
'''
import numpy as np
import pytest
import torch

from oracle import geometry as og, wiener as ow, ops as oops, sht as osht
from optweight_amd import solvers, preconditioners, workloads, alm_utils, mixed, _lib

pytestmark = pytest.mark.gpu


def rel(a, b):
    a = a.cpu().numpy() if isinstance(a, torch.Tensor) else np.asarray(a)
    b = b.cpu().numpy() if isinstance(b, torch.Tensor) else np.asarray(b)
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def oracle_geom(minfo):
    return og.Geom(minfo.theta, minfo.weight, minfo.nphi, phi0=minfo.phi0)


def oracle_synth(a, minfo, lmax, spin):
    out = np.zeros((a.shape[0], minfo.npix))
    osht.alm2map(a, out, lmax, oracle_geom(minfo), spin)
    return out


@pytest.fixture(scope='module')
def prob():
    return workloads.polarized(24, seed=5, ndisk=10, synth=oracle_synth)


@pytest.fixture(scope='module')
def data(prob):
    return workloads.batch_data(prob, 3, seed=1, synth=oracle_synth)


def add_precs(solver, prob, dev):
    solver.add_preconditioner(preconditioners.PseudoInvPreconditioner(
        prob.ainfo, prob.icov_ell, prob.icov_pix, prob.minfo, prob.spin))
    mg = preconditioners.MaskedPreconditioner(
        prob.ainfo, prob.icov_ell[0:1, 0:1], 0, prob.mask_bool[0], prob.minfo,
        min_pix=100, n_jacobi=1, lmax_r_ell=6000, device=dev)
    solver.add_preconditioner(mg, sel=np.s_[0])
    return solver


def rand_alm_t(rng, shape, lmax, dtype, dev):
    v = rng.normal(size=shape) + 1j * rng.normal(size=shape)
    v[..., :lmax + 1] = v[..., :lmax + 1].real
    return torch.as_tensor(v, device=dev).to(dtype)


@pytest.mark.parametrize('dtype', [torch.complex128, torch.complex64])
def test_dot_batch(dev, dtype):
    lmax = 40
    nelem = alm_utils.AlmInfo(lmax).nelem
    rng = np.random.default_rng(0)
    a = rand_alm_t(rng, (5, 3, nelem), lmax, dtype, dev)
    b = rand_alm_t(rng, (5, 3, nelem), lmax, dtype, dev)
    got = alm_utils.contract_almxblm_batch(a, b).cpu().numpy()
    for i in range(5):
        ref = alm_utils.contract_almxblm(a[i], b[i])
        assert abs(got[i] - ref) <= 1e-12 * abs(ref)
    # closed form: <a, a> = sum_{m=0} |a|^2 + 2 sum_{m>0} |a|^2
    an = a.to(torch.complex128).cpu().numpy()
    cf = 2 * np.sum(np.abs(an[1]) ** 2) - np.sum(np.abs(an[1, :, :lmax + 1]) ** 2)
    assert abs(alm_utils.contract_almxblm_batch(a, a).cpu().numpy()[1] - cf) <= 1e-12 * cf


@pytest.mark.parametrize('dtype', [torch.complex128, torch.complex64])
@pytest.mark.parametrize('lmax, nb, ncomp', [(500, 16, 3), (700, 3, 3), (1500, 1, 1)])
def test_dot_batch_sizes(dev, dtype, lmax, nb, ncomp):
    '''Sizes where every thread of the partial-sum kernel takes several
    strided steps per row (and a remainder), batched and single, against a
    float64 numpy sum: <a, b> = 2 sum Re(a conj b) - sum_{m=0} Re(a conj b).'''
    nelem = alm_utils.AlmInfo(lmax).nelem
    rng = np.random.default_rng(lmax)
    a = rand_alm_t(rng, (nb, ncomp, nelem), lmax, dtype, dev)
    b = rand_alm_t(rng, (nb, ncomp, nelem), lmax, dtype, dev)
    an = a.to(torch.complex128).cpu().numpy()
    bn = b.to(torch.complex128).cpu().numpy()
    prod = (an * np.conj(bn)).real
    ref = 2 * prod.sum(axis=(1, 2)) - prod[..., :lmax + 1].sum(axis=(1, 2))
    scale = 2 * np.abs(prod).sum(axis=(1, 2))
    got = alm_utils.contract_almxblm_batch(a, b).cpu().numpy()
    assert np.all(np.abs(got - ref) <= 1e-13 * scale)
    for i in range(nb):
        assert abs(alm_utils.contract_almxblm(a[i], b[i]) - ref[i]) <= 1e-13 * scale[i]


def test_cg_batch_updates(dev):
    rng = np.random.default_rng(1)
    nb, n = 4, 1000
    x, r, p, q = (torch.as_tensor(rng.normal(size=(nb, n)) + 1j * rng.normal(size=(nb, n)),
                                  device=dev) for _ in range(4))
    num = torch.as_tensor(rng.uniform(1, 2, nb), device=dev)
    den = torch.as_tensor(rng.uniform(1, 2, nb), device=dev)
    den[2] = 0
    x0, r0, p0 = x.clone(), r.clone(), p.clone()
    st = _lib.stream_ptr(dev)
    _lib.call('owmi_cg_xr_batch', _lib.F64, _lib.ptr(x), _lib.F64, _lib.ptr(r), _lib.ptr(p),
              _lib.ptr(q), _lib.ptr(num), _lib.ptr(den), n, nb, st)
    alpha = (num / den).cpu().numpy()
    alpha[2] = 0
    for i in range(nb):
        assert rel(x[i], x0[i].cpu().numpy() + alpha[i] * p0[i].cpu().numpy()) < 1e-15
        assert rel(r[i], r0[i].cpu().numpy() - alpha[i] * q[i].cpu().numpy()) < 1e-15
    z = torch.as_tensor(rng.normal(size=(nb, n)) + 1j * rng.normal(size=(nb, n)), device=dev)
    _lib.call('owmi_cg_p_batch', _lib.F64, _lib.ptr(p), _lib.ptr(z), _lib.ptr(num),
              _lib.ptr(den), n, nb, st)
    for i in range(nb):
        exp = z[i].cpu().numpy() + alpha[i] * p0[i].cpu().numpy()
        assert rel(p[i], exp) < 1e-15


@pytest.mark.parametrize('nb,nrows,lmax', [(1, 1, 30), (3, 3, 17), (16, 3, 40)])
@pytest.mark.parametrize('xdt,rdt', [(torch.complex128, torch.complex128),
                                     (torch.complex128, torch.complex64),
                                     (torch.complex64, torch.complex64)])
def test_cg_xr_rr_fused(dev, nb, nrows, lmax, xdt, rdt):
    '''The fused x/r update returns the update of owmi_cg_xr_batch bit for
    bit and <r, r> of the new residuals as the batched dot computes it
    (relative 1e-13; a right-hand side with a zero denominator keeps r).'''
    print(f'cg_xr_rr nb={nb} nrows={nrows} lmax={lmax} {xdt} {rdt}', flush=True)
    rng = np.random.default_rng(5)
    ne = (lmax + 1) * (lmax + 2) // 2
    shape = (nb, nrows, ne)

    def rand(dt):
        v = torch.as_tensor(rng.normal(size=shape) + 1j * rng.normal(size=shape), device=dev)
        return v.to(dt)
    x, r, p, q = rand(xdt), rand(rdt), rand(rdt), rand(rdt)
    num = torch.as_tensor(rng.uniform(1, 2, nb), device=dev)
    den = torch.as_tensor(rng.uniform(1, 2, nb), device=dev)
    den[nb // 2] = 0
    x1, r1 = x.clone(), r.clone()
    mixed._xr(x1, r1, p, q, num, den, nb)
    rr1 = alm_utils.contract_almxblm_batch(r1, r1).cpu().numpy()
    rr = torch.empty(nb, dtype=torch.float64, device=dev)
    mixed._xr_rr(x, r, p, q, num, den, rr, nb)
    assert torch.equal(x, x1) and torch.equal(r, r1)
    np.testing.assert_allclose(rr.cpu().numpy(), rr1, rtol=1e-13)


def test_batched_cg_matches_single(dev, prob, data):
    batch = add_precs(solvers.CGWienerMap.from_arrays(
        data, prob.minfo, prob.ainfo, prob.icov_ell, prob.icov_pix, device=dev), prob, dev)
    singles = [add_precs(solvers.CGWienerMap.from_arrays(
        data[i], prob.minfo, prob.ainfo, prob.icov_ell, prob.icov_pix, device=dev), prob, dev)
        for i in range(data.shape[0])]
    batch.init_solver()
    for s in singles:
        s.init_solver()
    assert batch.b_vec.shape == (3, 3, prob.ainfo.nelem)
    for step in range(5):
        batch.step()
        err = batch.err
        for i, s in enumerate(singles):
            s.step()
            assert rel(batch.x[i], s.x) < 1e-12, f'rhs {i} step {step}'
            assert abs(err[i] - s.err) <= 1e-10 * abs(s.err)


def true_res(solver, x):
    r = solver.b0 - solver.a_matrix(x)
    rr = alm_utils.contract_almxblm_batch(r, r).cpu().numpy()
    bb = alm_utils.contract_almxblm_batch(solver.b0, solver.b0).cpu().numpy()
    return np.sqrt(rr / bb)


def test_solve_fp64_and_mixed_batch(dev, prob, data):
    solver = add_precs(solvers.CGWienerMap.from_arrays(
        data, prob.minfo, prob.ainfo, prob.icov_ell, prob.icov_pix, device=dev), prob, dev)
    x64, info64 = mixed.solve_fp64(solver.a_matrix, solver.b0, solver.preconditioner,
                                   tol=1e-6, verify=True)
    assert np.all(info64.residual <= 1.5e-6)
    xm, infom = mixed.solve_mixed(solver.a_matrix, solver.b0, solver.preconditioner, tol=1e-6)
    res = true_res(solver, xm)
    assert np.all(res <= 1e-6), res
    assert np.allclose(res, infom.residual, rtol=1e-6)
    assert infom.replacements >= 1
    # against the oracle's converged float64 solution, RHS 0
    g = oracle_geom(prob.minfo)
    osol = ow.WienerMap.from_arrays(data[0], g, prob.lmax, prob.icov_ell, prob.icov_pix)
    osol.add_preconditioner(ow.PseudoInvPreconditioner(
        prob.lmax, prob.icov_ell, prob.icov_pix, g, prob.spin))
    cg = osol.init_solver()
    b0 = np.sqrt(oops.contract_almxblm(osol.b0, osol.b0))
    for _ in range(300):
        cg.step()
        if osol.get_residual() / b0 < 1e-11:
            break
    em, e64 = rel(xm[0], cg.x), rel(x64[0], cg.x)
    print(f'relative L2 vs oracle: mixed {em:.3e}, fp64 {e64:.3e}')
    assert em < 1e-4 and e64 < 1e-4


def test_mixed_single_rhs(dev, prob):
    solver = add_precs(solvers.CGWienerMap.from_arrays(
        prob.imap, prob.minfo, prob.ainfo, prob.icov_ell, prob.icov_pix, device=dev), prob, dev)
    x, info = mixed.solve_mixed(solver.a_matrix, solver.b0, solver.preconditioner, tol=1e-6,
                                delta=1e-2)
    assert x.shape == solver.b0.shape
    r = solver.b0 - solver.a_matrix(x)
    assert np.sqrt(alm_utils.contract_almxblm(r, r)
                   / alm_utils.contract_almxblm(solver.b0, solver.b0)) <= 1e-6


@pytest.mark.parametrize('nb', [3, 8])
def test_mixed_graph_replay_matches_eager(dev, prob, nb):
    '''The HIP-graph replay of the mixed iteration phases (recorded after one
    eager iteration, replayed every later iteration with host syncs and
    eager float64 residual replacements in between) gives the eager solve:
    same iteration and replacement counts, same final residuals and
    solutions to relative 1e-12 (nb 8 runs the matrix-core Legendre
    kernels inside the graphs).'''
    print(f'graph replay nb={nb}', flush=True)
    dat = workloads.batch_data(prob, nb, seed=3, synth=oracle_synth)
    solver = add_precs(solvers.CGWienerMap.from_arrays(
        dat, prob.minfo, prob.ainfo, prob.icov_ell, prob.icov_pix, device=dev), prob, dev)
    xe, ie = mixed.solve_mixed(solver.a_matrix, solver.b0, solver.preconditioner, tol=1e-6,
                               delta=1e-2, graphs=False)
    xg, ig = mixed.solve_mixed(solver.a_matrix, solver.b0, solver.preconditioner, tol=1e-6,
                               delta=1e-2, graphs=True)
    assert ig.graphs and not ie.graphs
    assert ig.iterations == ie.iterations and ig.replacements == ie.replacements
    assert list(ig.iterations_rhs) == list(ie.iterations_rhs)
    assert ig.iterations >= 3
    np.testing.assert_allclose(ig.residual, ie.residual, rtol=1e-10)
    assert rel(xg, xe) < 1e-12
    # a second graph solve in the same process (fresh recording)
    xg2, ig2 = mixed.solve_mixed(solver.a_matrix, solver.b0, solver.preconditioner, tol=1e-6,
                                 delta=1e-2, graphs=True)
    assert ig2.iterations == ie.iterations and rel(xg2, xe) < 1e-12


def test_plan_released_during_graph_recording(dev):
    '''An SHT plan whose last reference goes away inside a HIP-graph
    recording is destroyed after it (freeing device memory during the
    recording would invalidate it); the recorded graph replays correctly.'''
    print('plan release during graph recording', flush=True)
    from optweight_amd import sht, map_utils, alm_utils
    minfo = map_utils.MapInfo.map_info_gauss_legendre(9, 17)
    ainfo = alm_utils.AlmInfo(8)
    plan = sht.get_plan(minfo, 8, dev)
    del minfo.__dict__['_plans']
    x = torch.ones(1024, device=dev)
    y = torch.empty_like(x)
    y.copy_(x * 2)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        del plan
        y.copy_(x * 2)
    assert len(sht._DEFERRED) == 1
    g.replay()
    torch.cuda.synchronize()
    assert torch.equal(y, x * 2)
    sht.get_plan(map_utils.MapInfo.map_info_gauss_legendre(5, 9), 4, dev)
    assert len(sht._DEFERRED) == 0
    del ainfo
