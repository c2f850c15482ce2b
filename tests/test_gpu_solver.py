'''
GPU parity of the Wiener-filter operators, preconditioners and CG solver
against the oracle, on seeded problems of the benchmark's construction at
small lmax. Float64 throughout:
  * single applications of A, the pseudo-inverse and the multigrid
    preconditioner: relative L2 <= 1e-12;
  * 5-step CG trajectories: relative L2 of x <= 1e-10 at every step;
  * closed forms of the reference solver tests (test_solvers.py:45-175) and
    the full-sky white-noise exact inverse (CG converges in one step).
'''
import numpy as np
import pytest
import torch

from oracle import geometry as og, wiener as ow, ops as oops, sht as osht
from optweight_amd import solvers, preconditioners, workloads, alm_utils

pytestmark = pytest.mark.gpu


def rel(a, b):
    a = a.cpu().numpy() if isinstance(a, torch.Tensor) else np.asarray(a)
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def oracle_geom(minfo):
    return og.Geom(minfo.theta, minfo.weight, minfo.nphi, phi0=minfo.phi0)


def oracle_synth(a, minfo, lmax, spin):
    out = np.zeros((a.shape[0], minfo.npix))
    osht.alm2map(a, out, lmax, oracle_geom(minfo), spin)
    return out


@pytest.fixture(scope='module')
def pol_problem():
    return workloads.polarized(24, seed=3, ndisk=10, synth=oracle_synth)


@pytest.fixture(scope='module')
def t_problem():
    return workloads.temperature(40, seed=4, ndisk=10, synth=oracle_synth)


def build_pair(prob, dev, with_pinv=True):
    solver = solvers.CGWienerMap.from_arrays(prob.imap, prob.minfo, prob.ainfo, prob.icov_ell,
                                             prob.icov_pix, device=dev)
    g = oracle_geom(prob.minfo)
    osol = ow.WienerMap.from_arrays(prob.imap, g, prob.lmax, prob.icov_ell, prob.icov_pix)
    if with_pinv:
        solver.add_preconditioner(preconditioners.PseudoInvPreconditioner(
            prob.ainfo, prob.icov_ell, prob.icov_pix, prob.minfo, prob.spin))
        osol.add_preconditioner(ow.PseudoInvPreconditioner(
            prob.lmax, prob.icov_ell, prob.icov_pix, g, prob.spin))
    return solver, osol


def test_b_vec_and_a_matrix(dev, pol_problem):
    prob = pol_problem
    solver, osol = build_pair(prob, dev, with_pinv=False)
    assert rel(solver.b0, osol.b0) < 1e-12
    rng = np.random.default_rng(1)
    v = rng.normal(size=osol.b0.shape) + 1j * rng.normal(size=osol.b0.shape)
    v[:, :prob.lmax + 1] = v[:, :prob.lmax + 1].real
    ref = osol.a_matrix(v)
    got = solver.a_matrix(torch.as_tensor(v, device=dev))
    assert solver._fast and all(v is not None for v in solver._fast.values()), \
        'fused A path not taken'
    assert rel(got, ref) < 1e-12
    # the unfused (general callable) path gives the same result
    solver._fast = {k: None for k in solver._fast}
    got2 = solver.a_matrix(torch.as_tensor(v, device=dev))
    assert rel(got2, ref) < 1e-12


def test_a_symmetric(dev, pol_problem):
    prob = pol_problem
    solver, _ = build_pair(prob, dev, with_pinv=False)
    rng = np.random.default_rng(2)
    shape = (3, prob.ainfo.nelem)
    u = torch.as_tensor(rng.normal(size=shape) + 1j * rng.normal(size=shape), device=dev)
    w = torch.as_tensor(rng.normal(size=shape) + 1j * rng.normal(size=shape), device=dev)
    u[:, :prob.lmax + 1] = u[:, :prob.lmax + 1].real
    w[:, :prob.lmax + 1] = w[:, :prob.lmax + 1].real
    a = alm_utils.contract_almxblm(u, solver.a_matrix(w))
    b = alm_utils.contract_almxblm(solver.a_matrix(u), w)
    assert abs(a - b) <= 1e-12 * abs(a)


def test_pinv_application(dev, pol_problem):
    prob = pol_problem
    g = oracle_geom(prob.minfo)
    p = preconditioners.PseudoInvPreconditioner(prob.ainfo, prob.icov_ell, prob.icov_pix,
                                                prob.minfo, prob.spin)
    po = ow.PseudoInvPreconditioner(prob.lmax, prob.icov_ell, prob.icov_pix, g, prob.spin)
    rng = np.random.default_rng(3)
    v = rng.normal(size=(3, prob.ainfo.nelem)) + 1j * rng.normal(size=(3, prob.ainfo.nelem))
    vt = torch.as_tensor(v, device=dev)
    got = p(vt)
    assert rel(vt, v) == 0, 'preconditioner modified its input'
    assert rel(got, po(v)) < 1e-12


def test_cg_trajectory(dev, pol_problem):
    prob = pol_problem
    solver, osol = build_pair(prob, dev)
    solver.init_solver()
    cg = osol.init_solver()
    assert abs(solver.rz0 - cg.rz0) <= 1e-12 * abs(cg.rz0)
    for i in range(5):
        solver.step()
        cg.step()
        assert rel(solver.x, cg.x) < 1e-10, f'step {i}'
        assert abs(solver.err - cg.err) <= 1e-9 * abs(cg.err)


@pytest.mark.parametrize('precond', ['none', 'returns_input'])
def test_cg_trajectory_identity_preconditioner(dev, pol_problem, precond):
    '''Without a preconditioner (or with one that returns its input) the
    direction starts as the residual array itself; the in-place direction
    update must leave the residual alone, as the reference's out-of-place
    update does (regression test).'''
    prob = pol_problem
    solver, osol = build_pair(prob, dev, with_pinv=False)
    if precond == 'returns_input':
        solver.add_preconditioner(lambda a: a)
        osol.add_preconditioner(lambda a: a)
    solver.init_solver()
    cg = osol.init_solver()
    for i in range(4):
        solver.step()
        cg.step()
        assert rel(solver.x, cg.x) < 1e-10, f'step {i}'


def test_converged_solution(dev, t_problem):
    '''Both solvers driven to true relative residual <= 1e-12.'''
    prob = t_problem
    solver, osol = build_pair(prob, dev)
    solver.init_solver()
    cg = osol.init_solver()
    b0 = np.sqrt(oops.contract_almxblm(osol.b0, osol.b0))
    for _ in range(200):
        cg.step()
        if osol.get_residual() / b0 < 1e-12:
            break
    for _ in range(200):
        solver.step()
        if solver.get_residual() / b0 < 1e-12:
            break
    assert solver.get_residual() / b0 < 1e-12
    assert rel(solver.x, cg.x) < 1e-10


class TestReferenceClosedForms:
    '''test_solvers.py:45-175 re-expressed with device vectors.'''

    def test_run(self, dev):
        alm_data = torch.as_tensor(np.asarray([1, 2, 3, 4j, 5j, 6], dtype=np.complex128),
                                   device=dev)
        solver = solvers.CGWienerMap(alm_data, lambda a: 2 * a, lambda a: 3 * a,
                                     (lambda a: a, lambda a: a))
        solver.init_solver()
        b_in = solver.b.clone()
        while solver.i < 1:
            solver.step()
        np.testing.assert_allclose(solver.x.cpu().numpy(), alm_data.cpu().numpy() * 3 / 5,
                                   atol=1e-12)
        np.testing.assert_allclose(solver.A(solver.x).cpu().numpy(), b_in.cpu().numpy(),
                                   atol=1e-12)

    def test_run_beam_icov_qform(self, dev):
        alm_data = torch.as_tensor(np.asarray([1, 2, 3, 4j, 5j, 6], dtype=np.complex128),
                                   device=dev)
        solver = solvers.CGWienerMap(alm_data, lambda a: 2 * a, lambda a: 3 * a,
                                     (lambda a: a, lambda a: a), beam=lambda a: 0.1 * a)
        solver.init_solver()
        while solver.i < 1:
            solver.step()
        x_exp = alm_data.cpu().numpy() * (0.1 * 3 / (2 + 0.01 * 3))
        np.testing.assert_allclose(solver.x.cpu().numpy(), x_exp, atol=1e-12)
        icov_exp = alm_data.cpu().numpy() * (10 / (1 / 2 + 100 * 1 / 3))
        np.testing.assert_allclose(solver.get_icov().cpu().numpy(), icov_exp, atol=1e-12)

    def test_qform(self, dev):
        alm_data = torch.as_tensor(np.asarray([1, 2, 3, 4j, 5j, 6], dtype=np.complex128),
                                   device=dev)
        solver = solvers.CGWienerMap(alm_data, lambda a: 2 * a, lambda a: 3 * a,
                                     (lambda a: a, lambda a: a))
        solver.init_solver()
        while solver.i < 1:
            solver.step()
        qform_exp = 0.5 * 168 * (3 / 5) ** 2 * 5 - 168 * 3 * (3 / 5)
        assert abs(solver.get_qform() - qform_exp) < 1e-9

    def test_constrained(self, dev):
        alm_data = torch.ones(6, dtype=torch.complex128, device=dev)
        ris = torch.arange(6, dtype=torch.float64).to(torch.complex128).to(dev)
        rin = torch.flip(ris, [0])
        solver = solvers.CGWienerMap(alm_data, lambda a: 2 * a, lambda a: 3 * a,
                                     (lambda a: a, lambda a: a), beam=lambda a: 0.1 * a,
                                     rand_isignal=ris, rand_inoise=rin)
        solver.init_solver()
        while solver.i < 1:
            solver.step()
        x_exp = (0.1 * 3 * alm_data + ris + 0.1 * rin) / (0.01 * 3 + 2)
        np.testing.assert_allclose(solver.x.cpu().numpy(), x_exp.cpu().numpy(), atol=1e-12)

    def test_add_preconditioner(self, dev):
        alm_data = torch.arange(12, dtype=torch.float32).to(torch.complex64).to(dev).reshape(2, 6)
        solver = solvers.CGWienerMap(alm_data, lambda a: a, lambda a: a,
                                     (lambda a: a, lambda a: a))
        assert solver.preconditioner is None
        solver.init_solver()
        np.testing.assert_allclose(solver.M(alm_data).cpu(), alm_data.cpu())
        solver.add_preconditioner(lambda a: 2 * a)
        np.testing.assert_allclose(solver.preconditioner(alm_data).cpu(), 2 * alm_data.cpu())
        solver.add_preconditioner(lambda a: 2 * a, sel=np.s_[1])
        exp = alm_data.clone()
        exp[0] *= 2
        exp[1] *= 4
        np.testing.assert_allclose(solver.preconditioner(alm_data).cpu(), exp.cpu())
        solver.init_solver()
        assert solver.M is solver.preconditioner
        solver.reset_preconditioner()
        assert solver.preconditioner is None


def test_fullsky_white_noise_one_step(dev):
    '''Full sky, white noise icov = w / sigma^2: the pseudo-inverse is
    A^-1 exactly, so preconditioned CG converges in one step.'''
    lmax = 32
    minfo = workloads.gl_minfo(lmax)
    ainfo = alm_utils.AlmInfo(lmax)
    cov = workloads.cov_ell_TQU(lmax)
    icov_ell = workloads.pinv_ell(cov)
    w = np.repeat(minfo.weight, minfo.nphi)
    icov_pix = np.stack([w / 4., w / 8., w / 8.])
    rng = np.random.default_rng(5)
    imap = rng.normal(size=(3, minfo.npix))
    solver = solvers.CGWienerMap.from_arrays(imap, minfo, ainfo, icov_ell, icov_pix, device=dev)
    solver.add_preconditioner(preconditioners.PseudoInvPreconditioner(
        ainfo, icov_ell, icov_pix, minfo, [0, 2]))
    solver.init_solver()
    solver.step()
    b0 = np.sqrt(alm_utils.contract_almxblm(solver.b0, solver.b0))
    assert solver.get_residual() / b0 < 1e-11


def test_multigrid_parity(dev, t_problem):
    '''Level geometry and masks bit-identical; the smoother's probed
    diagonal equal (both round the float64 probe to float32); dense G to
    1e-12; one preconditioner application to 1e-10, each side with its own
    smoother and its own pseudo-inverse of the coarsest G.'''
    prob = t_problem
    mask = prob.mask_bool[0]
    prec = preconditioners.MaskedPreconditioner(prob.ainfo, prob.icov_ell[0:1], 0, mask,
                                                prob.minfo, min_pix=100, n_jacobi=1,
                                                lmax_r_ell=prob.lmax)
    g = oracle_geom(prob.minfo)
    olevels = ow.get_levels(mask, g, prob.icov_ell[0:1], 0, min_pix=100,
                            lmax_r_ell=prob.lmax)
    assert len(olevels) == len(prec.levels) >= 2
    for lv, ol in zip(prec.levels, olevels):
        np.testing.assert_array_equal(lv.mask_unobs, ol.mask_unobs)
        np.testing.assert_array_equal(lv.minfo.nphi, ol.geom.nphi)
        np.testing.assert_array_equal(lv.minfo.weight, ol.geom.weight)
        np.testing.assert_allclose(lv.d_ell, ol.d_ell, rtol=1e-15)
        if lv.pinv_g is None:
            np.testing.assert_allclose(lv.m_prec, ol.m_prec, rtol=1e-15)
        else:
            assert rel(lv.dense_g(), ol.dense_g()) < 1e-12
            print(f'coarsest pinv(G) GPU vs oracle: {rel(lv.pinv_g, ol.pinv_g):.3e}',
                  flush=True)
    oprec = ow.MaskedPreconditioner(prob.lmax, prob.icov_ell[0:1], 0, mask, g,
                                    lmax_r_ell=prob.lmax, levels=olevels)
    rng = np.random.default_rng(6)
    v = rng.normal(size=(1, prob.ainfo.nelem)) + 1j * rng.normal(size=(1, prob.ainfo.nelem))
    v[:, :prob.lmax + 1] = v[:, :prob.lmax + 1].real
    got = prec(torch.as_tensor(v, device=dev))
    err = rel(got, oprec(v))
    print(f'multigrid application GPU vs oracle: {err:.3e}', flush=True)
    assert err < 1e-10
    # sliced (1-d) input as used through add_preconditioner(sel=np.s_[0])
    got1 = prec(torch.as_tensor(v[0], device=dev))
    assert rel(got1, oprec(v)[0]) < 1e-10
