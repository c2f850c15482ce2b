'''
Batched constrained realisations of uneven band limit (two buckets, two
right-hand sides each), float64 operators:
  * the constrained right-hand sides b = P^T N^-1 d + S^-1/2 w + P^T N^-1/2 w'
    built from per-RHS seeds equal the oracle's (relative 1e-12);
  * every right-hand side of the mixed-precision bucketed solve reaches
    float64 true relative residual 1e-6 and agrees with its single-RHS
    float64 solve to relative L2 1e-4;
  * the lockstep solve (all buckets in every phase, one stream per bucket,
    the fork-join phases recorded into HIP graphs and replayed) against one
    solve per bucket: float64 solutions within relative 1e-12 with the same
    iteration counts per right-hand side, mixed precision within 1e-12 with
    the same iteration and replacement counts; with the multigrid
    preconditioner on T inside the graphs too; a Lockstep object replays
    its graphs for new right-hand sides;
  * a non-finite residual norm stops the solve with FloatingPointError.
'''
import numpy as np
import pytest
import torch

from oracle import geometry as og, sht as osht, wiener as ow
from optweight_amd import workloads, solvers, preconditioners, mixed, alm_utils

pytestmark = pytest.mark.gpu


def oracle_synth(a, minfo, lmax, spin):
    g = og.Geom(minfo.theta, minfo.weight, minfo.nphi, phi0=minfo.phi0)
    out = np.zeros((a.shape[0], minfo.npix))
    osht.alm2map(a, out, lmax, g, spin)
    return out


def rel(a, b):
    a = a.detach().cpu().numpy() if isinstance(a, torch.Tensor) else np.asarray(a)
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def build(prob, data, dev, seed):
    sol = solvers.CGWienerMap.from_arrays(data, prob.minfo, prob.ainfo, prob.icov_ell,
                                          prob.icov_pix, draw_constr=True, seed=seed,
                                          device=dev)
    sol.add_preconditioner(preconditioners.PseudoInvPreconditioner(
        prob.ainfo, prob.icov_ell, prob.icov_pix, prob.minfo, prob.spin))
    return sol


def test_ragged_constrained(dev):
    buckets = workloads.ragged(lmaxs=(12, 20), per=2, seed=3, synth=oracle_synth)
    sols, work = [], []
    for k, (prob, data) in enumerate(buckets):
        print(f'bucket lmax {prob.lmax}', flush=True)
        sol = build(prob, data, dev, seed=40 + k)
        g = og.Geom(prob.minfo.theta, prob.minfo.weight, prob.minfo.nphi, phi0=prob.minfo.phi0)
        for i in range(data.shape[0]):
            osol = ow.WienerMap.from_arrays(data[i], g, prob.lmax, prob.icov_ell, prob.icov_pix,
                                            draw_constr=True, seed=[40 + k, i])
            assert rel(sol.b0[i], osol.b0) < 1e-12
        sols.append(sol)
        work.append((sol.a_matrix, sol.b0, sol.preconditioner))
    res = mixed.solve_ragged(work, tol=1e-6)
    for sol, (x, info) in zip(sols, res):
        r = sol.b0 - sol.a_matrix(x)
        rr = alm_utils.contract_almxblm_batch(r, r).cpu().numpy()
        bb = alm_utils.contract_almxblm_batch(sol.b0, sol.b0).cpu().numpy()
        assert np.all(np.sqrt(rr / bb) <= 1e-6)
        for i in range(x.shape[0]):
            x1, _ = mixed.solve_fp64(sol.a_matrix, sol.b0[i].contiguous(), sol.preconditioner,
                                     tol=1e-8)
            assert rel(x[i], x1.cpu().numpy()) < 1e-4


def lockstep_buckets(dev, lmaxs=(12, 20, 31), per=(2, 3, 2), mg=False):
    sols, work = [], []
    for k, (lmax, nb) in enumerate(zip(lmaxs, per)):
        prob, data = workloads.ragged(lmaxs=(lmax,), per=nb, seed=5 + k, device=dev)[0]
        sol = build(prob, data, dev, seed=60 + k)
        if mg:
            sol.add_preconditioner(preconditioners.MaskedPreconditioner(
                prob.ainfo, prob.icov_ell[0:1, 0:1], 0, prob.mask_bool[0], prob.minfo,
                min_pix=100, n_jacobi=1, lmax_r_ell=6000, device=dev), sel=np.s_[0])
        sols.append(sol)
        work.append((sol.a_matrix, sol.b0, sol.preconditioner))
    return sols, work


@pytest.mark.parametrize('mg', [False, True])
def test_lockstep_fp64_equals_per_bucket(dev, mg):
    print(f'lockstep fp64 mg={mg}', flush=True)
    sols, work = lockstep_buckets(dev, mg=mg)
    # a bucket of zero right-hand sides converges at the first check and
    # leaves while the others iterate on
    a0, b0, m0 = work[0]
    work.append((a0, torch.zeros_like(b0), m0))
    ref = mixed.solve_ragged(work, tol=1e-10, fp64=True, lockstep=False)
    got = mixed.solve_ragged(work, tol=1e-10, fp64=True, lockstep=True, graphs=True)
    xz, iz = got.pop()
    ref.pop()
    assert iz.iterations == 1 and not torch.any(xz != 0)
    for (xr, ir), (xg, ig) in zip(ref, got):
        err = rel(xg, xr.cpu().numpy())
        print(f'  lmax bucket: iterations {ig.iterations_rhs.tolist()} rel {err:.2e} '
              f'residual max {np.max(ig.residual):.2e}', flush=True)
        assert ig.graphs and not ir.graphs
        assert list(ig.iterations_rhs) == list(ir.iterations_rhs)
        assert err <= 1e-12
        assert np.all(ig.residual <= 1.1e-10)


def test_lockstep_mixed_equals_per_bucket(dev):
    print('lockstep mixed', flush=True)
    sols, work = lockstep_buckets(dev, mg=True)
    ref = mixed.solve_ragged(work, tol=1e-6, delta=1e-2, lockstep=False, graphs=False)
    got = mixed.solve_ragged(work, tol=1e-6, delta=1e-2, lockstep=True, graphs=True)
    for sol, (xr, ir), (xg, ig) in zip(sols, ref, got):
        err = rel(xg, xr.cpu().numpy())
        print(f'  iterations {ig.iterations} replacements {ig.replacements} rel {err:.2e}',
              flush=True)
        assert ig.graphs
        assert ig.iterations == ir.iterations and ig.replacements == ir.replacements
        assert list(ig.iterations_rhs) == list(ir.iterations_rhs)
        assert err <= 1e-12
        r = sol.b0 - sol.a_matrix(xg)
        rr = alm_utils.contract_almxblm_batch(r, r).cpu().numpy()
        bb = alm_utils.contract_almxblm_batch(sol.b0, sol.b0).cpu().numpy()
        assert np.all(np.sqrt(rr / bb) <= 1e-6)


def test_lockstep_reuse(dev):
    '''A Lockstep object records its graphs in the first solve and replays
    them in later solves of new right-hand sides: 2 b gives exactly 2 x
    (scaling by 2 commutes with every rounding), and reversed batches equal
    their own per-bucket solves to 1e-12.'''
    print('lockstep reuse', flush=True)
    sols, work = lockstep_buckets(dev, mg=True)
    lk = mixed.Lockstep(work, tol=1e-6, delta=1e-2)
    first = [(x.clone(), info) for x, info in lk.solve()]
    doubled = lk.solve([2 * b for _, b, _ in work])
    for (x1, i1), (x2, i2) in zip(first, doubled):
        assert i2.graphs and i2.iterations == i1.iterations
        assert torch.equal(x2, 2 * x1)
    flipped = [b.flip(0).contiguous() for _, b, _ in work]
    got = lk.solve(flipped)
    ref = mixed.solve_ragged([(a, b, m) for (a, _, m), b in zip(work, flipped)], tol=1e-6,
                             delta=1e-2, lockstep=False, graphs=False)
    for (xg, ig), (xr, ir) in zip(got, ref):
        assert ig.iterations == ir.iterations and ig.replacements == ir.replacements
        assert rel(xg, xr.cpu().numpy()) <= 1e-12
    with pytest.raises(ValueError):
        lk.solve([b[:1] for _, b, _ in work])


def test_nonfinite_residual_raises(dev):
    print('non-finite residual', flush=True)
    sols, work = lockstep_buckets(dev, lmaxs=(12, 20), per=(2, 2))
    a_matrix, b, precond = work[1]

    def poisoned(alm):
        out = precond(alm)
        out[1] = float('nan')
        return out
    work[1] = (a_matrix, b, poisoned)
    with pytest.raises(FloatingPointError, match=r'right-hand sides \[1\]'):
        mixed.solve_ragged(work, tol=1e-8, fp64=True)
