'''
GPU parity of the constant-correlation noise model
(N = M N_pix^1/2 Y C_ell Y^T N_pix^1/2 M, solvers.py:459-575) against the
oracle, float64: the inner-CG inverse (InvPixEllPixMatVecMap, 5 steps) and
its inverse square root to relative L2 1e-9 (rounding carried through the
inner CG recurrences), batched right-hand sides equal to single ones to
1e-12, and the Wiener solver built by from_arrays_const_cor (b vector, A,
a 3-step CG trajectory, the constrained-realisation right-hand side) to
1e-8.
'''
import numpy as np
import pytest
import torch

from oracle import geometry as og, wiener as ow, sht as osht, layout
from optweight_amd import solvers, operators, workloads

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
def prob():
    return workloads.polarized(20, seed=5, ndisk=10, synth=oracle_synth)


def noise_ell(lmax):
    '''Inverse noise spectrum: white plus 1/f-like, T-E correlated.'''
    ell = np.arange(lmax + 1)
    n = 1.0 + (np.maximum(ell, 1) / 8.0) ** -1.5
    cov = np.zeros((3, 3, lmax + 1))
    cov[0, 0] = n
    cov[1, 1] = cov[2, 2] = 2 * n
    cov[0, 1] = cov[1, 0] = 0.3 * n
    return np.stack([np.linalg.inv(cov[:, :, l]) for l in range(lmax + 1)], axis=-1)


@pytest.mark.parametrize('masked', [False, True])
def test_inverse_noise_parity(dev, prob, masked):
    print(f'inverse noise masked={masked}', flush=True)
    g = oracle_geom(prob.minfo)
    inl = noise_ell(prob.lmax)
    mask = prob.mask_bool.astype(np.float64) if masked else None
    op = operators.InvPixEllPixMatVecMap(prob.icov_pix, inl, prob.minfo, prob.spin,
                                         power_m=-0.5, power_x=-1, mask=mask, nsteps=5)
    oop = ow.inv_pix_ell_pix_op(prob.icov_pix, inl, g, prob.spin, power_m=-0.5, power_x=-1,
                                mask=mask, nsteps=5)
    rng = np.random.default_rng(1)
    m = rng.normal(size=(3, prob.minfo.npix))
    mt = torch.as_tensor(m, device=dev)
    got = op(mt)
    assert rel(mt, m) == 0, 'operator modified its input'
    assert rel(got, oop(m)) < 1e-9
    m2 = rng.normal(size=(3, prob.minfo.npix))
    gb = op(torch.as_tensor(np.stack([m, m2]), device=dev))
    assert rel(gb[0], got.cpu().numpy()) < 1e-12
    assert rel(gb[1], op(torch.as_tensor(m2, device=dev)).cpu().numpy()) < 1e-12
    sq = operators.InvSqrtPixEllPixMatVecMap(prob.icov_pix, inl, prob.minfo, prob.spin,
                                             power_m=-0.5, power_x=-1, mask=mask, inv_op=op)
    osq = ow.inv_sqrt_pix_ell_pix_op(prob.icov_pix, inl, g, prob.spin, power_m=-0.5,
                                     power_x=-1, mask=mask, inv_op=oop)
    a = rng.normal(size=(3, layout.nelem(prob.lmax))) * (1 + 1j)
    ell, em = layout.lm_arrays(prob.lmax)
    a[:, em == 0] = a[:, em == 0].real
    a[1:, ell < 2] = 0
    assert rel(sq(torch.as_tensor(a, device=dev)), osq(a)) < 1e-9


@pytest.mark.parametrize('draw_constr', [False, True])
def test_const_cor_solver(dev, prob, draw_constr):
    print(f'const-cor solver draw_constr={draw_constr}', flush=True)
    g = oracle_geom(prob.minfo)
    inl = noise_ell(prob.lmax)
    kw = dict(nsteps=5, mask_pix=prob.mask_bool.astype(np.float64), draw_constr=draw_constr,
              seed=11 if draw_constr else None)
    solver = solvers.CGWienerMap.from_arrays_const_cor(prob.imap, prob.minfo, prob.ainfo,
                                                      prob.icov_ell, prob.icov_pix, inl,
                                                      device=dev, **kw)
    osol = ow.from_arrays_const_cor(prob.imap, g, prob.lmax, prob.icov_ell, prob.icov_pix, inl,
                                    **kw)
    assert rel(solver.b0, osol.b0) < 1e-9
    if draw_constr:
        return
    rng = np.random.default_rng(2)
    v = rng.normal(size=osol.b0.shape) + 1j * rng.normal(size=osol.b0.shape)
    v[:, :prob.lmax + 1] = v[:, :prob.lmax + 1].real
    assert rel(solver.a_matrix(torch.as_tensor(v, device=dev)), osol.a_matrix(v)) < 1e-9
    solver.init_solver()
    cg = osol.init_solver()
    for i in range(3):
        solver.step()
        cg.step()
        assert rel(solver.x, cg.x) < 1e-8, f'step {i}'
