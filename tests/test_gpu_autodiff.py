'''
Differentiable Wiener solve (optweight_amd.autodiff) on a small TQU problem
of the benchmark's construction, float64:
  * the loss L = 1/2 d^T (P S P^T + N)^-1 d + 1/2 <x, x> differentiated by
    the adjoint-solve backward agrees with central finite differences of
    the GPU forward (step 1e-6, relative 1e-5) along random directions in
    the data map and in the band amplitudes;
  * closed forms of the first term: dL1/dd = N^-1 (d - P x) and
    dL1/da_k = -1/2 <x, S_k^-1 x> / a_k (x the Wiener solution), relative
    1e-8;
  * the oracle's loss (CPU float64 solver to 1e-13) differentiated by
    central finite differences matches the GPU gradient (relative 1e-5).
'''
import numpy as np
import pytest
import torch

from oracle import geometry as og, sht as osht, wiener as ow, ops as oops
from optweight_amd import workloads, autodiff

pytestmark = pytest.mark.gpu


def oracle_synth(a, minfo, lmax, spin):
    g = og.Geom(minfo.theta, minfo.weight, minfo.nphi, phi0=minfo.phi0)
    out = np.zeros((a.shape[0], minfo.npix))
    osht.alm2map(a, out, lmax, g, spin)
    return out


@pytest.fixture(scope='module')
def setup(dev):
    lmax = 16
    prob = workloads.polarized(lmax, seed=6, ndisk=8, synth=oracle_synth)
    edges = np.array([2, 4, 7, 11, 17])
    model = autodiff.WienerModel(prob.minfo, prob.ainfo, prob.icov_ell, prob.icov_pix, edges,
                                 tol=1e-13, device=dev)
    amps0 = np.array([1.0, 0.8, 1.3, 0.9])
    return prob, model, amps0


def loss_value(model, d_np, amps_np, dev):
    d = torch.as_tensor(d_np, device=dev)
    a = torch.as_tensor(amps_np, device=dev)
    return float(autodiff.config_loss(d, a, model).item())


def test_grad_vs_finite_differences(dev, setup):
    prob, model, amps0 = setup
    d = torch.as_tensor(prob.imap, device=dev).requires_grad_()
    a = torch.as_tensor(amps0, device=dev).requires_grad_()
    L = autodiff.config_loss(d, a, model)
    L.backward()
    gd = d.grad.cpu().numpy()
    ga = a.grad.cpu().numpy()
    rng = np.random.default_rng(1)
    ed = rng.normal(size=prob.imap.shape) * (prob.icov_pix > 0)
    ed *= np.linalg.norm(prob.imap) / np.linalg.norm(ed)
    h = 1e-6
    fd = (loss_value(model, prob.imap + h * ed, amps0, dev)
          - loss_value(model, prob.imap - h * ed, amps0, dev)) / (2 * h)
    an = float(np.sum(gd * ed))
    print(f'data direction: autograd {an:.10e} fd {fd:.10e}')
    assert abs(an - fd) <= 1e-5 * abs(fd)
    ea = rng.normal(size=amps0.shape)
    fd = (loss_value(model, prob.imap, amps0 + h * ea, dev)
          - loss_value(model, prob.imap, amps0 - h * ea, dev)) / (2 * h)
    an = float(np.sum(ga * ea))
    print(f'amplitude direction: autograd {an:.10e} fd {fd:.10e}')
    assert abs(an - fd) <= 1e-5 * abs(fd)


def test_first_term_closed_forms(dev, setup):
    prob, model, amps0 = setup
    d = torch.as_tensor(prob.imap, device=dev).requires_grad_()
    a = torch.as_tensor(amps0, device=dev).requires_grad_()
    b = autodiff.BVec.apply(d, model)
    x = autodiff.WienerSolve.apply(b, a, model)
    ninv = torch.as_tensor(prob.icov_pix, device=dev)
    L1 = 0.5 * (torch.sum(d * ninv * d) - autodiff.contract(b, x, model.weight))
    L1.backward()
    xs = x.detach()
    exp_d = ninv * (d.detach() - model.base.proj(xs))
    err = (torch.linalg.vector_norm(d.grad - exp_d) / torch.linalg.vector_norm(exp_d)).item()
    assert err < 1e-8, err
    icov = model.scaled_icov(amps0)
    exp_a = []
    for k in range(model.nband):
        sel = (model.band == k).astype(float)
        from optweight_amd import operators
        v = operators.EllMatVecAlm(prob.ainfo, icov * sel)(xs)
        exp_a.append(-0.5 * autodiff.contract(xs, v, model.weight).item() / amps0[k])
    np.testing.assert_allclose(a.grad.cpu().numpy(), exp_a, rtol=1e-8)


def oracle_loss(prob, amps, d):
    g = og.Geom(prob.minfo.theta, prob.minfo.weight, prob.minfo.nphi, phi0=prob.minfo.phi0)
    edges = np.array([2, 4, 7, 11, 17])
    scale = np.ones(prob.lmax + 1)
    for k in range(len(amps)):
        scale[edges[k]:edges[k + 1]] = 1 / amps[k]
    icov = oops.full_matrix(prob.icov_ell) * scale
    sol = ow.WienerMap.from_arrays(d, g, prob.lmax, icov, prob.icov_pix)
    sol.add_preconditioner(ow.PseudoInvPreconditioner(prob.lmax, icov, prob.icov_pix, g,
                                                      prob.spin))
    cg = sol.init_solver()
    bn = np.sqrt(oops.contract_almxblm(sol.b0, sol.b0))
    for _ in range(200):
        cg.step()
        if sol.get_residual() / bn < 1e-13:
            break
    x = cg.x
    return 0.5 * (np.sum(d * prob.icov_pix * d) - oops.contract_almxblm(sol.b0, x)) \
        + 0.5 * oops.contract_almxblm(x, x)


def test_grad_vs_oracle(dev, setup):
    prob, model, amps0 = setup
    d = torch.as_tensor(prob.imap, device=dev).requires_grad_()
    a = torch.as_tensor(amps0, device=dev).requires_grad_()
    autodiff.config_loss(d, a, model).backward()
    rng = np.random.default_rng(2)
    ed = rng.normal(size=prob.imap.shape) * (prob.icov_pix > 0)
    ed *= np.linalg.norm(prob.imap) / np.linalg.norm(ed)
    ea = rng.normal(size=amps0.shape)
    h = 1e-6
    fd = (oracle_loss(prob, amps0 + h * ea, prob.imap + h * ed)
          - oracle_loss(prob, amps0 - h * ea, prob.imap - h * ed)) / (2 * h)
    an = float(np.sum(d.grad.cpu().numpy() * ed) + np.sum(a.grad.cpu().numpy() * ea))
    print(f'joint direction: autograd {an:.10e} oracle fd {fd:.10e}')
    assert abs(an - fd) <= 1e-5 * abs(fd)


def test_batched_forward(dev, setup):
    '''A batch of data maps gives the single-map Wiener solutions.'''
    prob, model, amps0 = setup
    d2 = np.stack([prob.imap, 0.5 * prob.imap[::-1].copy()])
    a = torch.as_tensor(amps0, device=dev)
    xb = autodiff.wiener_filter(torch.as_tensor(d2, device=dev), a, model)
    for i in range(2):
        x1 = autodiff.wiener_filter(torch.as_tensor(d2[i], device=dev), a, model)
        assert (torch.linalg.vector_norm(xb[i] - x1) / torch.linalg.vector_norm(x1)).item() \
            < 1e-10


def test_full_matrix_noise(dev, setup):
    '''icov_pix as a full (npol, npol, npix) matrix with a T-Q correlation:
    the loss's quadratic term is sum_p d_i N^-1_ij d_j and dL/dd of the first
    term is N^-1 (d - P x) with the full matrix (relative 1e-10 / 1e-8).'''
    from optweight_amd import operators, preconditioners, alm_utils
    prob, _, amps0 = setup
    npix = prob.minfo.npix
    diag = prob.icov_pix
    full = np.zeros((3, 3, npix))
    for i in range(3):
        full[i, i] = diag[i]
    full[0, 1] = full[1, 0] = 0.3 * np.sqrt(diag[0] * diag[1])
    itau = np.array([np.sum(diag[i]) / np.sum(diag[i] > 0) for i in range(3)])

    class FullModel(autodiff.WienerModel):
        def set_amps(self, amps):
            icov = self.scaled_icov(amps)
            self.base.icov_signal = operators.EllMatVecAlm(self.ainfo, icov)
            self.base._fast = None
            self.prec = preconditioners.HarmonicPreconditioner(self.ainfo, icov, itau=itau)
            return icov

    edges = np.array([2, 4, 7, 11, 17])
    model = FullModel(prob.minfo, prob.ainfo, prob.icov_ell, full, edges, tol=1e-13,
                      device=dev)
    d = torch.as_tensor(prob.imap, device=dev).requires_grad_()
    a = torch.as_tensor(amps0, device=dev)
    L = autodiff.config_loss(d, a, model)
    b = model.b_vec(d.detach().contiguous())
    x = model.solve(b, amps0)
    dn = prob.imap
    quad = np.einsum('ip,ijp,jp->', dn, full, dn)
    w = model.weight
    exp = 0.5 * (quad - autodiff.contract(b, x, w).item()) + 0.5 * autodiff.contract(x, x, w).item()
    assert abs(L.item() - exp) <= 1e-10 * abs(exp)
    # a diagonal-only quadratic term differs (the check above is not vacuous)
    assert abs(np.sum(dn * diag * dn) - quad) > 1e-6 * abs(quad)
    d2 = torch.as_tensor(prob.imap, device=dev).requires_grad_()
    b2 = autodiff.BVec.apply(d2, model)
    x2 = autodiff.WienerSolve.apply(b2, a, model)
    ninv = torch.as_tensor(full, device=dev)
    L1 = 0.5 * (torch.einsum('ip,ijp,jp->', d2, ninv, d2) - autodiff.contract(b2, x2, w))
    L1.backward()
    res = d2.detach() - model.base.proj(x2.detach())
    exp_d = torch.einsum('ijp,jp->ip', ninv, res)
    err = (torch.linalg.vector_norm(d2.grad - exp_d) / torch.linalg.vector_norm(exp_d)).item()
    assert err < 1e-8, err
    with pytest.raises(ValueError):
        autodiff.config_loss(d.detach()[:2], a, model)
