'''
Differentiable Wiener solve: torch.autograd.Functions whose backward is an
adjoint solve through the same symmetric operator and the adjoint
transforms. No counterpart in the reference; built on its solver
(solvers.CGWienerMap, solvers.py:8-456).

Model: data d (map), signal covariance S(a) = C_ell * a_k(l) with band
amplitudes a (band k(l) of each multipole), P = M Y B, noise N:
    b(d) = P^T N^-1 d,   A(a) = S(a)^-1 + P^T N^-1 P,   x = A(a)^-1 b.

Inner products: alm vectors use the real-field product
<u, v>_A = sum_{m=0} Re(u v*) + 2 sum_{m>0} Re(u v*) (contract_almxblm), for
which A is symmetric and Y^T is the adjoint of Y. torch's gradient of a real
loss with respect to a complex tensor z is g with dL = sum Re(conj(g) dz),
i.e. dL = <C^-1 g, dz>_A with C = 1 (m = 0), 2 (m > 0). Hence, for
g = dL/dx:
    y = A^-1 C^-1 g                          (adjoint solve, float64 CG)
    dL/db = C y
    dL/da_k = <y, (S^-1 restricted to band k) x>_A / a_k
    dL/dd = N^-1 P C^-1 (dL/db)               (backward of b(d))
'''
import numpy as np
import torch

from optweight_amd import solvers, operators, preconditioners, mixed


def real_field_weight(ainfo, device):
    '''C: 1 for m = 0, 2 for m > 0, per alm index (float64 device).'''
    _, em = ainfo.lm_arrays()
    return torch.as_tensor(np.where(em == 0, 1.0, 2.0), device=device)


def contract(a, b, weight):
    '''<a, b>_A as a differentiable torch scalar (weight from
    real_field_weight).'''
    return torch.sum(weight * (a * b.conj()).real)


class WienerModel:
    '''Operators of the differentiable solve for one data geometry.

    band_edges: increasing multipoles; band k covers
    band_edges[k] <= l < band_edges[k + 1]. Multipoles outside every band
    keep amplitude 1.'''

    def __init__(self, minfo, ainfo, icov_ell, icov_pix, band_edges, b_ell=None, mask_pix=None,
                 spin=None, tol=1e-10, device=None):
        if device is None:
            device = torch.device('cuda', torch.cuda.current_device())
        self.device = device
        self.minfo = minfo
        self.ainfo = ainfo
        self.icov_ell = mat_full(icov_ell)
        self.icov_pix = np.asarray(icov_pix)
        self.b_ell = b_ell
        self.tol = tol
        npol = self.icov_ell.shape[0]
        self.spin = spin
        zero = np.zeros((npol, minfo.npix))
        self.base = solvers.CGWienerMap.from_arrays(zero, minfo, ainfo, self.icov_ell,
                                                    self.icov_pix, b_ell=b_ell,
                                                    mask_pix=mask_pix, spin=spin,
                                                    device=device)
        self.spin = self.base.sht.spin
        edges = np.asarray(band_edges)
        ells = np.arange(ainfo.lmax + 1)
        self.band = np.full(ainfo.lmax + 1, -1)
        for k in range(edges.size - 1):
            self.band[(ells >= edges[k]) & (ells < edges[k + 1])] = k
        self.nband = edges.size - 1
        ell, _ = ainfo.lm_arrays()
        self.ell_of = torch.as_tensor(ell, device=device)
        self.band_of_ell = torch.as_tensor(self.band, device=device)
        self.weight = real_field_weight(ainfo, device)

    def scaled_icov(self, amps):
        '''S(a)^-1 = C_ell^-1 / a_k(l), (npol, npol, nell).'''
        scale = np.ones(self.ainfo.lmax + 1)
        sel = self.band >= 0
        scale[sel] = 1.0 / np.asarray(amps, dtype=np.float64)[self.band[sel]]
        return self.icov_ell * scale

    def set_amps(self, amps):
        icov = self.scaled_icov(amps)
        self.base.icov_signal = operators.EllMatVecAlm(self.ainfo, icov)
        self.base._fast = None
        self.prec = preconditioners.PseudoInvPreconditioner(
            self.ainfo, icov, self.icov_pix, self.minfo, self.spin, b_ell=self.b_ell)
        return icov

    def solve(self, b, amps):
        self.set_amps(amps)
        x, _ = mixed.solve_fp64(self.base.a_matrix, b, self.prec, tol=self.tol)
        return x

    def b_vec(self, d):
        return self.base.get_b_vec(d)

    def b_vec_adjoint(self, v):
        '''N^-1 P v: the transpose of b(d) (map out).'''
        return self.base.icov_noise(self.base.proj(v))

    def band_gradient(self, y, x, amps):
        '''<y, (S^-1 restricted to band k) x>_A / a_k for every band.'''
        icov = self.scaled_icov(amps)
        v = operators.EllMatVecAlm(self.ainfo, icov)(x)
        per = (self.weight * (y * v.conj()).real)
        per = per.reshape(-1, per.shape[-1]).sum(0)
        tot = torch.zeros(self.ainfo.lmax + 1, dtype=torch.float64, device=self.device)
        tot.index_add_(0, self.ell_of, per)
        out = torch.zeros(self.nband, dtype=torch.float64, device=self.device)
        sel = self.band_of_ell >= 0
        out.index_add_(0, self.band_of_ell[sel], tot[sel])
        return out / torch.as_tensor(np.asarray(amps, dtype=np.float64), device=self.device)


def mat_full(icov_ell):
    icov_ell = np.asarray(icov_ell)
    if icov_ell.ndim == 1:
        icov_ell = icov_ell[np.newaxis]
    if icov_ell.ndim == 2:
        icov_ell = icov_ell[:, np.newaxis, :] * np.eye(icov_ell.shape[0])[:, :, np.newaxis]
    return icov_ell


class BVec(torch.autograd.Function):
    '''b = P^T N^-1 d (alm) from a data map d.'''

    @staticmethod
    def forward(ctx, d, model):
        ctx.model = model
        return model.b_vec(d.detach().contiguous())

    @staticmethod
    def backward(ctx, g):
        model = ctx.model
        return model.b_vec_adjoint((g / model.weight).contiguous()), None


class WienerSolve(torch.autograd.Function):
    '''x = A(a)^-1 b, float64 CG to the model's tolerance; the backward is
    one adjoint solve.'''

    @staticmethod
    def forward(ctx, b, amps, model):
        a = amps.detach().cpu().numpy()
        x = model.solve(b.detach().contiguous(), a)
        ctx.model = model
        ctx.amps = a
        ctx.save_for_backward(x)
        return x

    @staticmethod
    def backward(ctx, g):
        model = ctx.model
        (x,) = ctx.saved_tensors
        model.set_amps(ctx.amps)
        y, _ = mixed.solve_fp64(model.base.a_matrix, (g / model.weight).contiguous(),
                                model.prec, tol=model.tol)
        grad_b = y * model.weight if ctx.needs_input_grad[0] else None
        grad_a = None
        if ctx.needs_input_grad[1]:
            grad_a = model.band_gradient(y, x, ctx.amps).to(torch.float64)
        return grad_b, grad_a, None


def wiener_filter(d, amps, model):
    '''Differentiable Wiener-filtered alm x(d, a) of a data map d.'''
    return WienerSolve.apply(BVec.apply(d, model), amps, model)


def config_loss(d, amps, model):
    '''L = 1/2 d^T (P S P^T + N)^-1 d + 1/2 <x, x>_A with x = x(d, a),
    using (P S P^T + N)^-1 d = N^-1 (d - P x) (Woodbury):
    d^T (P S P^T + N)^-1 d = d^T N^-1 d - <b, x>_A. The inverse noise
    covariance is per-pixel diagonal (npol, npix) or a full per-pixel
    matrix (npol, npol, npix).'''
    ninv = torch.as_tensor(model.icov_pix, device=d.device, dtype=d.dtype)
    full = ninv.ndim == d.ndim + 1 and tuple(ninv.shape[:2]) == (d.shape[-2], d.shape[-2]) \
        and ninv.shape[-1] == d.shape[-1]
    if not full and tuple(ninv.shape) != tuple(d.shape):
        raise ValueError(f'icov_pix of shape {tuple(ninv.shape)} does not match data '
                         f'{tuple(d.shape)}: need (npol, npix) or (npol, npol, npix)')
    b = BVec.apply(d, model)
    x = WienerSolve.apply(b, amps, model)
    # full matrix: sum_p d_i N^-1_ij d_j
    quad = torch.einsum('ip,ijp,jp->', d, ninv, d) if full else torch.sum(d * ninv * d)
    return 0.5 * (quad - contract(b, x, model.weight)) + 0.5 * contract(x, x, model.weight)
