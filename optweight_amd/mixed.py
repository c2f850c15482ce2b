'''
Mixed-precision Wiener solve: float32 operators and Krylov vectors with a
float64 solution and float64 residual replacement ("reliable updates",
Sleijpen & van der Vorst 1996; the lattice-QCD mixed-precision CG of Clark
et al. 2010).

One iteration is the pixell CG step of optweight's solver (solvers.py
:127-141) with A and M applied in float32:
    q = A32 p;  alpha = <r,z>/<p,q>;  x64 += alpha p;  r32 -= alpha q;
    z = M32 r32;  beta = <r,z>'/<r,z>;  p = z + beta p
All inner products accumulate in float64 on the device. Whenever the
recursive residual has dropped by ``delta`` since the last replacement (or
reached the target), the true residual r64 = b - A64 x64 is formed with the
float64 operator and replaces r32; the solve stops when that float64
residual satisfies ||r64|| <= tol ||b||. The solution therefore meets the
same float64 residual criterion as the plain float64 solve.
'''
import numpy as np
import torch

from optweight_amd import alm_utils, _lib


def _convert(src, dtype):
    out = torch.empty(src.shape, dtype=dtype, device=src.device)
    _lib.call('owmi_convert', _lib.dtype_code(src), _lib.ptr(src), _lib.dtype_code(out),
              _lib.ptr(out), src.numel(), _lib.stream_ptr(src.device))
    return out


class MixedSolveInfo:
    def __init__(self, iterations, replacements, residual):
        self.iterations = iterations
        self.replacements = replacements
        self.residual = residual


def solve_mixed(a_matrix, b, precond=None, tol=1e-6, delta=0.1, maxiter=1000,
                low=torch.complex64):
    '''Solve A x = b (b complex128 device alm) to float64 true relative
    residual tol. a_matrix and precond must accept both complex128 and
    ``low`` inputs. Returns (x complex128, MixedSolveInfo).'''
    if precond is None:
        precond = lambda v: v.clone()
    dot = alm_utils.contract_almxblm_dev
    dev = b.device
    st = _lib.stream_ptr(dev)
    code_lo = _lib.dtype_code(torch.empty(0, dtype=low))
    x = torch.zeros_like(b)
    bn2 = float(dot(b, b).item())
    target = tol * tol * bn2
    r = _convert(b, low)
    z = precond(r)
    rz = dot(r, z)
    p = z
    ref2 = bn2
    rr = torch.empty((), dtype=torch.float64, device=dev)
    it = nrep = 0
    res = 1.0
    while it < maxiter:
        q = a_matrix(p)
        pq = dot(p, q)
        _lib.call('owmi_cg_xr_mixed', _lib.F64, _lib.ptr(x), code_lo, _lib.ptr(r), _lib.ptr(p),
                  _lib.ptr(q), _lib.ptr(rz), _lib.ptr(pq), x.numel(), st)
        it += 1
        dot(r, r, out=rr)
        rr_h = float(rr.item())
        if rr_h <= delta * delta * ref2 or rr_h <= target:
            r64 = a_matrix(x)
            _lib.call('owmi_axpby', _lib.F64, 1.0, _lib.ptr(b), -1.0, _lib.ptr(r64), b.numel(), st)
            rr64 = float(dot(r64, r64).item())
            nrep += 1
            res = np.sqrt(rr64 / bn2)
            if rr64 <= target:
                break
            r = _convert(r64, low)
            ref2 = rr64
        z = precond(r)
        rz_new = dot(r, z)
        _lib.call('owmi_cg_p', code_lo, _lib.ptr(p), _lib.ptr(z), _lib.ptr(rz_new), _lib.ptr(rz),
                  p.numel(), st)
        rz = rz_new
    return x, MixedSolveInfo(it, nrep, res)


def solve_fp64(a_matrix, b, precond=None, tol=1e-6, maxiter=1000):
    '''Plain float64 preconditioned CG (the same recurrence, all float64),
    stopping on the recursive residual ||r|| <= tol ||b||.'''
    if precond is None:
        precond = lambda v: v.clone()
    dot = alm_utils.contract_almxblm_dev
    dev = b.device
    st = _lib.stream_ptr(dev)
    x = torch.zeros_like(b)
    bn2 = float(dot(b, b).item())
    target = tol * tol * bn2
    r = b.clone()
    z = precond(r)
    rz = dot(r, z)
    p = z
    rr = torch.empty((), dtype=torch.float64, device=dev)
    it = 0
    while it < maxiter:
        q = a_matrix(p)
        pq = dot(p, q)
        _lib.call('owmi_cg_xr', _lib.F64, _lib.ptr(x), _lib.ptr(r), _lib.ptr(p), _lib.ptr(q),
                  _lib.ptr(rz), _lib.ptr(pq), x.numel(), st)
        it += 1
        dot(r, r, out=rr)
        if float(rr.item()) <= target:
            break
        z = precond(r)
        rz_new = dot(r, z)
        _lib.call('owmi_cg_p', _lib.F64, _lib.ptr(p), _lib.ptr(z), _lib.ptr(rz_new), _lib.ptr(rz),
                  p.numel(), st)
        rz = rz_new
    return x, it
