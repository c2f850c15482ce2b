'''
Wiener solves of one or many right-hand sides on the device.

``solve_fp64``: the preconditioned CG recurrence of optweight's solver
(pixell CG as used at solvers.py:127-141), everything float64.

``solve_mixed``: float32 operators and Krylov vectors with a float64
solution and float64 residual replacement ("reliable updates", Sleijpen &
van der Vorst 1996; the mixed-precision CG of Clark et al. 2010). One
iteration is the pixell step with A and M applied in float32:
    q = A32 p;  alpha = <r,z>/<p,q>;  x64 += alpha p;  r32 -= alpha q;
    z = M32 r32;  beta = <r,z>'/<r,z>;  p = z + beta p
Inner products accumulate in float64 on the device. Whenever the recursive
residual of a right-hand side has dropped by ``delta`` since its last
replacement (or reached the target), its true residual r64 = b - A64 x64 is
formed with the float64 operator (on the sub-batch of such right-hand
sides) and replaces r32; a
right-hand side stops when that float64 residual satisfies
||r64|| <= tol ||b||, the same criterion the float64 solve is verified
against.

Batches: ``b`` of shape (nbatch, ncomp, nelem) runs nbatch independent
solves through one operator application per iteration; every right-hand
side has its own CG scalars (device vectors). A right-hand side that has
converged is frozen (its residual and search direction are cleared, its
coefficients become 0) while the others continue. An unbatched ``b``
(ncomp, nelem) is a batch of one.

HIP graphs (opt-in: graphs=True or OWMI_GRAPHS=1): one mixed iteration is
two device phases separated by the host's look at the residual norms
(q = A32 p through the x, r update and <r, r>; z = M32 r through the p
update). After one eager iteration (which sizes every workspace and cache),
each phase is recorded once into a HIP graph and replayed, so an iteration
costs two graph launches instead of the operators' many kernel launches
from Python. The replayed kernels are the eager ones on the same buffers
(x, r, p and the CG scalars are updated in place), and the residual
replacements between replays stay eager. On the TQU lmax 1023 batch of 16
the launches already overlap the kernels and the recording costs more than
it saves (49.4 against 50.7 solves/s), hence off by default.
'''
import gc
import os

import numpy as np
import torch

from optweight_amd import alm_utils, _lib


def _convert(src, dtype):
    out = torch.empty(src.shape, dtype=dtype, device=src.device)
    _lib.call('owmi_convert', _lib.dtype_code(src), _lib.ptr(src), _lib.dtype_code(out),
              _lib.ptr(out), src.numel(), _lib.stream_ptr(src.device))
    return out


class SolveInfo:
    '''iterations: CG iterations of the batch; iterations_rhs: per
    right-hand side; replacements: float64 residual replacements;
    residual: final float64 true relative residual per right-hand side
    (None when not computed).'''

    def __init__(self, iterations, iterations_rhs, replacements, residual):
        self.iterations = iterations
        self.iterations_rhs = iterations_rhs
        self.replacements = replacements
        self.residual = residual

    def as_dict(self):
        return {'iterations': self.iterations,
                'iterations_rhs': [int(i) for i in self.iterations_rhs],
                'replacements': self.replacements,
                'residual': None if self.residual is None
                else [float(r) for r in self.residual]}


def _as_batch(b):
    if b.ndim == 3:
        return b, False
    if b.ndim == 2:
        return b[None], True
    raise ValueError(f'b must be (ncomp, nelem) or (nbatch, ncomp, nelem), got {tuple(b.shape)}')


def _xr(x, r, p, q, rz, pq, nb):
    _lib.call('owmi_cg_xr_batch', _lib.dtype_code(x), _lib.ptr(x), _lib.dtype_code(r),
              _lib.ptr(r), _lib.ptr(p), _lib.ptr(q), _lib.ptr(rz), _lib.ptr(pq),
              r.numel() // nb, nb, _lib.stream_ptr(r.device))


def _pupd(p, z, rz_new, rz, nb):
    _lib.call('owmi_cg_p_batch', _lib.dtype_code(p), _lib.ptr(p), _lib.ptr(z), _lib.ptr(rz_new),
              _lib.ptr(rz), p.numel() // nb, nb, _lib.stream_ptr(p.device))


def _identity(v):
    return v.clone()


class _Phases:
    '''Device phases of a CG iteration: the first call of each runs eagerly,
    later calls replay a HIP graph recorded from it (use_graph), else they
    run eagerly every time.'''

    def __init__(self, fns, use_graph):
        self.fns = fns
        self.use_graph = use_graph
        self.graphs = [None] * len(fns)
        self.calls = [0] * len(fns)

    def run(self, k):
        self.calls[k] += 1
        if not self.use_graph or self.calls[k] == 1:
            self.fns[k]()
            return
        if self.graphs[k] is None:
            g = torch.cuda.CUDAGraph()
            # no garbage collection inside the recording: a finalizer that
            # frees device memory (an SHT plan of an unreachable geometry)
            # would invalidate it
            gc.collect()
            gc_was_enabled = gc.isenabled()
            gc.disable()
            try:
                with torch.cuda.graph(g):
                    self.fns[k]()
            except RuntimeError as err:
                # an operator that cannot be recorded (a host copy or sync
                # inside it): graphs were asked for, so this is an error, not
                # a silent switch to eager execution
                raise RuntimeError(f'HIP graph recording of CG phase {k} failed (run with '
                                   f'graphs=False for eager execution): {err}') from err
            finally:
                if gc_was_enabled:
                    gc.enable()
            self.graphs[k] = g
        self.graphs[k].replay()


def _graphs_default(t):
    return t.is_cuda and os.environ.get('OWMI_GRAPHS', '0') == '1'


def solve_mixed(a_matrix, b, precond=None, tol=1e-6, delta=0.1, maxiter=1000,
                low=torch.complex64, a_matrix_low=None, precond_low=None, graphs=None):
    '''Solve A x = b (b complex128 device alm, (ncomp, nelem) or (nbatch,
    ncomp, nelem)) to float64 true relative residual tol per right-hand
    side. a_matrix (and precond) must accept complex128 and ``low`` input,
    or the ``low`` versions are given separately. graphs: replay the
    iteration phases from HIP graphs (default: OWMI_GRAPHS=1 in the
    environment). Returns (x complex128, SolveInfo).'''
    b3, squeeze = _as_batch(b)
    precond = precond or _identity
    a_lo = a_matrix_low or a_matrix
    m_lo = precond_low or precond
    if graphs is None:
        graphs = _graphs_default(b3)
    dot = alm_utils.contract_almxblm_batch
    nb = b3.shape[0]
    dev = b3.device
    x = torch.zeros_like(b3)
    bn2 = dot(b3, b3).cpu().numpy()
    target = tol * tol * bn2
    r = _convert(b3, low)
    p = m_lo(r)
    if p.data_ptr() == r.data_ptr():
        # a preconditioner that returns its input: p is updated in place
        p = p.clone()
    rz = dot(r, p)
    # CG scalars on the device, updated in place
    pq = torch.empty((nb,), dtype=torch.float64, device=dev)
    rr = torch.empty((nb,), dtype=torch.float64, device=dev)
    rzn = torch.empty((nb,), dtype=torch.float64, device=dev)

    def step_a():
        # q = A32 p; x += alpha p; r -= alpha q; <r, r>
        q = a_lo(p)
        dot(p, q, out=pq)
        _xr(x, r, p, q, rz, pq, nb)
        dot(r, r, out=rr)

    def step_m():
        # z = M32 r; p = z + beta p
        z = m_lo(r)
        dot(r, z, out=rzn)
        _pupd(p, z, rzn, rz, nb)
        rz.copy_(rzn)

    phases = _Phases((step_a, step_m), graphs)
    ref2 = bn2.copy()
    active = np.ones(nb, dtype=bool)
    it_rhs = np.zeros(nb, dtype=np.int64)
    res = np.full(nb, np.nan)
    it = nrep = 0
    while it < maxiter and active.any():
        phases.run(0)
        it += 1
        it_rhs[active] += 1
        rr_h = rr.cpu().numpy()
        trig = active & ((rr_h <= delta * delta * ref2) | (rr_h <= target))
        if trig.any():
            # float64 true residual of the triggered right-hand sides only
            idx = np.nonzero(trig)[0]
            it_t = torch.as_tensor(idx, device=dev)
            r64 = b3[it_t] - a_matrix(x[it_t].contiguous())
            rr64 = dot(r64, r64).cpu().numpy()
            nrep += 1
            res[idx] = np.sqrt(rr64 / bn2[idx])
            conv = rr64 <= target[idx]
            active[idx[conv]] = False
            if not active.any():
                break
            r[it_t] = _convert(r64, low)
            del r64
            ref2[idx] = rr64
            if conv.any():
                done = torch.as_tensor(idx[conv], device=dev)
                r[done] = 0
                p[done] = 0
        phases.run(1)
    info = SolveInfo(it, it_rhs, nrep, res)
    # whether the iterations were replayed from recorded graphs
    info.graphs = phases.use_graph and any(g is not None for g in phases.graphs)
    return (x[0] if squeeze else x), info


def solve_fp64(a_matrix, b, precond=None, tol=1e-6, maxiter=1000, verify=False):
    '''Plain float64 preconditioned CG (the same recurrence, all float64),
    stopping each right-hand side on its recursive residual
    ||r|| <= tol ||b||. With verify=True the float64 true residual of every
    right-hand side is computed at the end. Returns (x, SolveInfo).'''
    b3, squeeze = _as_batch(b)
    precond = precond or _identity
    dot = alm_utils.contract_almxblm_batch
    nb = b3.shape[0]
    x = torch.zeros_like(b3)
    bn2 = dot(b3, b3).cpu().numpy()
    target = tol * tol * bn2
    r = b3.clone()
    z = precond(r)
    rz = dot(r, z)
    # p is updated in place: never as the same array as r
    p = z.clone() if z.data_ptr() == r.data_ptr() else z
    rr = torch.empty((nb,), dtype=torch.float64, device=b3.device)
    active = np.ones(nb, dtype=bool)
    it_rhs = np.zeros(nb, dtype=np.int64)
    it = 0
    while it < maxiter and active.any():
        q = a_matrix(p)
        pq = dot(p, q)
        _xr(x, r, p, q, rz, pq, nb)
        del q
        it += 1
        it_rhs[active] += 1
        rr_h = dot(r, r, out=rr).cpu().numpy()
        newly = active & (rr_h <= target)
        if newly.any():
            active &= ~newly
            if not active.any():
                break
            done = torch.as_tensor(~active, device=b3.device)
            r[done] = 0
            p[done] = 0
        z = precond(r)
        rz_new = dot(r, z)
        _pupd(p, z, rz_new, rz, nb)
        del z
        rz = rz_new
    res = None
    if verify:
        r64 = b3 - a_matrix(x)
        res = np.sqrt(dot(r64, r64).cpu().numpy() / bn2)
    info = SolveInfo(it, it_rhs, 0, res)
    return (x[0] if squeeze else x), info


def solve_ragged(buckets, tol=1e-6, delta=0.01, fp64=False):
    """Right-hand sides of uneven band limit: one batched solve per bucket
    (the operators differ per band limit). buckets: list of
    (a_matrix, b, precond) with b (nbatch, ncomp, nelem) at the bucket's
    lmax. Mixed precision unless fp64. Returns [(x, SolveInfo)]."""
    out = []
    for a_matrix, b, precond in buckets:
        if fp64:
            out.append(solve_fp64(a_matrix, b, precond, tol=tol, verify=True))
        else:
            out.append(solve_mixed(a_matrix, b, precond, tol=tol, delta=delta))
    return out
