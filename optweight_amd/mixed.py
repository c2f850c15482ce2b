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
import inspect
import os

import numpy as np
import torch

from optweight_amd import alm_utils, _lib


def _convert(src, dtype, out=None):
    if out is None:
        out = torch.empty(src.shape, dtype=dtype, device=src.device)
    elif tuple(out.shape) != tuple(src.shape) or out.dtype != dtype or not out.is_contiguous():
        raise ValueError('conversion output must be contiguous, of the input shape and dtype')
    _lib.call('owmi_convert', _lib.dtype_code(src), _lib.ptr(src), _lib.dtype_code(out),
              _lib.ptr(out), src.numel(), _lib.stream_ptr(src.device))
    return out


class SolveInfo:
    '''iterations: CG iterations of the batch; iterations_rhs: per
    right-hand side; replacements: float64 residual replacements;
    residual: final float64 true relative residual per right-hand side
    (None when not computed); graphs: whether the iterations were replayed
    from recorded HIP graphs.'''

    def __init__(self, iterations, iterations_rhs, replacements, residual, graphs=False):
        self.iterations = iterations
        self.iterations_rhs = iterations_rhs
        self.replacements = replacements
        self.residual = residual
        self.graphs = graphs

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


def _xr_rr(x, r, p, q, rz, pq, rr, nb):
    '''_xr with <r, r> of the new residuals into rr (one pass over r).'''
    nelem = r.shape[-1]
    lmax = alm_utils.getlmax(nelem)
    _lib.call('owmi_cg_xr_rr_batch', _lib.dtype_code(x), _lib.ptr(x), _lib.dtype_code(r),
              _lib.ptr(r), _lib.ptr(p), _lib.ptr(q), _lib.ptr(rz), _lib.ptr(pq), lmax,
              r.numel() // (nb * nelem), nb, _lib.ptr(rr), _lib.stream_ptr(r.device))


def _pupd(p, z, rz_new, rz, nb):
    _lib.call('owmi_cg_p_batch', _lib.dtype_code(p), _lib.ptr(p), _lib.ptr(z), _lib.ptr(rz_new),
              _lib.ptr(rz), p.numel() // nb, nb, _lib.stream_ptr(p.device))


def _identity(v):
    return v.clone()


def _dot(a, b, out=None):
    return alm_utils.contract_almxblm_batch(a, b, out=out)


def _takes_dot_out(op):
    '''Whether the operator callable accepts dot_out= (it then writes the
    real-field <p, A p> of every right-hand side itself, as
    CGWienerMap.a_matrix does).'''
    try:
        return 'dot_out' in inspect.signature(op).parameters
    except (TypeError, ValueError):
        return False


def _relative(rr, bb):
    '''sqrt(rr / bb) per right-hand side; 0 for b = 0 (solved by x = 0).'''
    return np.sqrt(np.divide(rr, bb, out=np.zeros_like(rr), where=bb > 0))


class _State:
    '''One batch of right-hand sides of a batched CG solve. An iteration is
    phase 0 on the device (q = A p, <p, q>, x += alpha p, r -= alpha q, and
    <r, r> into ``rr``), the host's look at <r, r> (``check``: convergence,
    residual replacement, freezing), then phase 1 on the device (z = M r,
    <r, z>, p = z + beta p). Phases only update the state's buffers in
    place, so they can be recorded into HIP graphs and replayed, also in
    later solves: ``reset`` starts a new solve in the same buffers.'''

    def __init__(self, b3, rr, tol, low):
        self.nb = b3.shape[0]
        self.dev = b3.device
        self.tol = tol
        self.rr = rr
        self.pq = torch.empty((self.nb,), dtype=torch.float64, device=self.dev)
        self.rzn = torch.empty_like(self.pq)
        self.rz = torch.empty_like(self.pq)
        self.x = torch.zeros_like(b3)
        self.r = torch.empty(b3.shape, dtype=low, device=self.dev)
        self.p = torch.empty_like(self.r)
        self.shape = tuple(b3.shape)

    def reset(self, b3):
        '''A new solve of b3 (the shape of the first): x = 0, r = b3,
        p = z = M r, rz = <r, z>.'''
        if tuple(b3.shape) != self.shape or b3.dtype != self.x.dtype:
            raise ValueError(f'right-hand sides {tuple(b3.shape)} {b3.dtype}: this solve '
                             f'was set up for {self.shape} {self.x.dtype}')
        self.b3 = b3
        self.bn2 = _dot(b3, b3).cpu().numpy()
        self.target = self.tol * self.tol * self.bn2
        self.active = np.ones(self.nb, dtype=bool)
        self.it = 0
        self.it_rhs = np.zeros(self.nb, dtype=np.int64)
        self.x.zero_()
        if self.r.dtype == b3.dtype:
            self.r.copy_(b3)
        else:
            _convert(b3, self.r.dtype, out=self.r)
        self.p.copy_(self.m_op(self.r))
        _dot(self.r, self.p, out=self.rz)

    def phase(self, k):
        if k == 0:
            if getattr(self, '_a_dot', None) is None:
                self._a_dot = _takes_dot_out(self.a_op)
            if self._a_dot:
                q = self.a_op(self.p, dot_out=self.pq)  # <p, A p> in A's last pass
            else:
                q = self.a_op(self.p)
                _dot(self.p, q, out=self.pq)
            _xr_rr(self.x, self.r, self.p, q, self.rz, self.pq, self.rr, self.nb)
        else:
            z = self.m_op(self.r)
            _dot(self.r, z, out=self.rzn)
            _pupd(self.p, z, self.rzn, self.rz, self.nb)
            self.rz.copy_(self.rzn)

    def _count(self, rr_h):
        self.it += 1
        self.it_rhs[self.active] += 1
        bad = np.flatnonzero(self.active & ~np.isfinite(rr_h))
        if bad.size:
            raise FloatingPointError(f'CG: non-finite residual norm of right-hand sides '
                                     f'{bad.tolist()} at iteration {self.it}')

    def _freeze(self, idx):
        '''Converged right-hand sides idx: residual and search direction
        cleared, so every later coefficient of theirs is 0.'''
        done = torch.as_tensor(idx, device=self.dev)
        self.r[done] = 0
        self.p[done] = 0


class _Fp64State(_State):
    '''Float64 CG; a right-hand side stops on its recursive residual
    ||r|| <= tol ||b||.'''

    def __init__(self, a_matrix, b3, precond, tol, rr):
        super().__init__(b3, rr, tol, b3.dtype)
        self.a_op = a_matrix
        self.m_op = precond

    def check(self, rr_h):
        self._count(rr_h)
        newly = self.active & (rr_h <= self.target)
        if newly.any():
            self.active &= ~newly
            if self.active.any():
                self._freeze(np.flatnonzero(newly))
        return bool(self.active.any())

    def info(self, verify=False):
        res = None
        if verify:
            r64 = self.b3 - self.a_op(self.x)
            res = _relative(_dot(r64, r64).cpu().numpy(), self.bn2)
        return SolveInfo(self.it, self.it_rhs.copy(), 0, res)


class _MixedState(_State):
    '''Float32 (``low``) operators and Krylov vectors, float64 solution;
    a right-hand side whose recursive residual dropped by delta since its
    last replacement (or reached the target) gets its float64 true residual
    b - A64 x, which replaces r and decides convergence.'''

    def __init__(self, a_matrix, b3, precond, tol, delta, low, a_lo, m_lo, rr):
        super().__init__(b3, rr, tol, low)
        self.a64 = a_matrix
        self.a_op = a_lo
        self.m_op = m_lo
        self.low = low
        self.delta2 = delta * delta

    def reset(self, b3):
        super().reset(b3)
        self.ref2 = self.bn2.copy()
        self.nrep = 0
        self.res = np.full(self.nb, np.nan)

    def check(self, rr_h):
        self._count(rr_h)
        trig = self.active & ((rr_h <= self.delta2 * self.ref2) | (rr_h <= self.target))
        if not trig.any():
            return True
        idx = np.flatnonzero(trig)
        sel = torch.as_tensor(idx, device=self.dev)
        r64 = self.b3[sel] - self.a64(self.x[sel].contiguous())
        rr64 = _dot(r64, r64).cpu().numpy()
        self.nrep += 1
        self.res[idx] = _relative(rr64, self.bn2[idx])
        conv = rr64 <= self.target[idx]
        self.active[idx[conv]] = False
        if not self.active.any():
            return False
        self.r[sel] = _convert(r64, self.low)
        self.ref2[idx] = rr64
        if conv.any():
            self._freeze(idx[conv])
        return True

    def info(self):
        return SolveInfo(self.it, self.it_rhs.copy(), self.nrep, self.res.copy())


class _Runner:
    '''Runs phase k of the running states. One state runs on the current
    stream; several run as a fork-join, each on a stream of its own (default
    priority) that waits for the current stream, which then waits for all
    of them, so their kernels overlap. With use_graph the first run of each
    (phase, running set) is eager (it sizes every workspace and cache on the
    streams it uses), the second records it, fork-join included, into one
    HIP graph (one memory pool shared by all graphs of the runner: the
    phases never run concurrently and free their temporaries), and later
    runs replay that graph. A recording failure raises.'''

    def __init__(self, states, use_graph):
        self.states = states
        self.use_graph = use_graph
        dev = states[0].dev
        self.streams = ([torch.cuda.Stream(device=dev) for _ in states] if len(states) > 1
                        else None)
        self.calls = {}
        self.graphs = {}
        self.pool = None

    def _eager(self, k, running):
        if self.streams is None:
            for i in running:
                self.states[i].phase(k)
            return
        main = torch.cuda.current_stream()
        for i in running:
            side = self.streams[i]
            side.wait_stream(main)
            with torch.cuda.stream(side):
                self.states[i].phase(k)
        for i in running:
            main.wait_stream(self.streams[i])

    def run(self, k, running):
        key = (k, tuple(running))
        self.calls[key] = self.calls.get(key, 0) + 1
        if not self.use_graph or self.calls[key] == 1:
            self._eager(k, running)
            return
        if key not in self.graphs:
            if self.pool is None:
                self.pool = torch.cuda.graph_pool_handle()
            g = torch.cuda.CUDAGraph()
            # no garbage collection inside the recording: a finalizer that
            # frees device memory (an SHT plan of an unreachable geometry)
            # would invalidate it
            gc.collect()
            gc_was_enabled = gc.isenabled()
            gc.disable()
            try:
                with torch.cuda.graph(g, pool=self.pool):
                    self._eager(k, running)
            except RuntimeError as err:
                # an operator that cannot be recorded (a host copy or sync
                # inside it): graphs were asked for, so this is an error, not
                # a silent switch to eager execution
                raise RuntimeError(f'HIP graph recording of CG phase {k} failed (run with '
                                   f'graphs=False for eager execution): {err}') from err
            finally:
                if gc_was_enabled:
                    gc.enable()
            self.graphs[key] = g
        self.graphs[key].replay()

    @property
    def replayed(self):
        return bool(self.graphs)


class Lockstep:
    '''Batched CG solves of several buckets of right-hand sides that iterate
    together (the buckets differ in band limit and operators). Every phase
    of an iteration runs the running buckets at once, each on a stream of
    its own, then the host reads all their <r, r> with one copy, each bucket
    checks convergence, and a bucket leaves once all its right-hand sides
    have converged. Each bucket follows the recurrence and stopping rule of
    its own batched solve (solve_fp64 / solve_mixed), so the results equal
    one solve per bucket.

    buckets: [(a_matrix, b, precond)], b (nbatch, ncomp, nelem) complex128
    (or (ncomp, nelem)) fixing each bucket's shape; mixed precision
    (complex64 operators) unless fp64. The buffers and the HIP graphs of
    the phases (graphs=True: recorded at the second run of each phase and
    running set) persist: ``solve`` of new right-hand sides of the same
    shapes replays them.'''

    def __init__(self, buckets, tol=1e-6, delta=0.01, fp64=False, graphs=True, maxiter=1000):
        if not buckets:
            raise ValueError('no buckets')
        b3s = [_as_batch(b)[0] for _, b, _ in buckets]
        self.squeeze = [b.ndim == 2 for _, b, _ in buckets]
        self.fp64 = fp64
        self.maxiter = maxiter
        self.rr_all = torch.empty((sum(b.shape[0] for b in b3s),), dtype=torch.float64,
                                  device=b3s[0].device)
        self.offsets = np.cumsum([0] + [b.shape[0] for b in b3s])
        self.states = []
        for k, ((a_matrix, _, precond), b3) in enumerate(zip(buckets, b3s)):
            rr = self.rr_all[self.offsets[k]:self.offsets[k + 1]]
            precond = precond or _identity
            if fp64:
                self.states.append(_Fp64State(a_matrix, b3, precond, tol, rr))
            else:
                self.states.append(_MixedState(a_matrix, b3, precond, tol, delta,
                                               torch.complex64, a_matrix, precond, rr))
        self.runner = _Runner(self.states, graphs)
        self._first = [b for _, b, _ in buckets]

    def solve(self, bs=None, verify=True):
        '''Solve the buckets' right-hand sides bs (default: those given at
        construction). Returns [(x, SolveInfo)] in bucket order; x is a view
        of the solver's buffer, overwritten by the next solve. verify: float64
        true residuals of a float64 solve.'''
        bs = self._first if bs is None else bs
        if len(bs) != len(self.states):
            raise ValueError(f'{len(bs)} right-hand side batches for {len(self.states)} buckets')
        for st, b in zip(self.states, bs):
            st.reset(_as_batch(b)[0])
        running = [i for i, st in enumerate(self.states) if st.active.any()]
        it = 0
        while it < self.maxiter and running:
            self.runner.run(0, running)
            it += 1
            rr_h = self.rr_all.cpu().numpy()
            running = [i for i in running
                       if self.states[i].check(rr_h[self.offsets[i]:self.offsets[i + 1]])]
            if running:
                self.runner.run(1, running)
        out = []
        for st, squeeze in zip(self.states, self.squeeze):
            info = st.info(verify) if self.fp64 else st.info()
            info.graphs = self.runner.replayed
            out.append((st.x[0] if squeeze else st.x, info))
        return out


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
    if graphs is None:
        graphs = _graphs_default(b3)
    rr = torch.empty((b3.shape[0],), dtype=torch.float64, device=b3.device)
    st = _MixedState(a_matrix, b3, precond, tol, delta, low, a_matrix_low or a_matrix,
                     precond_low or precond, rr)
    return _solve_one(st, b3, squeeze, graphs, maxiter, st.info)


def solve_fp64(a_matrix, b, precond=None, tol=1e-6, maxiter=1000, verify=False):
    '''Plain float64 preconditioned CG (the same recurrence, all float64),
    stopping each right-hand side on its recursive residual
    ||r|| <= tol ||b||. With verify=True the float64 true residual of every
    right-hand side is computed at the end. Returns (x, SolveInfo).'''
    b3, squeeze = _as_batch(b)
    rr = torch.empty((b3.shape[0],), dtype=torch.float64, device=b3.device)
    st = _Fp64State(a_matrix, b3, precond or _identity, tol, rr)
    return _solve_one(st, b3, squeeze, False, maxiter, lambda: st.info(verify))


def _solve_one(st, b3, squeeze, graphs, maxiter, info_fn):
    '''One state on the current stream: iterate, then (x, SolveInfo).'''
    st.reset(b3)
    runner = _Runner([st], graphs)
    it = 0
    while it < maxiter:
        runner.run(0, [0])
        it += 1
        if not st.check(st.rr.cpu().numpy()):
            break
        runner.run(1, [0])
    info = info_fn()
    info.graphs = runner.replayed
    return (st.x[0] if squeeze else st.x), info


def solve_ragged(buckets, tol=1e-6, delta=0.01, fp64=False, lockstep=True, graphs=None,
                 maxiter=1000):
    """Right-hand sides of uneven band limit. buckets: list of
    (a_matrix, b, precond) with b (nbatch, ncomp, nelem) at the bucket's
    lmax; mixed precision unless fp64. Returns [(x, SolveInfo)] in bucket
    order. lockstep (default): one Lockstep solve of all buckets (graphs
    default True there; for repeated solves keep a Lockstep object, which
    records its graphs once); else one solve per bucket in turn (graphs as
    solve_mixed). Both give the same iterates."""
    if not buckets:
        return []
    if lockstep:
        return Lockstep(buckets, tol=tol, delta=delta, fp64=fp64,
                        graphs=True if graphs is None else graphs, maxiter=maxiter).solve()
    out = []
    for a_matrix, b, precond in buckets:
        if fp64:
            out.append(solve_fp64(a_matrix, b, precond, tol=tol, verify=True, maxiter=maxiter))
        else:
            out.append(solve_mixed(a_matrix, b, precond, tol=tol, delta=delta, graphs=graphs,
                                   maxiter=maxiter))
    return out
