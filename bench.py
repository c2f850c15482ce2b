'''
Benchmark: Wiener solves/s on the polarized (TQU) lmax=1023 workload,
GL 1024 x 2047 grid, to true relative residual ||b - A x|| / ||b|| <= 1e-6
(float64 operator, real-field alm norm), on one MI355X.

A step is one full solve of the workload (inputs resident in HBM).
Reported: the mixed-precision solver (float32 CG inner solves, float64
iterative refinement) as ``value`` when available, the plain float64 GPU
solve beside it; the roofline of the dominant kernel measured with HIP
events on the launch stream; a CPU baseline timed on a bounded sample.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--lmax 1023]
'''
import argparse
import json
import os
import sys
import time

import numpy as np
import torch

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from optweight_amd import (solvers, preconditioners, workloads, alm_utils, sht,  # noqa: E402
                           mixed, _lib, autotune)

PEAK_FP64_TFLOPS = 78.6
PEAK_FP32_TFLOPS = 157.3
PEAK_HBM_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def build_solver(prob, dev, imap=None, use_mg=True):
    """CGWienerMap of the workload (imap (npol, npix), or (nbatch, npol, npix)
    for a batch of right-hand sides) with the pinv preconditioner on TQU and
    the multigrid preconditioner on T."""
    imap = prob.imap if imap is None else imap
    solver = solvers.CGWienerMap.from_arrays(imap, prob.minfo, prob.ainfo, prob.icov_ell,
                                             prob.icov_pix, device=dev)
    solver.add_preconditioner(preconditioners.PseudoInvPreconditioner(
        prob.ainfo, prob.icov_ell, prob.icov_pix, prob.minfo, prob.spin))
    if use_mg:
        mg = preconditioners.MaskedPreconditioner(
            prob.ainfo, prob.icov_ell[0:1, 0:1], 0, prob.mask_bool[0], prob.minfo,
            min_pix=1000, n_jacobi=1, lmax_r_ell=6000, device=dev)
        solver.add_preconditioner(mg, sel=np.s_[0])
    return solver


def true_residual(solver, x):
    """float64 true relative residual ||b - A x|| / ||b|| per right-hand side."""
    r = solver.b0 - solver.a_matrix(x)
    if r.ndim == 3:
        rr = alm_utils.contract_almxblm_batch(r, r).cpu().numpy()
        bb = alm_utils.contract_almxblm_batch(solver.b0, solver.b0).cpu().numpy()
        return np.sqrt(rr / bb)
    return np.sqrt(alm_utils.contract_almxblm(r, r) / alm_utils.contract_almxblm(solver.b0,
                                                                                 solver.b0))


def run_solver(kind, solver, tol=1e-6, delta=0.1):
    b = solver.b0
    if kind == 'fp64':
        return mixed.solve_fp64(solver.a_matrix, b, solver.preconditioner, tol=tol)
    return mixed.solve_mixed(solver.a_matrix, b, solver.preconditioner, tol=tol, delta=delta)


def time_solves(kind, solver, steps, warmup, delta=0.1, count_flops=False):
    """Warm-up, then ``steps`` timed solves. With count_flops, also the
    useful Legendre flops the timed solves issued, by precision
    (sht.LEGENDRE_FLOPS: N_lm * N_theta * (2 n0 + 8 n2) per direction)."""
    for w in range(warmup):
        x, info = run_solver(kind, solver, delta=delta)
        log(f'{kind} warmup solve {w}: iterations {info.iterations} '
            f'(per RHS {list(info.iterations_rhs)}), replacements {info.replacements}, '
            f'true residual {np.max(true_residual(solver, x)):.3e}')
    torch.cuda.synchronize()
    sht.reset_legendre_flops()
    t0 = time.perf_counter()
    infos = []
    for _ in range(steps):
        x, info = run_solver(kind, solver, delta=delta)
        infos.append(info)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    flops = dict(sht.LEGENDRE_FLOPS)
    if count_flops:
        return x, elapsed, infos, flops
    return x, elapsed, infos


def calibration_copy(dev):
    """One float32 copy of 256 MiB (read once, written once), the first
    dispatch of a counter run: FETCH_SIZE and WRITE_SIZE of the other
    kernels are scaled by the bytes this copy moves over what the counters
    read for it."""
    src = torch.ones(64 << 20, dtype=torch.float32, device=dev)
    dst = torch.empty_like(src)
    dst.copy_(src)
    torch.cuda.synchronize()


def time_kernel(fn, reps=10):
    '''Average duration (ms) of fn() measured with events on the current
    stream.'''
    fn()
    s = torch.cuda.Event(enable_timing=True)
    e = torch.cuda.Event(enable_timing=True)
    s.record()
    for _ in range(reps):
        fn()
    e.record()
    e.synchronize()
    return s.elapsed_time(e) / reps


def legendre_roofline(prob, dev, batch, reps=10):
    """Dominant kernel of the timed (mixed, batched) solve: the float32
    spin-2 Legendre adjoint at the workload's shape (TQU lmax 1023 on the
    1024-ring GL grid, ``batch`` right-hand sides; from 8 right-hand sides
    the matrix-core kernel adj_mfma_kernel, 8 waves per block over the 512
    ring pairs), launched alone through owmi_sht_legendre and timed with HIP
    events on its stream (16 transforms per block from a batch of 16).
    Flops per launch: N_lm * N_theta * 8 * batch, the products the matrix
    cores execute (no padding: (2lam, s(-2)lam) x (S+, s S-) columns)."""
    plan = sht.get_plan(prob.minfo, prob.lmax, dev)
    nelem = prob.ainfo.nelem
    alm = torch.zeros((batch, 3, nelem), dtype=torch.complex64, device=dev)
    st = _lib.stream_ptr(dev)
    # fill the ring-coefficient workspace with a synthesis of unit alm first
    alm[:, 1:] = 1.0
    _lib.check(plan._lib.owmi_sht_legendre(plan.handle, _lib.F32, 2, 0, _lib.ptr(alm[0, 1]),
                                           3 * nelem, batch, st))

    def run():
        _lib.check(plan._lib.owmi_sht_legendre(plan.handle, _lib.F32, 2, 1,
                                               _lib.ptr(alm[0, 1]), 3 * nelem, batch, st))
    ms = time_kernel(run, reps=reps)
    flops = nelem * prob.minfo.nrow * 8 * batch
    achieved = flops / (ms * 1e-3) / 1e12
    # 512 ring pairs: 8 waves per block; 16 transforms per block from a
    # batch of 16 (column tiles CT = 4), else 8
    kernel = ('adj_mfma_kernel<2, 8, 4>' if batch >= 16 else
              'adj_mfma_kernel<2, 8, 2>' if batch >= 8 else 'adj2_kernel<float, 2, 4, 4>')
    return {'kernel': kernel, 'ms': ms, 'flops': flops, 'achieved': achieved}


# HBM bytes per launch of the roofline kernel (adj_mfma_kernel<float, 2, 8,
# 4>, 128 right-hand sides): FETCH_SIZE (x 2.0) + WRITE_SIZE (x 1.0), factors
# from a 256 MiB copy in the same process, measured with `rocprofv3 --pmc` on
# `python bench.py --roofline-only`
# (profiles/adj_mfma2_f32_tqu1023_b128_roofline.md): 2.763 GB + 1.452 GB.
ROOFLINE_TRAFFIC_BYTES = 4.215e9
ROOFLINE_TRAFFIC_SHAPE = (1023, 128)

# right-hand sides per batch: the headline solve, and the batched timings of
# the other modes (config 3: 16 per band limit)
HEADLINE_BATCH = 128
AUX_BATCH = 16


def cpu_model():
    try:
        with open('/proc/cpuinfo') as f:
            for line in f:
                if line.startswith('model name'):
                    return line.split(':', 1)[1].strip()
    except OSError:
        pass
    return 'unknown'


def cpu_baseline_solve(prob, iterations, nstep=5):
    """The oracle's float64 solve of the same workload on the host cores:
    pixell's CG recurrence with the pseudo-inverse on TQU and the multigrid
    preconditioner on T (oracle/wiener.py), transforms with the Legendre
    stage in C/OpenMP on every core this process may use and the ring FFTs
    in scipy (oracle/fastsht.py). After one untimed set-up and init, nstep
    CG iterations are timed one by one; solves/s = 1 / (median iteration
    time x ``iterations``, the float64 GPU solve's iteration count, the
    same recurrence to the same stopping rule). Returns (baseline dict,
    oracle iterate after nstep steps)."""
    from oracle import wiener as ow, geometry as og, fastsht
    fastsht.install()
    try:
        t0 = time.perf_counter()
        g = og.Geom(prob.minfo.theta, prob.minfo.weight, prob.minfo.nphi, phi0=prob.minfo.phi0)
        osol = ow.WienerMap.from_arrays(prob.imap, g, prob.lmax, prob.icov_ell, prob.icov_pix)
        osol.add_preconditioner(ow.PseudoInvPreconditioner(prob.lmax, prob.icov_ell,
                                                           prob.icov_pix, g, prob.spin))
        osol.add_preconditioner(ow.MaskedPreconditioner(
            prob.lmax, prob.icov_ell[0:1, 0:1], 0, prob.mask_bool[0], g, min_pix=1000,
            n_jacobi=1, lmax_r_ell=6000), sel=np.s_[0:1])
        cg = osol.init_solver()
        setup = time.perf_counter() - t0
        times = []
        for _ in range(nstep):
            t1 = time.perf_counter()
            cg.step()
            times.append(time.perf_counter() - t1)
    finally:
        fastsht.uninstall()
    med = float(np.median(times))
    cores = fastsht.threads()
    out = {'value': 1.0 / (med * iterations), 'unit': 'solves/s', 'cores': cores,
           'kind': 'port', 'cpu_model': cpu_model(),
           'iteration_s_median': med, 'iteration_s': times, 'setup_s': setup,
           'sample': (f'oracle float64 CG (pixell recurrence, pinv TQU + multigrid T) on the '
                      f'same TQU lmax {prob.lmax} workload, Legendre stage in C/OpenMP on '
                      f'{cores} threads, ring FFTs in scipy: median of {nstep} timed '
                      f'iterations ({med:.2f}s) x {iterations:.0f} iterations of the float64 '
                      f'solve; set-up ({setup:.1f}s) not included')}
    return out, cg.x


def wavelet_bench(args, dev):
    '''Harmonic-space temperature solve with the wavelet noise model:
    CGWiener.from_arrays_wav + PseudoInvPreconditionerWav, mixed precision
    (float32 operators, float64 residual replacement), FP64-verified.'''
    lmax = args.wavelet_lmax
    t0 = time.perf_counter()
    wp = workloads.wavelet_temperature(lmax=lmax, device=dev)
    log(f'wavelet workload lmax {lmax}: {time.perf_counter() - t0:.1f}s')
    sol = solvers.CGWiener.from_arrays_wav(wp.alm_data, wp.ainfo, wp.icov_ell, wp.icov_wav,
                                           wp.w_ell, mask_pix=wp.mask, minfo_mask=wp.minfo,
                                           spin=0, device=dev)
    prec = preconditioners.PseudoInvPreconditionerWav(wp.ainfo, wp.icov_ell, wp.icov_wav,
                                                      wp.w_ell, 0, mask_pix=wp.mask,
                                                      minfo_mask=wp.minfo)
    b = sol.b0.to(torch.complex128)
    log(f'solver set-up done: {time.perf_counter() - t0:.1f}s')
    for _ in range(args.warmup):
        x, info = mixed.solve_mixed(sol.a_matrix, b, prec, tol=1e-6, delta=args.delta)
        log(f'warmup: iterations {info.iterations}, replacements {info.replacements}, '
            f'residual {info.residual}')
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    for _ in range(args.steps):
        x, info = mixed.solve_mixed(sol.a_matrix, b, prec, tol=1e-6, delta=args.delta)
    torch.cuda.synchronize()
    el = time.perf_counter() - t1
    r = b - sol.a_matrix(x)
    res = float(np.sqrt(alm_utils.contract_almxblm(r, r) / alm_utils.contract_almxblm(b, b)))
    out = {'metric': 'Wiener solves/sec, T wavelet noise, resid 1e-6 (FP64-verified)',
           'value': args.steps / el, 'unit': 'solves/s', 'n_gpus': 1, 'steps': args.steps,
           'warmup': args.warmup, 'ms_per_step': 1e3 * el / args.steps,
           'higher_is_better': True, 'scaling': 'weak', 'vs_baseline': None,
           'dtype': 'f32+f64', 'data': 'synthetic',
           'config': {'workload': f'T lmax {lmax}, GL {lmax + 1}x{2 * lmax + 1}, '
                                  f'{wp.w_ell.shape[0]} needlet scales (lambda 2), galactic '
                                  'mask + 200 holes, wavelet pinv',
                      'iterations': info.iterations, 'replacements': info.replacements,
                      'true_residual': res}}
    print(json.dumps(out))


def streaming_bench(args, dev):
    """HBM-bound kernels of the CG iteration at the headline shapes (TQU,
    lmax 1023, GL 1024x2047, --batch right-hand sides in float32; one right-hand
    side in float64), each launched alone and timed with HIP events on its
    stream. Algorithmic bytes per launch: every array read and written once.
    Prints one JSON line per kernel and one summary line."""
    from optweight_amd import alm_c_utils, mixed
    lmax, nb = args.lmax, args.batch
    ainfo = alm_utils.AlmInfo(lmax)
    nelem = ainfo.nelem
    npix = (lmax + 1) * (2 * lmax + 1)
    rng = np.random.default_rng(0)
    lines = []

    def record(name, fn, nbytes, reps=20):
        ms = time_kernel(fn, reps=reps)
        gbs = nbytes / (ms * 1e-3) / 1e9
        lines.append({'kernel': name, 'ms': ms, 'bytes': nbytes, 'GB/s': gbs,
                      'frac_hbm': gbs / PEAK_HBM_GBS})
        log(f'{name}: {ms * 1e3:.1f} us, {nbytes / 1e6:.1f} MB, {gbs:.0f} GB/s')

    for cdt, rdt, n in ((torch.complex64, torch.float32, nb), (torch.complex128, torch.float64, 1)):
        esz = 8 if cdt == torch.complex64 else 16
        rsz = esz // 2
        tag = 'f32 x%d' % n if rdt == torch.float32 else 'f64 x1'
        alm = torch.randn((n, 3, nelem), dtype=cdt, device=dev)
        out = torch.empty_like(alm)
        lmat = np.abs(rng.normal(size=(3, 3, lmax + 1)))
        lmat = torch.as_tensor(lmat + lmat.transpose(1, 0, 2), dtype=rdt, device=dev)
        alm_c_utils.lmul(alm, lmat, ainfo, alm_out=out)
        record(f'lmul 3x3 ({tag})',
               lambda: _lib.call('owmi_lmul', _lib.dtype_code(alm), _lib.ptr(lmat), 0,
                                 _lib.ptr(alm), _lib.ptr(out), lmax, 3, n,
                                 _lib.stream_ptr(dev)),
               2 * alm.numel() * esz)
        record(f'lmul_add 3x3 + 3x3 ({tag})',
               lambda: _lib.call('owmi_lmul_add', _lib.dtype_code(alm), _lib.ptr(lmat),
                                 _lib.ptr(out), _lib.ptr(lmat), _lib.ptr(alm), _lib.ptr(out),
                                 lmax, 3, n, _lib.stream_ptr(dev)),
               3 * alm.numel() * esz)
        ldiag = torch.as_tensor(np.abs(rng.normal(size=(3, lmax + 1))), dtype=rdt, device=dev)
        record(f'lmul diagonal ({tag})',
               lambda: _lib.call('owmi_lmul', _lib.dtype_code(alm), _lib.ptr(ldiag), 1,
                                 _lib.ptr(alm), _lib.ptr(out), lmax, 3, n,
                                 _lib.stream_ptr(dev)),
               2 * alm.numel() * esz)
        imap = torch.randn((n, 3, npix), dtype=rdt, device=dev)
        d = torch.rand((3, npix), dtype=rdt, device=dev)
        omap = torch.empty_like(imap)
        tmap = torch.randn_like(imap)
        bmap = torch.randn_like(imap)
        dcomp = torch.rand((3,), dtype=rdt, device=dev)
        # x (in place), a, b read, x written: 4 arrays of the map size
        record(f'map_fused x += d (a - b) ({tag})',
               lambda: _lib.call('owmi_map_fused', _lib.dtype_code(imap), _lib.ptr(tmap),
                                 _lib.ptr(tmap), _lib.ptr(imap), _lib.ptr(bmap), _lib.ptr(dcomp),
                                 None, 3, npix, n, _lib.stream_ptr(dev)),
               4 * imap.numel() * imap.element_size())
        del bmap
        record(f'pixel diagonal N^-1 ({tag})',
               lambda: _lib.call('owmi_pix_matvec', _lib.dtype_code(imap), _lib.ptr(d), 0,
                                 _lib.ptr(imap), _lib.ptr(omap), 1, d.numel(), n,
                                 _lib.stream_ptr(dev)),
               (2 * imap.numel() + d.numel()) * rsz)
        b = torch.randn_like(alm)
        dots = alm_utils.contract_almxblm_batch(alm, b)
        record(f'dot <a, b> per RHS ({tag})',
               lambda: _lib.call('owmi_alm_dot_batch', _lib.dtype_code(alm), _lib.ptr(alm),
                                 _lib.ptr(b), lmax, 3, n, _lib.ptr(dots),
                                 _lib.stream_ptr(dev)),
               2 * alm.numel() * esz)
    # the mixed CG updates: x (complex128) += alpha p, r -= alpha q (complex64)
    x = torch.zeros((nb, 3, nelem), dtype=torch.complex128, device=dev)
    r = torch.randn((nb, 3, nelem), dtype=torch.complex64, device=dev)
    pv = torch.randn_like(r)
    q = torch.randn_like(r)
    rz = torch.rand((nb,), dtype=torch.float64, device=dev)
    pq = torch.rand((nb,), dtype=torch.float64, device=dev) + 1
    record(f'CG x, r update (x f64, r/p/q f32 x{nb})',
           lambda: mixed._xr(x, r, pv, q, rz, pq, nb),
           2 * x.numel() * 16 + (2 * r.numel() + pv.numel() + q.numel()) * 8)
    rzn = torch.rand((nb,), dtype=torch.float64, device=dev)
    record(f'CG p update (f32 x{nb})', lambda: mixed._pupd(pv, q, rzn, rz, nb),
           3 * pv.numel() * 8)
    record(f'convert f64 -> f32 (x{nb})', lambda: mixed._convert(x, torch.complex64),
           x.numel() * (16 + 8))
    for ln in lines:
        print(json.dumps({'streaming': ln}))


def phase1_bench(args, dev):
    """Phase 1 of the reference's two-phase filter (filters.py:138-143,
    221-229): CG with the pseudo-inverse plus MaskedPreconditionerCG (15
    inner steps, no low-pass, band limit --prec-cg-lmax, mask of T) on the
    TQU workload. Times the reference's schedule of 5 float64 iterations of
    one right-hand side, one masked-CG application in float64 (one
    right-hand side) and in float32 (a batch), and reports the true
    relative residual after the 5 iterations with and without the masked CG
    preconditioner. Prints one JSON line."""
    prob = workloads.polarized(args.lmax, seed=0, device=dev)
    lmax_cg = args.prec_cg_lmax or args.lmax
    mask = np.asarray(prob.mask_bool)[0].astype(bool)
    t0 = time.perf_counter()
    mcg = preconditioners.MaskedPreconditionerCG(prob.ainfo, prob.icov_ell, prob.spin, mask,
                                                 prob.minfo, lmax=lmax_cg, nsteps=15,
                                                 lmax_r_ell=None)
    pinv = preconditioners.PseudoInvPreconditioner(prob.ainfo, prob.icov_ell, prob.icov_pix,
                                                   prob.minfo, prob.spin)
    log(f'masked CG set-up {time.perf_counter() - t0:.1f}s (lmax {lmax_cg}, '
        f'{int(mcg.mask_unobs[0].sum())} unobserved pixels of {mcg.minfo.npix})')
    niter = 5
    res = {}
    times = {}
    # first applications build the plans of the equal-area geometry
    mcg(solvers.CGWienerMap.from_arrays(prob.imap, prob.minfo, prob.ainfo, prob.icov_ell,
                                        prob.icov_pix, device=dev).b0)
    for name, precs in (('pinv+masked_cg', (pinv, mcg)), ('pinv', (pinv,))):
        solver = solvers.CGWienerMap.from_arrays(prob.imap, prob.minfo, prob.ainfo,
                                                 prob.icov_ell, prob.icov_pix, device=dev)
        for pr in precs:
            solver.add_preconditioner(pr)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        solver.init_solver()
        for _ in range(niter):
            solver.step()
        torch.cuda.synchronize()
        times[name] = time.perf_counter() - t1
        res[name] = float(true_residual(solver, solver.x))
        log(f'{name}: {niter} float64 iterations (with the initial preconditioner application) '
            f'{times[name]:.3f}s, true residual {res[name]:.3e}')
    b64 = solver.b0
    mcg(b64)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    mcg(b64)
    torch.cuda.synchronize()
    t_app64 = time.perf_counter() - t1
    data = workloads.batch_data(prob, args.batch, seed=0, device=dev)
    bsolver = solvers.CGWienerMap.from_arrays(data, prob.minfo, prob.ainfo, prob.icov_ell,
                                              prob.icov_pix, device=dev)
    b32 = bsolver.b0.to(torch.complex64)
    mcg(b32)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    mcg(b32)
    torch.cuda.synchronize()
    t_app32 = time.perf_counter() - t1
    del bsolver, data, b32
    log(f'masked CG application: float64 x1 {t_app64 * 1e3:.1f} ms, float32 x{args.batch} '
        f'{t_app32 * 1e3:.1f} ms ({t_app32 * 1e3 / args.batch:.1f} ms per right-hand side)')
    out = {'phase1': {
        'workload': f'TQU lmax {args.lmax}, GL {args.lmax + 1}x{2 * args.lmax + 1}, pinv + '
                    f'MaskedPreconditionerCG (lmax {lmax_cg}, 15 inner steps)',
        'iterations': niter,
        'fp64_seconds_per_iteration': times['pinv+masked_cg'] / niter,
        'fp64_pinv_only_seconds_per_iteration': times['pinv'] / niter,
        'true_residual_after': res['pinv+masked_cg'],
        'true_residual_after_pinv_only': res['pinv'],
        'masked_cg_application_ms_fp64_x1': t_app64 * 1e3,
        f'masked_cg_application_ms_fp32_x{args.batch}': t_app32 * 1e3}}
    print(json.dumps(out))


def filter_bench(args, dev):
    """The two-phase CAR filter of the reference (filters.CGPixFilter,
    filters.py:9-301) on a full-sky Clenshaw-Curtis CAR map of
    (lmax + 2) x 2 (lmax + 1) pixels, TQU at band limit lmax (config-2
    spectra, depth and mask models, 10 arcmin beam): set-up, then the
    default schedule of 5 phase-1 (pinv + masked CG) and 15 phase-2 (pinv +
    multigrid on T) float64 iterations of one map. Prints one JSON line."""
    from optweight_amd import filters
    ny, nx = args.lmax + 2, 2 * (args.lmax + 1)
    t0 = time.perf_counter()
    wl = workloads.car_polarized(ny, nx, args.lmax, seed=0, device=dev)
    t1 = time.perf_counter()
    f = filters.CGPixFilter(wl['theory_cls'], wl['b_ell'], wl['icov_pix'], wl['mask'],
                            wcs=wl['wcs'], lmax=args.lmax, device=dev)
    f.filter(wl['imap'], niter=1, niter_masked_cg=1, verbose=False)
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    log(f'CAR workload {ny}x{nx} built in {t1 - t0:.1f}s, filter set-up and warm-up '
        f'{t2 - t1:.1f}s')
    out = f.filter(wl['imap'], niter=15, niter_masked_cg=5, verbose=False)
    torch.cuda.synchronize()
    el = time.perf_counter() - t2
    errs = [float(e) for e in out['errors'][1:]]
    log(f'filter: 5 + 15 iterations in {el:.3f}s, errors after phase 1 {errs[4]:.3e}, '
        f'final {errs[-1]:.3e}')
    print(json.dumps({'filter': {
        'workload': f'CGPixFilter, TQU, CC CAR {ny}x{nx}, lmax {args.lmax}, 10 arcmin beam, '
                    'pinv + masked CG (5 steps) then pinv + multigrid T (15 steps), float64',
        'seconds': el, 'iterations': len(errs), 'errors': errs}}))


def const_cor_bench(args, dev):
    """The constant-correlation noise model N = M N_pix^1/2 Y C_ell Y^T
    N_pix^1/2 M (solvers.py:459-575) on the TQU lmax workload: icov_pix and
    mask of the headline workload, a 1/f-like noise spectrum C_ell (knee at
    l = 200, slope -2, TE-correlated), N^-1 by 5 inner CG steps, the
    harmonic preconditioner with the isotropic inverse variance of icov_pix.
    Times 10 float64 iterations of one right-hand side and one float32
    application of A to a batch; prints one JSON line."""
    prob = workloads.polarized(args.lmax, seed=0, device=dev)
    ell = np.arange(args.lmax + 1)
    nl = 1.0 + (np.maximum(ell, 1) / 200.0) ** -2.0
    cov = np.zeros((3, 3, args.lmax + 1))
    cov[0, 0] = nl
    cov[1, 1] = cov[2, 2] = nl
    cov[0, 1] = cov[1, 0] = 0.2 * nl
    icov_noise_ell = np.linalg.inv(cov.transpose(2, 0, 1)).transpose(1, 2, 0)
    mask = prob.mask_bool.astype(np.float64)
    solver = solvers.CGWienerMap.from_arrays_const_cor(
        prob.imap, prob.minfo, prob.ainfo, prob.icov_ell, prob.icov_pix, icov_noise_ell,
        nsteps=5, mask_pix=mask, device=dev)
    solver.add_preconditioner(preconditioners.HarmonicPreconditioner(
        prob.ainfo, prob.icov_ell, icov_pix=prob.icov_pix, minfo=prob.minfo))
    solver.init_solver()
    solver.step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    niter = 10
    for _ in range(niter):
        solver.step()
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    res = float(true_residual(solver, solver.x))
    log(f'const-cor: {niter} float64 iterations {el:.3f}s ({el / niter * 1e3:.1f} ms each), '
        f'true residual after 11: {res:.3e}')
    data = workloads.batch_data(prob, args.batch, seed=0, device=dev)
    bsolver = solvers.CGWienerMap.from_arrays_const_cor(
        data.astype(np.float32), prob.minfo, prob.ainfo, prob.icov_ell, prob.icov_pix,
        icov_noise_ell, nsteps=5, mask_pix=mask, device=dev)
    v = bsolver.b0
    bsolver.a_matrix(v)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    bsolver.a_matrix(v)
    torch.cuda.synchronize()
    t_a32 = time.perf_counter() - t1
    log(f'const-cor A in float32 for {args.batch} right-hand sides: {t_a32 * 1e3:.1f} ms '
        f'({t_a32 * 1e3 / args.batch:.2f} ms each)')
    print(json.dumps({'const_cor': {
        'workload': f'TQU lmax {args.lmax}, GL {args.lmax + 1}x{2 * args.lmax + 1}, '
                    'constant-correlation noise (5 inner CG steps), harmonic preconditioner',
        'fp64_ms_per_iteration': el / niter * 1e3, 'true_residual_after_11': res,
        f'a_matrix_ms_fp32_x{args.batch}': t_a32 * 1e3}}))


def noise_est_bench(args, dev):
    """Wavelet noise-model estimation (noise_utils.estimate_cov_wav) at the
    config-4 shape: temperature noise alm at lmax --wavelet-lmax, 12
    needlet kernels (lambda 2), each scale's map on its Gauss-Legendre grid,
    squared, smoothed and normalised on the GPU (float64). Prints one JSON
    line."""
    from optweight_amd import noise_utils
    lmax = args.wavelet_lmax
    w_ell = workloads.needlet_kernels(lmax)
    ainfo = alm_utils.AlmInfo(lmax)
    rng = np.random.default_rng(0)
    a = (rng.normal(size=(1, 1, ainfo.nelem)) + 1j * rng.normal(size=(1, 1, ainfo.nelem)))
    a[..., :lmax + 1] = a[..., :lmax + 1].real
    at = torch.as_tensor(a, device=dev)
    noise_utils.estimate_cov_wav(at, ainfo, w_ell, 0, diag=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    cov = noise_utils.estimate_cov_wav(at, ainfo, w_ell, 0, diag=True)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    npix = sum(int(cov.minfos[k].npix) for k in cov.maps)
    log(f'estimate_cov_wav: {w_ell.shape[0]} scales, lmax {lmax}, {npix} pixels in total: '
        f'{el:.3f}s')
    print(json.dumps({'noise_estimation': {
        'workload': f'estimate_cov_wav, T, lmax {lmax}, {w_ell.shape[0]} needlet scales, float64',
        'seconds': el, 'pixels': npix}}))


def fwav_bench(args, dev):
    """The 2-D Fourier-wavelet noise model (solvers.py:577-695) at the
    full-sky CAR shape of band limit --wavelet-lmax: temperature map on the
    Clenshaw-Curtis grid of L + 1 rings by 2 L columns (float32), digital
    scale-discrete kernels (lambda 2) on its Fourier plane, a lognormal
    per-pixel covariance per scale. Times N^-1 m (FInvFWavFMatVecMap, 3
    inner CG steps: 7 Fourier-wavelet round trips) and its kernels with HIP
    events (rfft, irfft, slice gather and scatter, inner product, Fourier
    diagonal), and the oracle's float64 numpy version of the same operator
    on a bounded sample (L / 2). Prints one JSON line."""
    from optweight_amd import dft, fkernel, wavtrans, operators, map_utils
    L = args.wavelet_lmax + 1
    ny, nx = L + 1, 2 * L
    wcs = map_utils.CarWCS(cdelt=[-360. / nx, 180. / L], crpix=[nx // 2 + 1, ny // 2 + 1])
    t0 = time.perf_counter()
    ly, lx = dft.laxes_real((ny, nx), wcs)
    fks = fkernel.get_sd_kernels_fourier(ly, lx, 2, digital=True)
    rng = np.random.default_rng(0)
    cov = wavtrans.Wav(2, dtype=np.float32)
    for j, fk in fks:
        ny_j, nx_j = fk.fkernel.shape
        minfo_j = map_utils.get_minfo('CC', ny_j, 2 * (nx_j - 1) + 1)
        cov.add((j, j), np.exp(0.3 * rng.normal(size=(1, minfo_j.npix))).astype(np.float32),
                minfo_j)
    minfo = map_utils.get_minfo('CC', ny, nx)
    op = operators.FInvFWavFMatVecMap(minfo, cov, fks, None, power_x=0, nsteps=3)
    m = torch.as_tensor(rng.normal(size=(1, minfo.npix)), dtype=torch.float32, device=dev)
    op(m)
    torch.cuda.synchronize()
    log(f'fwav setup {time.perf_counter() - t0:.1f}s, {len(fks)} scales, '
        f'kernel blocks {[s[2:] for s in fks.device_plan(dev, torch.float32)["shapes"]]}')
    reps = max(args.steps, 1)
    t_apply = time_kernel(lambda: op(m), reps=reps)
    # kernels at this shape
    m2 = m.reshape(1, ny, nx)
    f = dft.allocate_fmap(tuple(m2.shape), torch.float32, device=dev)
    out = torch.empty_like(m2)
    t_rfft = time_kernel(lambda: dft.rfft(m2, f))
    t_irfft = time_kernel(lambda: dft.irfft(f, out))
    plan = fks.device_plan(dev, torch.float32)
    nlx = nx // 2 + 1
    small = torch.empty(plan['total'], dtype=torch.complex64, device=dev)
    st = _lib.stream_ptr(dev)
    nsc = len(plan['shapes'])
    t_gather = time_kernel(lambda: _lib.call(
        'owmi_fwav_gather', _lib.F32, _lib.ptr(f), 1, ny, nlx, nsc, _lib.ptr(plan['desc']),
        _lib.ptr(plan['rows']), _lib.ptr(plan['kern']), plan['nblocks'], _lib.ptr(small), st))
    t_scatter = time_kernel(lambda: _lib.call(
        'owmi_fwav_scatter', _lib.F32, _lib.ptr(small), 1, ny, nlx, nsc, _lib.ptr(plan['desc']),
        _lib.ptr(plan['invrows']), _lib.ptr(plan['kern']), _lib.ptr(f), st))
    dot_out = torch.empty((), dtype=torch.float64, device=dev)
    t_dot = time_kernel(lambda: dft.contract_fxg_dev(f, f, out=dot_out))
    fdiag = torch.ones((ny, nlx), dtype=torch.float32, device=dev)
    t_fmul = time_kernel(lambda: dft.fmul_2d(f, fdiag, out=f))
    nf = ny * nlx
    ns = plan['total']
    kern_bytes = {
        'fwav_gather': (ns * (8 + 4 + 8) + ns * 4, t_gather),
        'fwav_scatter': (ns * (8 + 4) + nf * 8, t_scatter),
        'fourier_dot': (2 * nf * 8, t_dot),
        'fmul_diag': (nf * (8 + 4 + 8), t_fmul),
    }
    kernels = {}
    for name, (nbytes, ms) in kern_bytes.items():
        gbs = nbytes / (ms * 1e-3) / 1e9
        kernels[name] = {'ms': ms, 'algorithmic_bytes': nbytes, 'GB/s': gbs,
                         'frac_hbm': gbs / PEAK_HBM_GBS}
    log(f'fwav: N^-1 apply {t_apply:.2f} ms; rfft {t_rfft:.3f} ms, irfft {t_irfft:.3f} ms; '
        + ', '.join(f'{k} {v["ms"]:.3f} ms ({v["GB/s"]:.0f} GB/s)' for k, v in kernels.items()))
    cpu = None
    if args.cpu_baseline == 'on':
        from oracle import fourier as ofour
        Ls = L // 2
        nys, nxs = Ls + 1, 2 * Ls
        wcs_s = map_utils.CarWCS(cdelt=[-360. / nxs, 180. / Ls], crpix=[nxs // 2 + 1, nys // 2 + 1])
        lys, lxs = dft.laxes_real((nys, nxs), wcs_s)
        fks_s = fkernel.get_sd_kernels_fourier(lys, lxs, 2, digital=True)
        kern_s = [(j, np.asarray(fk.fkernel, dtype=np.float64), fk.slices_y, fk.slice_x)
                  for j, fk in fks_s]
        cov_s = {j: np.exp(0.3 * rng.normal(size=(1, k.shape[0] * (2 * k.shape[1] - 1))))
                 for j, k, _, _ in kern_s}
        oop = ofour.finv_fwav_map_op(nys, nxs, kern_s, cov_s, None, power_x=0, nsteps=3)
        ms_ = rng.normal(size=(1, nys * nxs))
        t1 = time.perf_counter()
        oop(ms_)
        el = time.perf_counter() - t1
        cpu = {'value': 1.0 / el, 'unit': 'N^-1 applications/s', 'cores': 1, 'kind': 'port',
               'sample': f'oracle FInvFWavFMatVecMap (numpy, float64), T, CC {nys}x{nxs}, '
                         f'{len(kern_s)} scales, one application ({el:.1f} s)'}
        log(f'cpu baseline: {cpu["sample"]}')
    print(json.dumps({'fwav': {
        'workload': f'FInvFWavFMatVecMap (3 inner CG steps), T, CC {ny}x{nx} float32, '
                    f'{len(fks)} digital scale-discrete scales (lambda 2)',
        'ms_per_apply': t_apply, 'applications_per_s': 1e3 / t_apply,
        'rfft_ms': t_rfft, 'irfft_ms': t_irfft, 'kernels': kernels, 'cpu_baseline': cpu}}))


def ragged_bench(args, dev):
    '''Constrained realisations of uneven band limit: `batch` right-hand sides
    at each band limit (data + S^-1/2 w + P^T N^-1/2 w', seeds [k, i]), pinv +
    multigrid(T) preconditioning, mixed precision, FP64-verified residual.
    Timed: the lockstep solve (all band limits in every CG phase, one stream
    each, phases replayed from HIP graphs) and, for comparison, one batched
    solve per band limit in turn; both give the same iterates.'''
    lmaxs = [int(v) for v in args.ragged_lmaxs.split(',')]
    t0 = time.perf_counter()
    work, solvers_, tuned = [], [], {}
    for k, (prob, data) in enumerate(workloads.ragged(lmaxs, per=args.batch, seed=0,
                                                      device=dev)):
        sol = solvers.CGWienerMap.from_arrays(data, prob.minfo, prob.ainfo, prob.icov_ell,
                                              prob.icov_pix, draw_constr=True, seed=100 + k,
                                              device=dev)
        sol.add_preconditioner(preconditioners.PseudoInvPreconditioner(
            prob.ainfo, prob.icov_ell, prob.icov_pix, prob.minfo, prob.spin))
        if not args.no_mg:
            sol.add_preconditioner(preconditioners.MaskedPreconditioner(
                prob.ainfo, prob.icov_ell[0:1, 0:1], 0, prob.mask_bool[0], prob.minfo,
                min_pix=1000, n_jacobi=1, lmax_r_ell=6000, device=dev), sel=np.s_[0])
        if args.autotune == 'on':
            res = autotune.tune(prob.minfo, prob.lmax, torch.float32, args.batch, device=dev)
            tuned[prob.lmax] = res['params']
        solvers_.append(sol)
        work.append((sol.a_matrix, sol.b0, sol.preconditioner))
        log(f'bucket lmax {prob.lmax}: setup done ({time.perf_counter() - t0:.1f}s)')

    def timed(lockstep):
        # the lockstep solver records its HIP graphs in the first (warm-up)
        # solve and replays them in the timed ones
        lk = mixed.Lockstep(work, delta=args.delta) if lockstep else None

        def run():
            if lk is not None:
                return lk.solve()
            return mixed.solve_ragged(work, delta=args.delta, lockstep=False, graphs=False)
        for _ in range(max(1, args.warmup)):
            run()
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        for _ in range(args.steps):
            res = run()
        torch.cuda.synchronize()
        el = time.perf_counter() - t1
        log(f'{"lockstep" if lockstep else "per band limit"}: {1e3 * el / args.steps:.1f} ms '
            f'per step')
        return el, res

    el_seq, res_seq = timed(False)
    el, res = timed(True)
    worst, same = 0.0, 0.0
    for sol, (x, info), (xs, _) in zip(solvers_, res, res_seq):
        if not info.graphs:
            raise RuntimeError('the lockstep solve did not replay HIP graphs')
        worst = max(worst, float(np.max(true_residual(sol, x))))
        same = max(same, float(torch.linalg.vector_norm(x - xs) / torch.linalg.vector_norm(xs)))
        log(f'lmax {sol.sht.ainfo.lmax}: iterations {info.iterations}, '
            f'replacements {info.replacements}')
    n = args.steps * args.batch * len(lmaxs)
    out = {'metric': 'constrained realisations/sec, TQU, uneven lmax, resid 1e-6',
           'value': n / el, 'unit': 'solves/s', 'n_gpus': 1, 'steps': args.steps,
           'warmup': args.warmup, 'ms_per_step': 1e3 * el / args.steps,
           'higher_is_better': True, 'scaling': 'weak', 'vs_baseline': None,
           'dtype': 'f32+f64', 'data': 'synthetic',
           'config': {'workload': f'{args.batch} right-hand sides at each lmax {lmaxs}, '
                                  'pinv + multigrid(T), mixed precision, lockstep CG over '
                                  'the band limits (one stream each, phases replayed from '
                                  'HIP graphs recorded in the warm-up solve)',
                      'true_residual_max': worst,
                      'per_band_limit_solves_per_s': n / el_seq,
                      'speedup_lockstep_vs_per_band_limit': el_seq / el,
                      'rel_l2_lockstep_vs_per_band_limit_max': same,
                      'iterations': [info.iterations for _, info in res],
                      'autotune': tuned}}
    print(json.dumps(out))


def autodiff_bench(args, dev):
    '''Gradients of L = 1/2 d^T (P S P^T + N)^-1 d + 1/2 <x, x> (x the Wiener
    solution) with respect to the data map d and 30 log-spaced band
    amplitudes of S, on the TQU workload at --lmax: one step is the forward
    float64 solve, the loss and the backward (one adjoint float64 solve).
    Central finite differences of the loss along one random direction in d
    and one in the amplitudes check the gradients at this size.'''
    from optweight_amd import autodiff
    t0 = time.perf_counter()
    prob = workloads.polarized(args.lmax, seed=0, device=dev)
    edges = workloads.log_bands(2, args.lmax, 30)
    model = autodiff.WienerModel(prob.minfo, prob.ainfo, prob.icov_ell, prob.icov_pix, edges,
                                 tol=1e-10, device=dev)
    amps = 1.0 + 0.1 * np.sin(np.arange(30))
    d0 = torch.as_tensor(prob.imap, device=dev)
    a0 = torch.as_tensor(amps, device=dev)
    log(f'setup {time.perf_counter() - t0:.1f}s')

    def step():
        d = d0.clone().requires_grad_()
        a = a0.clone().requires_grad_()
        loss = autodiff.config_loss(d, a, model)
        loss.backward()
        return float(loss.item()), d.grad, a.grad

    for _ in range(max(1, args.warmup)):
        loss, gd, ga = step()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    for _ in range(args.steps):
        loss, gd, ga = step()
    torch.cuda.synchronize()
    el = time.perf_counter() - t1

    def value(d, a):
        with torch.no_grad():
            return float(autodiff.config_loss(d, a, model).item())
    rng = np.random.default_rng(1)
    ed = torch.as_tensor(rng.normal(size=prob.imap.shape) * (prob.icov_pix > 0), device=dev)
    ed *= torch.linalg.vector_norm(d0) / torch.linalg.vector_norm(ed)
    ea = torch.as_tensor(rng.normal(size=30), device=dev)
    # L is quadratic in d (central differences exact up to the solves);
    # in the amplitudes the step trades truncation against solve error
    hd, ha = 1e-3, 1e-4
    fd_d = (value(d0 + hd * ed, a0) - value(d0 - hd * ed, a0)) / (2 * hd)
    fd_a = (value(d0, a0 + ha * ea) - value(d0, a0 - ha * ea)) / (2 * ha)
    an_d = float(torch.sum(gd * ed).item())
    an_a = float(torch.sum(ga * ea).item())
    err_d = abs(an_d - fd_d) / abs(fd_d)
    err_a = abs(an_a - fd_a) / abs(fd_a)
    log(f'fd check: d {an_d:.10e} vs {fd_d:.10e} ({err_d:.1e}); '
        f'amplitudes {an_a:.10e} vs {fd_a:.10e} ({err_a:.1e})')
    out = {'metric': 'loss+gradient evaluations/sec (data map and 30 band amplitudes), TQU '
                     f'lmax={args.lmax}',
           'value': args.steps / el, 'unit': 'evaluations/s', 'n_gpus': 1, 'steps': args.steps,
           'warmup': args.warmup, 'ms_per_step': 1e3 * el / args.steps,
           'higher_is_better': True, 'scaling': 'weak', 'vs_baseline': None, 'dtype': 'f64',
           'data': 'synthetic',
           'config': {'workload': f'differentiable Wiener filter, TQU lmax {args.lmax}, GL '
                                  f'{prob.minfo.nrow}x{prob.minfo.nphi[0]}, pinv preconditioner, '
                                  'forward and adjoint float64 CG to 1e-10, 30 log-spaced bands '
                                  f'{edges.tolist()}',
                      'loss': loss,
                      'fd_rel_err_data': err_d, 'fd_rel_err_amplitudes': err_a,
                      'fd_steps': {'data': hd, 'amplitudes': ha}}}
    print(json.dumps(out))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=5)
    ap.add_argument('--warmup', type=int, default=1)
    ap.add_argument('--lmax', type=int, default=1023)
    ap.add_argument('--no-mg', action='store_true')
    ap.add_argument('--cpu-baseline', choices=['on', 'off'], default='on')
    ap.add_argument('--batch', type=int, default=None,
                    help='right-hand sides per mixed-precision batched solve (default 128 '
                         'for the headline solve, 16 per band limit for --ragged and for '
                         'the batched timings of the other modes)')
    ap.add_argument('--delta', type=float, default=0.001,
                    help='residual reduction between float64 residual replacements')
    ap.add_argument('--fp64-batch-steps', type=int, default=1,
                    help='timed plain float64 solves of the whole batch (after one untimed)')
    ap.add_argument('--autotune', choices=['on', 'off'], default='on',
                    help='tune the Legendre launch choices of the workload plan (float32 at '
                         'the batch, float64 single transforms) before timing; cached')
    ap.add_argument('--ragged', action='store_true',
                    help='time the batched constrained realisations of uneven band limit '
                         '(16 right-hand sides at each of lmax 255, 511, 1023, 2047) instead '
                         'of the headline workload; prints its own JSON line')
    ap.add_argument('--ragged-lmaxs', default='255,511,1023,2047')
    ap.add_argument('--wavelet', action='store_true',
                    help='time the temperature solve with the wavelet noise model (GL grid of '
                         '--wavelet-lmax, 12 needlet scales, float32 maps, mixed precision, '
                         'wavelet pseudo-inverse preconditioner); prints its own JSON line')
    ap.add_argument('--wavelet-lmax', type=int, default=4095)
    ap.add_argument('--streaming', action='store_true',
                    help='time the HBM-bound kernels of the CG iteration at the headline shapes '
                         '(lmul, pixel diagonal, dots, CG updates, conversion); prints one '
                         'JSON line per kernel')
    ap.add_argument('--filter', action='store_true',
                    help='time the two-phase CAR filter (CGPixFilter) on a full-sky CC map at '
                         '--lmax; prints its own JSON line')
    ap.add_argument('--noise-est', action='store_true',
                    help='time the wavelet noise-model estimation at --wavelet-lmax; prints its '
                         'own JSON line')
    ap.add_argument('--const-cor', action='store_true',
                    help='time the constant-correlation noise model solve on the TQU workload; '
                         'prints its own JSON line')
    ap.add_argument('--phase1', action='store_true',
                    help='time phase 1 of the two-phase filter (pseudo-inverse plus the masked '
                         'CG preconditioner) on the TQU workload; prints its own JSON line')
    ap.add_argument('--prec-cg-lmax', type=int, default=0,
                    help='band limit of the masked CG preconditioner (default: --lmax)')
    ap.add_argument('--fwav', action='store_true',
                    help='time the 2-D Fourier-wavelet noise model inverse on the full-sky CAR '
                         'map of band limit --wavelet-lmax; prints its own JSON line')
    ap.add_argument('--autodiff', action='store_true',
                    help='time loss + gradient evaluations of the differentiable Wiener filter '
                         '(data map and 30 log-spaced band amplitudes) on the TQU workload at '
                         '--lmax, with a finite-difference check; prints its own JSON line')
    ap.add_argument('--roofline-only', action='store_true',
                    help='only build the workload geometry and time the roofline kernel '
                         '(for counter runs), after one copy of known size')
    args = ap.parse_args()
    headline = not (args.ragged or args.wavelet or args.phase1 or args.filter or args.const_cor
                    or args.noise_est or args.fwav or args.streaming or args.autodiff)
    if args.batch is None:
        args.batch = HEADLINE_BATCH if headline else AUX_BATCH
    if args.gpus != 1:
        print('bench.py measures one GPU: this build has no multi-GPU path '
              '(independent right-hand sides shard trivially across GPUs).', file=sys.stderr)
        sys.exit(2)
    dev = torch.device('cuda:0')
    torch.cuda.set_device(dev)

    if args.ragged:
        ragged_bench(args, dev)
        return
    if args.autodiff:
        autodiff_bench(args, dev)
        return
    if args.wavelet:
        wavelet_bench(args, dev)
        return

    if args.phase1:
        phase1_bench(args, dev)
        return
    if args.filter:
        filter_bench(args, dev)
        return
    if args.const_cor:
        const_cor_bench(args, dev)
        return
    if args.noise_est:
        noise_est_bench(args, dev)
        return
    if args.fwav:
        fwav_bench(args, dev)
        return
    if args.streaming:
        calibration_copy(dev)
        streaming_bench(args, dev)
        return
    if args.roofline_only:
        calibration_copy(dev)
        prob = workloads.polarized(args.lmax, seed=0, device=dev)
        roof = legendre_roofline(prob, dev, args.batch, reps=3)
        print(json.dumps({'roofline_only': roof}))
        return

    log(f'building workload TQU lmax={args.lmax}, batch {args.batch}')
    t0 = time.perf_counter()
    prob = workloads.polarized(args.lmax, seed=0, device=dev)
    solver = build_solver(prob, dev, use_mg=not args.no_mg)
    data = workloads.batch_data(prob, args.batch, seed=0, device=dev)
    data[0] = prob.imap
    bsolver = build_solver(prob, dev, imap=data, use_mg=not args.no_mg)
    log(f'setup {time.perf_counter() - t0:.1f}s')
    tuned = {}
    if args.autotune == 'on':
        for dt, nb in ((torch.float32, args.batch), (torch.float64, 1),
                       (torch.float64, args.batch)):
            res = autotune.tune(prob.minfo, prob.lmax, dt, nb, device=dev)
            tuned[f'{str(dt)[6:]}_batch{nb}'] = {k: res[k] for k in ('params', 'cached',
                                                                        'tune_seconds')}
        log(f'autotune: {tuned}')

    # plain float64 GPU solve, one right-hand side per step
    x64, el64, inf64 = time_solves('fp64', solver, args.steps, args.warmup)
    res64 = float(true_residual(solver, x64))
    fp64_sps = args.steps / el64
    it64 = float(np.mean([i.iterations for i in inf64]))
    log(f'fp64: {args.steps} solves in {el64:.3f}s ({fp64_sps:.3f}/s), {it64:.1f} it/solve, '
        f'true residual {res64:.3e}')
    # mixed precision, a batch of right-hand sides per step
    xmx, elmx, infmx, flmx = time_solves('mixed', bsolver, args.steps, args.warmup,
                                         delta=args.delta, count_flops=True)
    resmx = true_residual(bsolver, xmx)
    nsolve = args.steps * args.batch
    mixed_sps = nsolve / elmx
    itmx = float(np.mean([i.iterations for i in infmx]))
    nrep = float(np.mean([i.replacements for i in infmx]))
    diff = float(torch.linalg.vector_norm(xmx[0] - x64) / torch.linalg.vector_norm(x64))
    log(f'mixed batch {args.batch}: {nsolve} solves in {elmx:.3f}s ({mixed_sps:.3f}/s), '
        f'{itmx:.1f} it/batch, {nrep:.1f} replacements, max true residual '
        f'{resmx.max():.3e}, RHS 0 rel L2 vs fp64 solution {diff:.3e}, '
        f'speedup {mixed_sps / fp64_sps:.2f}x')
    log(f'mixed: {itmx * args.steps / elmx:.1f} batched iterations/s, useful Legendre work '
        f'{flmx["float32"] / elmx / 1e12:.1f} TFLOP/s float32 + '
        f'{flmx["float64"] / elmx / 1e12:.2f} TFLOP/s float64 over the solve time')
    elapsed = elmx
    roof = legendre_roofline(prob, dev, args.batch)
    log(f'roofline: {roof}')

    # plain float64 CG on the same batch of right-hand sides (batching
    # without mixed precision), and the mixed solutions' error against it
    # per right-hand side
    # (--fp64-batch-steps 0 skips it, e.g. for a profile of the mixed solve)
    fp64b_sps, res64b, errs = None, np.array([np.nan]), np.array([np.nan])
    if args.fp64_batch_steps > 0:
        x64b, el64b, inf64b = time_solves('fp64', bsolver, args.fp64_batch_steps, 1)
        fp64b_sps = args.fp64_batch_steps * args.batch / el64b
        errs = (torch.linalg.vector_norm((xmx - x64b).reshape(args.batch, -1), dim=1)
                / torch.linalg.vector_norm(x64b.reshape(args.batch, -1), dim=1)).cpu().numpy()
        res64b = true_residual(bsolver, x64b)
        log(f'fp64 batch {args.batch}: {fp64b_sps:.3f} solves/s, max true residual '
            f'{res64b.max():.3e}; mixed vs fp64 per RHS rel L2 max {errs.max():.3e} '
            f'median {np.median(errs):.3e}')

    cpu = None
    if args.cpu_baseline == 'on':
        cpu, xo = cpu_baseline_solve(prob, it64)
        # the same 5 float64 CG steps on the GPU: full-size parity of the
        # iterate with the oracle
        solver.init_solver()
        for _ in range(len(cpu['iteration_s'])):
            solver.step()
        cpu['gpu_vs_oracle_rel_l2_5_steps'] = float(
            np.linalg.norm(solver.x.cpu().numpy() - xo) / np.linalg.norm(xo))
        log(f'cpu baseline: {cpu}')

    out = {
        'metric': 'Wiener solves/sec, TQU lmax=1023 to resid 1e-6: mixed/batched vs FP64 GPU',
        'value': mixed_sps,
        'unit': 'solves/s',
        'n_gpus': 1,
        'steps': args.steps,
        'warmup': args.warmup,
        'ms_per_step': 1e3 * elapsed / args.steps,
        'higher_is_better': True,
        'scaling': 'weak',
        'vs_baseline': None,
        'dtype': 'f32+f64',
        'data': 'synthetic',
        'config': {'workload': f'TQU Wiener filter, lmax {args.lmax}, GL {args.lmax + 1}x'
                               f'{2 * args.lmax + 1}, pinv + multigrid(T) preconditioner, '
                               f'true resid 1e-6',
                   'solver': f'mixed (f32 CG, f64 reliable updates, delta {args.delta}), '
                             f'{args.batch} right-hand sides per batch',
                   'batch': args.batch, 'iterations': itmx, 'replacements': nrep,
                   'true_residual_max': float(resmx.max()),
                   'rel_l2_vs_fp64_solution': diff,
                   'fp64_solves_per_s': fp64_sps, 'fp64_iterations': it64,
                   'fp64_true_residual': res64,
                   'speedup_vs_fp64': mixed_sps / fp64_sps,
                   # the speed-up split into batching and precision: plain
                   # float64 CG on the same batch, and per-RHS distance of
                   # the mixed solutions to those float64 solutions
                   'fp64_batched_solves_per_s': fp64b_sps,
                   'fp64_batched_true_residual_max': float(res64b.max()),
                   'speedup_fp64_batched_vs_fp64': (fp64b_sps / fp64_sps
                                                    if fp64b_sps else None),
                   'speedup_mixed_vs_fp64_batched': (mixed_sps / fp64b_sps
                                                     if fp64b_sps else None),
                   'rel_l2_vs_fp64_batched_max': float(errs.max()),
                   'rel_l2_vs_fp64_batched_median': float(np.median(errs)),
                   'autotune': tuned,
                   # batched iterations per second, and the useful Legendre
                   # flops of the timed mixed solves over their wall time
                   'iterations_per_s': itmx * args.steps / elmx,
                   'legendre_tflops_per_s_fp32': flmx['float32'] / elmx / 1e12,
                   'legendre_tflops_per_s_fp64': flmx['float64'] / elmx / 1e12},
        'roofline': {'bound': 'mfma', 'achieved': roof['achieved'],
                     'peak': PEAK_FP32_TFLOPS, 'unit': 'TFLOP/s',
                     'frac': roof['achieved'] / PEAK_FP32_TFLOPS,
                     # counters were read at the default shape only
                     'traffic': (ROOFLINE_TRAFFIC_BYTES
                                 if (args.lmax, args.batch) == ROOFLINE_TRAFFIC_SHAPE
                                 else None),
                     'kernel': roof['kernel'], 'kernel_ms': roof['ms'],
                     'flops_per_launch': roof['flops']},
        'cpu_baseline': cpu,
    }
    print(json.dumps(out))


if __name__ == '__main__':
    main()
