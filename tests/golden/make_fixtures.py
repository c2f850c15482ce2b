'''
Writes the full-size golden fixtures of tests/test_gpu_fullsize.py with the
oracle's long-double numpy transforms (oracle/sht.py, not the float64 C
backend), in the build container:

  config2_tqu1023.npz   TQU lmax 1023 on the GL 1024 x 2047 grid (the
      benchmark workload, workloads.polarized(1023, seed=0)): for seeded
      alm v (numpy default_rng(101)) and map f (default_rng(102)):
      A v (the Wiener operator), the pseudo-inverse P v, Y v (synthesis),
      Y^T W f (weighted adjoint); each as values at 4096 seeded indices, its
      norm, and <u_k, .> for 8 seeded u_k (default_rng(200 + k)).
  config1_t255.npz      T lmax 255 (workloads.temperature(255, seed=0) with
      the oracle synthesising the signal): the data map's b and the float64
      CG solution with the pseudo-inverse, driven to true relative residual
      <= 1e-13 (full vector).

    python tests/golden/make_fixtures.py [--only config1|config2]
'''
import argparse
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import geometry as og, wiener as ow, sht as osht, ops  # noqa: E402
from optweight_amd import workloads  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))
NSAMP = 4096
NPROJ = 8


def geom_of(minfo):
    return og.Geom(minfo.theta, minfo.weight, minfo.nphi, phi0=minfo.phi0)


def oracle_synth(a, minfo, lmax, spin):
    out = np.zeros((a.shape[0], minfo.npix))
    osht.alm2map(a, out, lmax, geom_of(minfo), spin)
    return out


def rand_alm(shape, lmax, seed):
    rng = np.random.default_rng(seed)
    v = rng.normal(size=shape) + 1j * rng.normal(size=shape)
    v[..., :lmax + 1] = v[..., :lmax + 1].real
    return v


def summary(name, arr, out):
    '''Samples at seeded flat indices, the norm, and seeded projections.'''
    flat = arr.reshape(-1)
    idx = np.sort(np.random.default_rng(300).choice(flat.size, NSAMP, replace=False))
    out[f'{name}_idx'] = idx
    out[f'{name}_val'] = flat[idx]
    out[f'{name}_norm'] = np.linalg.norm(flat)
    proj = []
    for k in range(NPROJ):
        rng = np.random.default_rng(200 + k)
        u = rng.normal(size=flat.size)
        if np.iscomplexobj(flat):
            u = u + 1j * rng.normal(size=flat.size)
        proj.append(np.vdot(u, flat))
    out[f'{name}_proj'] = np.asarray(proj)


def config2():
    lmax = 1023
    t0 = time.time()
    prob = workloads.polarized(lmax, seed=0, synth=lambda a, mi, l, s: np.zeros((3, mi.npix)))
    g = geom_of(prob.minfo)
    osol = ow.WienerMap.from_arrays(prob.imap, g, lmax, prob.icov_ell, prob.icov_pix)
    pinv = ow.PseudoInvPreconditioner(lmax, prob.icov_ell, prob.icov_pix, g, prob.spin)
    v = rand_alm((3, prob.ainfo.nelem), lmax, 101)
    f = np.random.default_rng(102).normal(size=(3, g.npix))
    out = {}
    summary('av', osol.a_matrix(v), out)
    print(f'A v {time.time() - t0:.0f}s', flush=True)
    summary('pv', pinv(v.copy()), out)
    print(f'P v {time.time() - t0:.0f}s', flush=True)
    m = np.zeros((3, g.npix))
    osht.alm2map(v, m, lmax, g, [0, 2])
    summary('yv', m, out)
    a = np.zeros_like(v)
    osht.map2alm(f, a, g, lmax, [0, 2])
    summary('ytwf', a, out)
    print(f'SHT pair {time.time() - t0:.0f}s', flush=True)
    np.savez_compressed(os.path.join(HERE, 'config2_tqu1023.npz'), **out)


def config1():
    lmax = 255
    prob = workloads.temperature(lmax, seed=0, synth=oracle_synth)
    g = geom_of(prob.minfo)
    osol = ow.WienerMap.from_arrays(prob.imap, g, lmax, prob.icov_ell, prob.icov_pix)
    osol.add_preconditioner(ow.PseudoInvPreconditioner(lmax, prob.icov_ell, prob.icov_pix, g, 0))
    cg = osol.init_solver()
    bn = np.sqrt(ops.contract_almxblm(osol.b0, osol.b0))
    for it in range(500):
        cg.step()
        if osol.get_residual() / bn < 1e-13:
            break
    res = osol.get_residual() / bn
    print(f'config 1: {cg.i} iterations, true residual {res:.2e}', flush=True)
    np.savez_compressed(os.path.join(HERE, 'config1_t255.npz'), imap=prob.imap, b=osol.b0,
                        x=cg.x, residual=res, iterations=cg.i)


if __name__ == '__main__':
    ap = argparse.ArgumentParser()
    ap.add_argument('--only', choices=['config1', 'config2'])
    args = ap.parse_args()
    if args.only in (None, 'config1'):
        config1()
    if args.only in (None, 'config2'):
        config2()
