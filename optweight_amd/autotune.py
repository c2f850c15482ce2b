'''
Autotuner of the Legendre kernels' launch choices. No counterpart in the
reference (ducc0 fixes its blocking internally); the choices change which
kernel variant runs and the order of float sums, never the transform.

For one plan (ring geometry and lmax), precision and batch size, ``tune``
times the Legendre stage of both directions (owmi_sht_legendre; spin 0 on
one component and spin 2 on two, as a TQU transform runs them) under each
candidate value of the parameters the dispatch consults at that batch, one
parameter at a time from the library's own rules (coordinate descent), and
keeps the fastest value of each:

  float32, batch >= 8:   matrix cores or vector kernels (OWMI_TUNE_MFMA_MIN),
                         then column tiles of the matrix-core adjoint per
                         spin (OWMI_TUNE_ADJ_CT0 / CT2: 2 or 4);
  float64, batch >= 4:   float64 matrix cores or vector kernels for the
                         synthesis and the adjoint (OWMI_TUNE_MFMA64_MIN),
                         then as below if the vector kernels won;
  otherwise:             transforms per thread of the vector synthesis and
                         per block of the vector adjoint (1, 2, 4), ring
                         pairs per lane of the vector adjoint (spin 0: 1, 2,
                         4; spin 2: 1, 2).

The winners are set on the plan for batches of at least ``batch``
(owmi_sht_plan_tune) and cached in a JSON file (OWMI_TUNE_CACHE, default
~/.cache/optweight_amd/tune.json) under the key (lmax, nrings, hash of the
rings' nphi and theta, dtype, batch, device name); sht.get_plan re-applies
every cached entry of its geometry to a new plan.
'''
import hashlib
import json
import os
import time

import numpy as np
import torch

from optweight_amd import _lib

TUNE_MFMA_MIN = 1
TUNE_ADJ_CT0 = 2
TUNE_ADJ_CT2 = 3
TUNE_SYNTH_NB = 4
TUNE_ADJ_NB = 5
TUNE_ADJ_K0 = 6
TUNE_ADJ_K2 = 7
TUNE_MFMA64_MIN = 8

PARAMS = {'mfma_min': TUNE_MFMA_MIN, 'adj_ct0': TUNE_ADJ_CT0, 'adj_ct2': TUNE_ADJ_CT2,
          'synth_nb': TUNE_SYNTH_NB, 'adj_nb': TUNE_ADJ_NB, 'adj_k0': TUNE_ADJ_K0,
          'adj_k2': TUNE_ADJ_K2, 'mfma64_min': TUNE_MFMA64_MIN}

_loaded = None


def cache_path():
    return os.environ.get('OWMI_TUNE_CACHE') or os.path.join(
        os.path.expanduser('~'), '.cache', 'optweight_amd', 'tune.json')


def geometry_key(minfo, lmax, rings=None):
    '''Geometry part of the cache key: lmax, ring count and a hash of the
    rings' nphi and theta.'''
    sel = slice(None) if rings is None else np.asarray(rings, dtype=np.int64)
    nphi = np.ascontiguousarray(np.asarray(minfo.nphi)[sel], dtype=np.int64)
    theta = np.ascontiguousarray(np.asarray(minfo.theta)[sel], dtype=np.float64)
    h = hashlib.sha1(nphi.tobytes() + theta.tobytes()).hexdigest()[:16]
    return f'lmax={int(lmax)};nrings={nphi.size};rings={h}'


def cache_key(minfo, lmax, dtype, batch, device, rings=None):
    dt = 'float64' if dtype in (torch.float64, torch.complex128) else 'float32'
    name = torch.cuda.get_device_name(device).replace(';', ',')
    return f'{geometry_key(minfo, lmax, rings)};dtype={dt};batch={int(batch)};device={name}'


def load_cache(path=None):
    path = path or cache_path()
    try:
        with open(path) as f:
            data = json.load(f)
        return data if isinstance(data, dict) else {}
    except (OSError, ValueError):
        return {}


def save_cache(data, path=None):
    '''Write the cache atomically (a temporary file renamed over the old).'''
    path = path or cache_path()
    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    tmp = f'{path}.{os.getpid()}.tmp'
    with open(tmp, 'w') as f:
        json.dump(data, f, indent=1, sort_keys=True)
    os.replace(tmp, path)


def _set(plan, batch, params):
    for name, value in params.items():
        _lib.call('owmi_sht_plan_tune', plan.handle, int(batch), PARAMS[name], int(value))


def apply_cached(plan, minfo, lmax, rings=None):
    '''Set every cached entry of this geometry (all batches and dtypes; the
    parameters of the two dtypes never overlap in meaning) on a new plan.'''
    global _loaded
    if _loaded is None or _loaded[0] != cache_path():
        _loaded = (cache_path(), load_cache())
    gk = geometry_key(minfo, lmax, rings) + ';'
    dev_name = torch.cuda.get_device_name(plan.device)
    for key, entry in _loaded[1].items():
        if key.startswith(gk) and key.endswith(f';device={dev_name}'):
            _set(plan, entry['batch'], entry['params'])


def _timer(plan, dtype, batch, reps):
    '''Seconds of the Legendre stage of a TQU transform pair (synthesis and
    adjoint, spin 0 then spin 2) at this batch, median of reps.'''
    cdt = torch.complex128 if dtype in (torch.float64, torch.complex128) else torch.complex64
    code = _lib.F64 if cdt == torch.complex128 else _lib.F32
    nelem = (plan.lmax + 1) * (plan.lmax + 2) // 2
    alm = torch.zeros((batch, 3, nelem), dtype=cdt, device=plan.device)
    st = _lib.stream_ptr(plan.device)

    def run():
        for direction in (0, 1):
            _lib.check(plan._lib.owmi_sht_legendre(plan.handle, code, 0, direction,
                                                   _lib.ptr(alm[0, 0]), 3 * nelem, batch, st))
            _lib.check(plan._lib.owmi_sht_legendre(plan.handle, code, 2, direction,
                                                   _lib.ptr(alm[0, 1]), 3 * nelem, batch, st))

    def measure():
        run()
        times = []
        for _ in range(reps):
            s = torch.cuda.Event(enable_timing=True)
            e = torch.cuda.Event(enable_timing=True)
            s.record()
            run()
            e.record()
            e.synchronize()
            times.append(s.elapsed_time(e) * 1e-3)
        return float(np.median(times))
    return measure


def _candidates(dtype, batch, mfma):
    '''The next parameters to search; mfma: None before the matrix-core
    choice, else whether the matrix cores won.'''
    f64 = dtype in (torch.float64, torch.complex128)
    if not f64 and batch >= 8:
        if mfma is None:
            # 1: the matrix cores at this batch; batch + 1: the vector kernels
            return [('mfma_min', [1, batch + 1])]
        if mfma:
            return [('adj_ct2', [2, 4]), ('adj_ct0', [2, 4])]
    if f64 and batch >= 4:
        if mfma is None:
            return [('mfma64_min', [1, batch + 1])]
        if mfma:
            return []  # both directions run on the float64 matrix cores
    nbs = [v for v in (1, 2, 4) if v <= batch]
    return [('synth_nb', nbs), ('adj_nb', nbs), ('adj_k0', [1, 2, 4]), ('adj_k2', [1, 2])]


def _searched(cands):
    '''The candidates with a choice to make. A parameter with one candidate
    value at this batch (transforms per thread at batch 1) is left out: the
    entry also covers larger batches, where the library's own rule applies
    to it.'''
    return [(name, values) for name, values in cands if len(values) > 1]


def tune(minfo, lmax, dtype, batch, device=None, rings=None, reps=3, use_cache=True,
         force=False):
    '''Tune the plan of (minfo, lmax[, rings]) on ``device`` for ``batch``
    right-hand sides of ``dtype`` (float32 / float64 or their complex
    types), set the result on the plan and cache it. Returns
    {'params': {name: value}, 'seconds': timing of the winner, 'cached':
    bool, 'tune_seconds': wall time of the search}.'''
    from optweight_amd import sht
    if device is None:
        device = torch.device('cuda', torch.cuda.current_device())
    plan = sht.get_plan(minfo, lmax, device, rings=rings)
    key = cache_key(minfo, lmax, dtype, batch, plan.device, rings)
    data = load_cache() if use_cache else {}
    if use_cache and not force and key in data:
        entry = data[key]
        _set(plan, batch, entry['params'])
        return dict(entry, cached=True, tune_seconds=0.0)
    t0 = time.perf_counter()
    measure = _timer(plan, dtype, batch, reps)
    best = {}
    _set(plan, batch, {name: 0 for name in PARAMS})
    mfma = None
    best_t = measure()
    for stage in range(3):
        cands = _candidates(dtype, batch, mfma)
        for name, values in _searched(cands):
            times = {}
            for v in values:
                _set(plan, batch, {name: v})
                times[v] = measure()
            v = min(times, key=times.get)
            best[name] = v
            best_t = times[v]
            _set(plan, batch, {name: v})
        if cands and cands[0][0] in ('mfma_min', 'mfma64_min'):
            mfma = best[cands[0][0]] == 1
            continue
        break
    entry = {'batch': int(batch), 'params': best, 'seconds': best_t}
    if use_cache:
        data = load_cache()
        data[key] = entry
        save_cache(data)
        global _loaded
        _loaded = None
    return dict(entry, cached=False, tune_seconds=time.perf_counter() - t0)


def reset(plan):
    '''Back to the library's rules on this plan.'''
    _lib.call('owmi_sht_plan_tune_reset', plan.handle)
