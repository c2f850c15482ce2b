'''
The Legendre autotuner (optweight_amd/autotune.py): the cache file round
trip and key on the CPU; on the GPU, tuning of a float32 batch (matrix-core
choices) and a float64 transform (vector-kernel choices) with a temporary
cache, the cached result re-applied without a search (also to a new plan
of the same geometry), and transforms with the tuned choices equal to the
library's defaults (float64 1e-12, float32 1e-5: the variants only reorder
float sums).
'''
import json

import numpy as np
import pytest
import torch

from optweight_amd import autotune, map_utils, alm_utils, sht


def test_cache_roundtrip(tmp_path, monkeypatch):
    path = tmp_path / 'tune.json'
    monkeypatch.setenv('OWMI_TUNE_CACHE', str(path))
    assert autotune.load_cache() == {}
    data = {'lmax=63;nrings=64;rings=abc;dtype=float32;batch=16;device=x':
            {'batch': 16, 'params': {'adj_ct2': 4, 'mfma_min': 1}, 'seconds': 1e-3}}
    autotune.save_cache(data)
    assert autotune.load_cache() == data
    assert json.loads(path.read_text()) == data
    # a corrupt file reads as empty
    path.write_text('{not json')
    assert autotune.load_cache() == {}


def test_geometry_key_distinguishes_rings():
    m1 = map_utils.MapInfo.map_info_gauss_legendre(64, 127)
    m2 = map_utils.MapInfo.map_info_gauss_legendre(64, 129)
    k1 = autotune.geometry_key(m1, 63)
    assert k1 == autotune.geometry_key(map_utils.MapInfo.map_info_gauss_legendre(64, 127), 63)
    assert k1 != autotune.geometry_key(m2, 63)
    assert k1 != autotune.geometry_key(m1, 62)
    assert k1 != autotune.geometry_key(m1, 63, rings=np.arange(10))
    assert k1.startswith('lmax=63;nrings=64;')


def test_single_value_parameters_are_not_set():
    # at batch 1 the transforms-per-thread choices have one candidate; a
    # tuned entry applies to every larger batch, so they must stay unset
    for dt in (torch.float32, torch.float64):
        names = [n for n, _ in autotune._searched(autotune._candidates(dt, 1, None))]
        assert 'synth_nb' not in names and 'adj_nb' not in names
        assert 'adj_k0' in names and 'adj_k2' in names
    names = [n for n, _ in autotune._searched(autotune._candidates(torch.float64, 2, False))]
    assert 'synth_nb' in names and 'adj_nb' in names
    # float64 batches: the matrix-core choice covers both directions; the
    # vector kernels' choices follow only when the vector kernels win
    assert autotune._candidates(torch.float64, 128, None) == [('mfma64_min', [1, 129])]
    assert autotune._candidates(torch.float64, 128, True) == []
    names = [n for n, _ in autotune._searched(autotune._candidates(torch.float64, 128, False))]
    assert 'synth_nb' in names and 'adj_k2' in names


def rand_alm(ntrans, lmax, seed):
    rng = np.random.default_rng(seed)
    n = (lmax + 1) * (lmax + 2) // 2
    a = rng.normal(size=(ntrans, 3, n)) + 1j * rng.normal(size=(ntrans, 3, n))
    a[..., :lmax + 1] = a[..., :lmax + 1].real
    return a


def transform_pair(a, minfo, ainfo):
    m = torch.zeros(a.shape[:2] + (minfo.npix,), device=a.device,
                    dtype=torch.float64 if a.dtype == torch.complex128 else torch.float32)
    sht.alm2map(a, m, ainfo, minfo, [0, 2])
    back = torch.zeros_like(a)
    sht.map2alm(m, back, minfo, ainfo, [0, 2])
    return m.cpu().numpy(), back.cpu().numpy()


@pytest.mark.gpu
@pytest.mark.parametrize('dtype,batch', [(torch.complex64, 16), (torch.complex128, 4)])
def test_tune_apply_and_equal_results(dev, tmp_path, monkeypatch, dtype, batch):
    monkeypatch.setenv('OWMI_TUNE_CACHE', str(tmp_path / 'tune.json'))
    lmax = 95
    minfo = map_utils.MapInfo.map_info_gauss_legendre(lmax + 1, 2 * lmax + 1)
    ainfo = alm_utils.AlmInfo(lmax)
    a = torch.as_tensor(rand_alm(batch, lmax, 3), device=dev).to(dtype)
    plan = sht.get_plan(minfo, lmax, dev)
    autotune.reset(plan)
    ref = transform_pair(a, minfo, ainfo)
    res = autotune.tune(minfo, lmax, dtype, batch, device=dev, reps=2)
    print(f'tuned {dtype} batch {batch}: {res}', flush=True)
    assert not res['cached'] and res['params']
    got = transform_pair(a, minfo, ainfo)
    tol = 1e-12 if dtype == torch.complex128 else 1e-5
    for g, r in zip(got, ref):
        assert np.linalg.norm(g - r) <= tol * np.linalg.norm(r)
    # the cache holds the entry; a second call applies it without a search
    again = autotune.tune(minfo, lmax, dtype, batch, device=dev)
    assert again['cached'] and again['params'] == res['params']
    # a new plan of the same geometry gets the cached choices at creation
    minfo2 = map_utils.MapInfo.map_info_gauss_legendre(lmax + 1, 2 * lmax + 1)
    got2 = transform_pair(a, minfo2, ainfo)
    for g, r in zip(got2, got):
        assert np.linalg.norm(g - r) <= 1e-15 * np.linalg.norm(r) + 0.0
