'''
GPU parity of the noise-model estimation (noise_utils.estimate_cov_pix /
estimate_cov_wav / norm_cov_est) against the oracle, float64: relative L2
1e-10 for the smoothed, normalised and positive-semi-definite covariance
maps (diag and full, spin 0 and TQU, band-limited or not), and the input
alm left unchanged by the wavelet estimate.
'''
import numpy as np
import pytest
import torch

from oracle import geometry as og, sht as osht, layout, noise as on
from optweight_amd import noise_utils, map_utils, alm_utils

pytestmark = pytest.mark.gpu


def rel(a, b):
    a = a.cpu().numpy() if isinstance(a, torch.Tensor) else np.asarray(a)
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def noise_alm(lmax, shape, seed):
    rng = np.random.default_rng(seed)
    a = rng.normal(size=shape + (layout.nelem(lmax),)) + 1j * rng.normal(
        size=shape + (layout.nelem(lmax),))
    ell, em = layout.lm_arrays(lmax)
    a[..., em == 0] = a[..., em == 0].real
    if shape[-1] == 3:
        a[..., 1:, ell < 2] = 0
    return a


@pytest.mark.parametrize('diag', [True, False])
@pytest.mark.parametrize('npol,spin', [(1, 0), (3, [0, 2])])
@pytest.mark.parametrize('band', ['lmax', 'kernel', 'none'])
def test_estimate_cov_pix_parity(dev, diag, npol, spin, band):
    print(f'estimate_cov_pix diag={diag} npol={npol} band={band}', flush=True)
    lmax = 20
    g = og.gauss_legendre(lmax + 1, 2 * lmax + 1)
    minfo = map_utils.MapInfo.map_info_gauss_legendre(lmax + 1, 2 * lmax + 1)
    a = noise_alm(lmax, (2, npol), 1)
    m = np.zeros((2, npol, g.npix))
    for c in range(2):
        osht.alm2map(a[c], m[c], lmax, g, spin)
    kw = {'lmax': dict(lmax=lmax), 'kernel': dict(kernel_ell=np.linspace(1, 0.2, lmax + 1)),
          'none': dict(fwhm=0.3)}[band]
    ref = on.estimate_cov_pix(m, g, diag=diag, **kw)
    got = noise_utils.estimate_cov_pix(torch.as_tensor(m, device=dev), minfo, diag=diag, **kw)
    assert tuple(got.shape) == ref.shape
    assert rel(got, ref) < 1e-10


@pytest.mark.parametrize('diag', [True, False])
def test_estimate_cov_wav_parity(dev, diag):
    print(f'estimate_cov_wav diag={diag}', flush=True)
    lmax = 40
    ell = np.arange(lmax + 1)
    w_ell = np.zeros((3, lmax + 1))
    edges = [0, 10, 20, 40]
    for j in range(3):
        lo, hi = edges[j], edges[j + 1]
        w_ell[j, lo:hi + 1] = np.sin(np.pi * (ell[lo:hi + 1] - lo) / (hi - lo)) + (j == 0) * 0.5
    w_ell = np.clip(w_ell, 0, None)
    a = noise_alm(lmax, (1, 3), 2)
    at = torch.as_tensor(a, device=dev)
    ref = on.estimate_cov_wav(a, lmax, w_ell, [0, 2], diag=diag)
    got = noise_utils.estimate_cov_wav(at, alm_utils.AlmInfo(lmax), w_ell, [0, 2], diag=diag)
    assert rel(at, a) == 0, 'input alm changed'
    for j, (g, cov) in enumerate(ref):
        assert np.allclose(got.minfos[(j, j)].theta, g.theta)
        assert rel(got.maps[(j, j)], cov) < 1e-10, j
