'''
Host logic and oracle pins of the 2-D Fourier-wavelet noise model (CPU):
  * wlm_utils: the known answers of test_wlm_utils.py:8-330 (band limits of
    Fig. 1 of arXiv:1211.1680, sum_j w_j^2 = 1, argument errors);
  * fkernel: test_fkernel.py:13-475 (digitisation, kernel slices with and
    without FFT-length padding, FKernel / FKernelSet, scale-discrete 2-D
    kernels adding up to one in quadrature);
  * dft host helpers: test_dft.py:142-599 (wavenumber axes, slices,
    optimal FFT lengths, cl2flat);
  * oracle: rfft against the DFT sum, irfft(rfft) = 1, contract_fxg against
    the full complex FFT for the four parities of (ny, nx)
    (test_dft.py:261-357) and against the real-space dot product,
    f2wav / wav2f adjoint and a partition of unity
    (test_wavtrans.py:940-1100), a cosine's flat-sky smoothing, the inverse
    Fourier-wavelet operator of a constant covariance, and the Wiener
    operator of that noise model against the pixel-diagonal one.
'''
import numpy as np
import pytest

from oracle import fourier as of, geometry as og, wiener as ow, layout
from optweight_amd import wlm_utils, fkernel, dft, map_utils


# ------------------------------------------------------------- wlm_utils

def test_sd_kernels_known_answers():
    w_ell, lmaxs = wlm_utils.get_sd_kernels(3, 30)
    assert w_ell.shape == (5, 31)
    np.testing.assert_array_equal(lmaxs, [3, 9, 27, 30, 30])
    np.testing.assert_allclose(np.sum(w_ell ** 2, axis=0), 1, atol=1e-6)
    np.testing.assert_array_equal(wlm_utils.get_lmax_array(w_ell), lmaxs)
    w, l = wlm_utils.get_sd_kernels(3, 30, lmin=4)
    np.testing.assert_array_equal(l, [9, 27, 30, 30])
    w, l = wlm_utils.get_sd_kernels(3, 30, jmax=3)
    np.testing.assert_array_equal(l, [3, 9, 27, 30])
    w, l = wlm_utils.get_sd_kernels(3, 30, lmax_j=25)
    np.testing.assert_array_equal(l, [3, 9, 27, 30])
    _, _, js = wlm_utils.get_sd_kernels(3, 30, lmin=4, return_j=True)
    np.testing.assert_array_equal(js, [1, 2, 3, 4])
    assert wlm_utils.lmax_to_j_scale(10, 3) == 3
    assert wlm_utils.j_scale_to_lmax(3, 3) == 81
    assert wlm_utils.j_scale_to_lmax(3, 1.3) == 3
    for kw in (dict(lmin=10, j0=1), dict(lmin=-10), dict(j0=-10), dict(j0=10),
               dict(j0=10, jmax=3, lmax_j=20)):
        with pytest.raises(ValueError):
            wlm_utils.get_sd_kernels(3, 30, **kw)
    np.testing.assert_array_equal(
        wlm_utils.get_lmax_array(np.array([[1, 1, 0, 0], [0, 1, 1, 0], [0, 0, 0, 1],
                                           [0, 0, 0, 0]], dtype=float)), [2, 3, 3, 0])
    big, _ = wlm_utils.get_sd_kernels(3, 6000)
    np.testing.assert_allclose(np.sum(big ** 2, axis=0), 1, atol=1e-10)


def test_lambda_functions():
    ts = np.asarray([0, 0.1, 1 / 3, 0.5, 1, 2])
    s = wlm_utils.s_lambda(ts, 3)
    np.testing.assert_allclose(s[[0, 1, 2, 4]], 0, atol=1e-12)
    assert abs(s[3] - np.exp(-1 / (1 - 0.5 ** 2))) < 1e-12
    k = wlm_utils.k_lambda(ts, 3)
    np.testing.assert_allclose(k[:3], 1)
    assert 0 < k[3] < 1 and abs(k[4]) < 1e-12
    z = wlm_utils.zeta_lambda(ts, 3)
    np.testing.assert_allclose(z[:3], 0, atol=1e-7)
    assert 0 < z[3] < 1 and abs(z[4] - 1) < 1e-12
    np.testing.assert_allclose(wlm_utils.schwarz(np.asarray([-1e40, 1e40])), 0)
    with pytest.raises(ValueError):
        wlm_utils.s_lambda(np.zeros(1), -1)


# --------------------------------------------------------------- fkernel

def test_digitize():
    np.testing.assert_array_equal(fkernel.digitize_1d(np.ones(10)), np.ones(10, bool))
    np.testing.assert_array_equal(fkernel.digitize_1d(np.zeros(10)), np.zeros(10, bool))
    arr = np.sin(np.linspace(0, np.pi / 3, 2000))
    arr /= arr.max()
    assert abs(arr.sum() - fkernel.digitize_1d(arr).sum()) < arr.sum() * 1e-3
    for bad in (np.ones(10) * 2, np.ones(10) - 1.1):
        with pytest.raises(ValueError):
            fkernel.digitize_1d(bad)
    w_ell, _ = wlm_utils.get_sd_kernels(3, 30)
    d_ell, ells = fkernel.digitize_kernels(w_ell, upsamp_fact=10)
    assert d_ell.dtype == bool and d_ell.shape == (5, 310)
    np.testing.assert_allclose(ells, np.linspace(0, 31, 310))
    np.testing.assert_array_equal(d_ell.sum(axis=0), 1)
    with pytest.raises(ValueError):
        fkernel.digitize_kernels(np.ones((3, 5)))


KERNEL_SLICE_CASES = [
    ([[0, 1, 1, 0], [1, 1, 0, 0], [1, 0, 0, 0], [0, 0, 0, 0], [0, 1, 0, 0], [0, 1, 1, 0]],
     False, slice(0, 3), slice(0, 3), slice(-2, None)),
    ([[0, 1, 1, 0], [1, 1, 0, 0], [1, 0, 0, 0], [1, 0, 0, 0], [0, 1, 0, 0], [0, 1, 1, 0]],
     False, slice(0, 3), slice(0, 3), slice(-3, None)),
    ([[0, 1, 1, 1], [1, 1, 0, 0], [1, 0, 0, 0], [1, 0, 0, 0], [0, 1, 0, 0], [0, 1, 1, 0]],
     False, slice(0, 4), slice(0, 3), slice(-3, None)),
    ([[1, 0, 0, 0]] + [[0, 0, 0, 0]] * 5, False, slice(0, 1), slice(0, 1), slice(0, 0)),
    ([[0, 0, 0, 0]] * 5 + [[1, 0, 0, 0]], False, slice(0, 1), slice(0, 2), slice(-1, None)),
    ([[0, 1, 1, 0], [1, 1, 0, 0], [1, 0, 0, 0], [1, 0, 0, 0], [0, 0, 0, 0], [0, 1, 0, 0],
      [0, 1, 1, 0]], False, slice(0, 3), slice(0, 4), slice(-3, None)),
    ([[0, 1, 1, 0], [1, 1, 0, 0], [1, 0, 0, 0], [0, 0, 0, 0], [1, 0, 0, 0], [0, 1, 0, 0],
      [0, 1, 1, 0]], False, slice(0, 3), slice(0, 4), slice(-3, None)),
    ([[0], [1], [0], [0], [1]], True, slice(0, 1), slice(0, 2), slice(-2, None)),
    ([[0], [1], [0], [0], [0], [1]], True, slice(0, 1), slice(0, 2), slice(-2, None)),
    ([[0], [1], [0], [1], [0], [1]], True, slice(0, 1), slice(0, 3), slice(-3, None)),
]


@pytest.mark.parametrize('case', range(len(KERNEL_SLICE_CASES)))
def test_find_kernel_slice(case):
    arr, opt, sx, sy0, sy1 = KERNEL_SLICE_CASES[case]
    (ypos, yneg), x = fkernel.find_kernel_slice(np.asarray(arr), optimize_len=opt)
    assert (x, ypos, yneg) == (sx, sy0, sy1)


def test_find_kernel_slice_errors():
    for bad in (np.zeros((2, 2)), np.ones((1, 2, 2)), np.ones((1, 2, 2)) * 1e-7):
        with pytest.raises(ValueError):
            fkernel.find_kernel_slice(bad)
    assert fkernel._prefix_len(np.asarray([0, 1, 0])) == 2
    assert fkernel._prefix_len(np.asarray([0, 1, 1])) == 3
    assert fkernel._prefix_len(np.asarray([1, 0, 0])) == 1
    assert fkernel._prefix_len(np.zeros(3)) == 0


def test_fkernel_and_set():
    full1 = np.asarray([[0, 1, 1, 0], [1, 1, 0, 0], [1, 0, 0, 0], [0, 0, 0, 0], [0, 1, 0, 0],
                        [0, 1, 1, 0]])
    fk_arr1 = np.asarray([[0, 1, 1], [1, 1, 0], [1, 0, 0], [0, 1, 0], [0, 1, 1]])
    ly, lx = np.asarray([0, 1, 2, -2, -1]), np.asarray([0, 1, 2])
    fk1 = fkernel.FKernel(fk_arr1, ly, lx, slices_y=(slice(0, 3), slice(-2, None)),
                          slice_x=slice(0, 3), shape_full=full1.shape)
    np.testing.assert_allclose(fk1.modlmap(), np.sqrt(ly[:, None] ** 2 + lx[None, :] ** 2))
    np.testing.assert_array_equal(fk1.to_full_array(), full1)
    np.testing.assert_array_equal(fk1.rows(), [0, 1, 2, 4, 5])
    full2 = np.zeros((6, 4), dtype=int)
    full2[0, 0] = 1
    fk2 = fkernel.FKernel(np.asarray([[1]]), np.asarray([0]), np.asarray([0]),
                          slices_y=(slice(0, 1), slice(-1, -1)), slice_x=slice(0, 1),
                          shape_full=full2.shape)
    fks = fkernel.FKernelSet()
    fks[0] = fk1
    fks[1] = fk2
    assert len(fks) == 2 and fks.dtype == np.int64
    assert np.shares_memory(fks[0].fkernel, fk_arr1)
    full = fks.to_full_array()
    np.testing.assert_array_equal(full[0], full1)
    np.testing.assert_array_equal(full[1], full2)
    fks = fks.astype(np.float32)
    assert fks.dtype == np.float32 and fks[1].fkernel.dtype == np.float32
    assert not np.shares_memory(fks[0].fkernel, fk_arr1)
    with pytest.raises(TypeError):
        fks[2] = fk_arr1
    with pytest.raises(ValueError):
        fkernel.FKernel(fk_arr1, ly[:3], lx)


def car_axes(lmax=30, nx=None):
    '''Wavenumber axes of a full-sky CAR map with lmax + 1 rings (both
    poles) and 2 lmax + 1 columns, x steps negative as in CAR maps.'''
    ny = lmax + 1
    nx = 2 * lmax + 1 if nx is None else nx
    wcs = map_utils.CarWCS(cdelt=[-360. / nx, 180. / lmax], crpix=[nx // 2 + 1, ny // 2 + 1])
    return (ny, nx), wcs


@pytest.mark.parametrize('digital', [False, True])
def test_sd_kernels_fourier(digital):
    shape, wcs = car_axes()
    ly, lx = dft.laxes_real(shape, wcs)
    modlmap = dft.laxes2modlmap(ly, lx)
    fks = fkernel.get_sd_kernels_fourier(ly, lx, 3, digital=digital, oversample=10)
    assert len(fks) == 5
    assert fks.dtype == (bool if digital else np.float64)
    arr = fks.to_full_array().astype(np.float64)
    assert arr.shape == (5,) + modlmap.shape
    np.testing.assert_allclose(np.sum(arr ** 2, axis=0), 1, atol=1e-6)


# ------------------------------------------------------------- dft host

def test_dft_host_helpers():
    shape, wcs = car_axes(10, nx=21)
    ly, lx = dft.laxes_real(shape, wcs)
    dy, dx = np.radians(180. / 10), np.radians(-360. / 21)
    np.testing.assert_allclose(ly, 2 * np.pi * np.fft.fftfreq(11, dy))
    np.testing.assert_allclose(lx, 2 * np.pi * np.fft.rfftfreq(21, dx))
    lmap = dft.lmap_real(shape, wcs)
    np.testing.assert_allclose(dft.modlmap_real(shape, wcs), np.hypot(lmap[0], lmap[1]))
    assert [dft.get_optimal_fftlen(n, **kw) for n, kw in
            [(0, {}), (1, {}), (1, dict(even=False)), (3, dict(even=False)),
             (4, dict(even=False)), (11, {}), (11, dict(even=False)), (97, {})]] == \
        [0, 2, 1, 4, 4, 12, 12, 98]
    arr = np.arange(24).reshape(6, 4)
    out, (ly2, lx2) = dft.slice_fmap(arr, (slice(0, 3), slice(-2, None)), slice(0, 3),
                                     laxes=(np.asarray([0, 1, 2, -3, -2, -1]), np.arange(4)))
    np.testing.assert_array_equal(out, arr[[0, 1, 2, 4, 5]][:, :3])
    np.testing.assert_array_equal(ly2, [0, 1, 2, -2, -1])
    modl = dft.laxes2modlmap(*dft.laxes_real((6, 11), map_utils.CarWCS([0.1, 0.1], [0, 0])))
    ells = np.linspace(modl.min(), modl.max(), 100)
    c_ell = (1 + (np.maximum(ells, 10) / 2000) ** -3.)[None, None]
    np.testing.assert_allclose(dft.cl2flat(c_ell, ells, modl)[0, 0],
                               1 + (np.maximum(modl, 10) / 2000) ** -3., rtol=1e-2)
    lw = dft.lwcs_real(shape, wcs)
    np.testing.assert_allclose(lw.wcs.cdelt, [2 * np.pi / (abs(dx) * 21), 2 * np.pi / (dy * 11)])


# ----------------------------------------------------------------- oracle

@pytest.mark.parametrize('ny,nx', [(4, 5), (4, 4), (5, 5), (5, 4), (6, 6)])
def test_oracle_dft_pins(ny, nx):
    rng = np.random.default_rng(ny * 10 + nx)
    f = rng.normal(size=(ny, nx))
    g = rng.normal(size=(ny, nx))
    yy, xx = np.mgrid[:ny, :nx]
    brute = np.array([[np.sum(f * np.exp(-2j * np.pi * (ky * yy / ny + kx * xx / nx)))
                       for kx in range(nx // 2 + 1)] for ky in range(ny)]) / np.sqrt(ny * nx)
    F, G = of.rfft(f), of.rfft(g)
    np.testing.assert_allclose(F, brute, atol=1e-12)
    np.testing.assert_allclose(of.irfft(F, nx), f, atol=1e-12)
    full = np.sum(np.fft.fft2(f, norm='ortho') * np.conj(np.fft.fft2(g, norm='ortho'))).real
    assert abs(of.contract_fxg(F, G) - full) < 1e-12
    assert abs(of.contract_fxg(F, G) - np.sum(f * g)) < 1e-12
    # a dropped column term or a wrong parity shows up here
    Fc = F.copy()
    Fc[..., 0, -1] += 1j
    assert abs(of.contract_fxg(Fc, G) - of.contract_fxg(F, G)) > 0 or nx == 1


def sd_kernel_list(fks):
    return [(widx, np.asarray(fk.fkernel, dtype=np.float64), fk.slices_y, fk.slice_x)
            for widx, fk in fks]


@pytest.mark.parametrize('digital', [False, True])
@pytest.mark.parametrize('nx', [40, 41])
def test_oracle_f2wav_adjoint_and_partition(digital, nx):
    shape, wcs = car_axes(20, nx=nx)
    ly, lx = dft.laxes_real(shape, wcs)
    fks = fkernel.get_sd_kernels_fourier(ly, lx, 2, digital=digital)
    kern = sd_kernel_list(fks)
    rng = np.random.default_rng(3)
    m = rng.normal(size=(2,) + shape)
    F = of.rfft(m)
    wav = of.f2wav(F, kern)
    np.testing.assert_allclose(of.wav2f(wav, kern, F.shape[-2:]), F, atol=1e-10)
    if nx % 2 == 0:
        # the half-plane column lx = nx / 2 counts once in the full plane but
        # twice in the odd-width wavelet grids: adjoint only for odd nx
        return
    w = {j: rng.normal(size=v.shape) for j, v in wav.items()}
    lhs = sum(np.sum(wav[j] * w[j]) for j in wav)
    rhs = of.contract_fxg(F, of.wav2f(w, kern, F.shape[-2:]))
    assert abs(lhs - rhs) < 1e-10 * abs(lhs)


def test_oracle_flat_smoothing_of_a_cosine():
    ny, nx = 12, 16
    yy = np.arange(ny)[:, None] * np.ones((1, nx))
    m = np.cos(2 * np.pi * 2 * yy / ny)
    ly = 2 * np.pi * np.fft.fftfreq(ny, 0.1)
    lx = 2 * np.pi * np.fft.rfftfreq(nx, 0.1)
    modl = dft.laxes2modlmap(ly, lx)
    # a kernel linear in l is sampled exactly by the linear interpolation
    fb = of.cl2flat(1 - np.arange(200) / 400., np.arange(200), modl)
    out = of.irfft(of.rfft(m) * fb, nx)
    np.testing.assert_allclose(out, (1 - abs(ly[2]) / 400.) * m, atol=1e-12)
    cov = of.estimate_cov_pix_flat(np.ones((1, 1, ny, nx)) * 3, 0.2, modl, diag=True)
    np.testing.assert_allclose(cov, 9, atol=1e-12)


def test_oracle_inverse_of_constant_covariance():
    # odd nx: for even nx the inner product's data-dependent parity test sees
    # the Fourier-wavelet outputs as odd and the input as even (as the
    # reference does), so one CG step is not exact there
    shape, wcs = car_axes(16, nx=33)
    ly, lx = dft.laxes_real(shape, wcs)
    kern = sd_kernel_list(fkernel.get_sd_kernels_fourier(ly, lx, 2, digital=True))
    F = of.rfft(np.random.default_rng(5).normal(size=(1,) + shape))
    wav = of.f2wav(F, kern)
    m_wav = {j: np.full((1, v[0].size), 4.0) for j, v in wav.items()}
    inv = of.inv_fwav_op(kern, F.shape[-2:], m_wav, nsteps=1)
    np.testing.assert_allclose(inv(F), F / 4, atol=1e-12)
    np.testing.assert_allclose(of.fwav_op(kern, F.shape[-2:], m_wav)(F), 4 * F, atol=1e-12)


def test_oracle_fwav_wiener_equals_white_noise():
    lmax = 8
    ny, nx = lmax + 1, 2 * lmax + 1
    minfo = map_utils.MapInfo.map_info_clenshaw_curtis(ny, nx)
    geom = og.Geom(minfo.theta, minfo.weight, minfo.nphi, phi0=minfo.phi0)
    shape, wcs = car_axes(lmax, nx=nx)
    ly, lx = dft.laxes_real(shape, wcs)
    kern = sd_kernel_list(fkernel.get_sd_kernels_fourier(ly, lx, 2, digital=True))
    F = of.rfft(np.zeros((1, ny, nx)))
    m_wav = {j: np.full((1, v[0].size), 2.0) for j, v in of.f2wav(F, kern).items()}
    rng = np.random.default_rng(2)
    imap = rng.normal(size=(1, minfo.npix))
    icov_ell = np.ones((1, lmax + 1)) * 0.5
    sol = of.from_arrays_fwav(imap, geom, lmax, ny, nx, icov_ell, m_wav, kern, spin=0)
    ref = ow.WienerMap(imap, ow.ell_op(icov_ell, lmax), ow.pix_op(np.full((1, minfo.npix), 0.5)),
                       (ow.y_op(lmax, geom, 0), ow.yt_op(geom, lmax, 0)))
    np.testing.assert_allclose(sol.b0, ref.b0, atol=1e-12 * np.abs(ref.b0).max())
    # complex alm: with real alm the band-limited map's coefficient at
    # (ly = 0, lx = last) is real and the parity test of the inner product
    # takes odd nx for even
    a = rng.normal(size=(1, layout.nelem(lmax))) + 1j * rng.normal(size=(1, layout.nelem(lmax)))
    a[:, :lmax + 1] = a[:, :lmax + 1].real
    np.testing.assert_allclose(sol.a_matrix(a), ref.a_matrix(a), atol=1e-11)
