'''
Synthetic Wiener-filter problems on Gauss-Legendre grids (seeded, host
numpy), used by the benchmark, the tests and the smoke check.

  temperature(lmax, ...)   T only, GL (lmax+1) x (2 lmax+1):
      C_l = 1e3 (l + 10)^-2.5 for l >= 2, 0 below; icov_l = pinv(C_l);
      data = mask * (Y a + n), a = C^(1/2) unit-variance alm, n white with
      sigma per pixel; icov_pix = w / sigma^2, zero in the mask.
  polarized(lmax, ...)     TQU on the same grid:
      TT as above, EE = 0.05 TT, BB = 0.005 TT, TE = 0.3 sqrt(TT EE);
      icov_T = depth * w / sigma^2 with depth = exp(g), g a seeded smooth
      Gaussian field (l < 30) mapped affinely onto [ln 0.5, ln 2];
      icov_Q = icov_U = icov_T / 2 (Q/U noise sqrt(2) x T).
  mask: zero inverse noise for |latitude| < 10 deg and inside ndisk disks
  of radius 0.5 deg centred on uniformly random sphere points.

The signal map Y a is synthesised with the device transform (``synth``
callable, default the HIP SHT); the noise and everything else is numpy.
'''
import numpy as np
import torch
from scipy.special import sph_harm_y

from optweight_amd import map_utils, alm_utils, sht, mat_utils


def cov_ell_T(lmax):
    ell = np.arange(lmax + 1, dtype=np.float64)
    c = np.zeros(lmax + 1)
    c[2:] = 1e3 * (ell[2:] + 10.) ** -2.5
    return c


def cov_ell_TQU(lmax):
    tt = cov_ell_T(lmax)
    ee = 0.05 * tt
    bb = 0.005 * tt
    te = 0.3 * np.sqrt(tt * ee)
    cov = np.zeros((3, 3, lmax + 1))
    cov[0, 0] = tt
    cov[1, 1] = ee
    cov[2, 2] = bb
    cov[0, 1] = cov[1, 0] = te
    return cov


def pinv_ell(cov):
    '''Per-l pseudo-inverse (as optweight's filters.py:107-109).'''
    cov = mat_utils.full_matrix(np.asarray(cov))
    out = np.zeros_like(cov)
    for l in range(cov.shape[-1]):
        out[:, :, l] = np.linalg.pinv(cov[:, :, l])
    return out


def gl_minfo(lmax):
    return map_utils.MapInfo.map_info_gauss_legendre(lmax + 1, 2 * lmax + 1)


def pixel_coords(minfo):
    theta = np.repeat(minfo.theta, minfo.nphi)
    phi = np.concatenate([minfo.phi0[r] + 2 * np.pi * np.arange(minfo.nphi[r]) / minfo.nphi[r]
                          for r in range(minfo.nrow)])
    return theta, phi


def make_mask(minfo, seed, ndisk=50, rdisk_deg=0.5, band_deg=10.):
    '''Observed = True. Galactic-like band plus point-source disks (each disk
    is tested only on the rings within its radius, which holds every
    pixel it can contain).'''
    rng = np.random.default_rng(seed)
    theta = np.repeat(minfo.theta, minfo.nphi)
    lat = np.pi / 2 - theta
    obs = np.abs(lat) >= np.radians(band_deg)
    z = rng.uniform(-1, 1, ndisk)
    ph = rng.uniform(0, 2 * np.pi, ndisk)
    th = np.arccos(z)
    rad = np.radians(rdisk_deg)
    cosr = np.cos(rad)
    for k in range(ndisk):
        rows = np.nonzero(np.abs(minfo.theta - th[k]) <= rad + 1e-9)[0]
        for r in rows:
            n = int(minfo.nphi[r])
            sl = slice(int(minfo.offsets[r]), int(minfo.offsets[r]) + n)
            phi = minfo.phi0[r] + 2 * np.pi * np.arange(n) / n
            t = minfo.theta[r]
            cosd = np.sin(t) * np.sin(th[k]) * np.cos(phi - ph[k]) + np.cos(t) * np.cos(th[k])
            obs[sl] &= cosd < cosr
    return obs


def smooth_field(minfo, seed, lmax_s=29):
    '''Seeded Gaussian field with unit C_l for l <= lmax_s, evaluated on the
    rings directly from scipy's Y_lm (no transform library involved).'''
    ainfo = alm_utils.AlmInfo(lmax_s)
    a = alm_utils.unit_var_alm(ainfo, (1,), seed)[0]
    ell, em = ainfo.lm_arrays()
    out = np.zeros(minfo.npix)
    for r in range(minfo.nrow):
        th = minfo.theta[r]
        n = minfo.nphi[r]
        ph = minfo.phi0[r] + 2 * np.pi * np.arange(n) / n
        F = np.zeros(lmax_s + 1, dtype=np.complex128)
        y = sph_harm_y(ell, em, th, 0.0)
        np.add.at(F, em, a * y)
        mm = np.arange(lmax_s + 1)
        fac = np.where(mm == 0, 1.0, 2.0)
        vals = np.real(np.exp(1j * np.outer(ph, mm)) @ (fac * F))
        out[minfo.offsets[r]:minfo.offsets[r] + n] = vals
    return out


def depth_map(minfo, seed, lo=0.5, hi=2.0):
    g = smooth_field(minfo, seed)
    g = (g - g.min()) / (g.max() - g.min())
    return np.exp(np.log(lo) + g * (np.log(hi) - np.log(lo)))


def _synth_device(alm, minfo, lmax, spin, device):
    ainfo = alm_utils.AlmInfo(lmax)
    a = torch.as_tensor(alm, device=device)
    out = torch.zeros((alm.shape[0], minfo.npix), dtype=torch.float64, device=device)
    sht.alm2map(a, out, ainfo, minfo, spin)
    return out.cpu().numpy()


class Problem:
    '''Arrays of one Wiener-filter problem (host numpy).'''

    def __init__(self, lmax, minfo, icov_ell, icov_pix, imap, mask_bool, spin, cov_ell,
                 signal_alm):
        self.lmax = lmax
        self.minfo = minfo
        self.ainfo = alm_utils.AlmInfo(lmax)
        self.icov_ell = icov_ell
        self.icov_pix = icov_pix
        self.imap = imap
        self.mask_bool = mask_bool
        self.spin = spin
        self.cov_ell = cov_ell
        self.signal_alm = signal_alm


def temperature(lmax, seed=0, sigma=10., ndisk=50, synth=None, device=None):
    minfo = gl_minfo(lmax)
    cov = cov_ell_T(lmax)[None, :]
    icov_ell = np.zeros_like(cov)
    icov_ell[0, 2:] = 1 / cov[0, 2:]
    ss = np.random.SeedSequence([1, seed])
    s_alm, s_noise, s_mask = [np.random.default_rng(s) for s in ss.spawn(3)]
    a = alm_utils.rand_alm(cov, alm_utils.AlmInfo(lmax), s_alm)
    if synth is None:
        dev = device or torch.device('cuda', torch.cuda.current_device())
        sig = _synth_device(a, minfo, lmax, 0, dev)
    else:
        sig = synth(a, minfo, lmax, 0)
    obs = make_mask(minfo, s_mask, ndisk=ndisk)
    w = np.repeat(minfo.weight, minfo.nphi)
    icov_pix = (w / sigma ** 2 * obs)[None, :]
    noise = s_noise.standard_normal(minfo.npix) * sigma
    imap = ((sig[0] + noise) * obs)[None, :]
    return Problem(lmax, minfo, icov_ell, icov_pix, imap, obs[None, :], 0, cov, a)


def polarized(lmax, seed=0, sigma=10., ndisk=50, synth=None, device=None):
    minfo = gl_minfo(lmax)
    cov = cov_ell_TQU(lmax)
    icov_ell = pinv_ell(cov)
    ss = np.random.SeedSequence([2, seed])
    s_alm, s_noise, s_mask, s_depth = [np.random.default_rng(s) for s in ss.spawn(4)]
    a = alm_utils.rand_alm(cov, alm_utils.AlmInfo(lmax), s_alm)
    if synth is None:
        dev = device or torch.device('cuda', torch.cuda.current_device())
        sig = _synth_device(a, minfo, lmax, [0, 2], dev)
    else:
        sig = synth(a, minfo, lmax, [0, 2])
    obs = make_mask(minfo, s_mask, ndisk=ndisk)
    depth = depth_map(minfo, s_depth)
    w = np.repeat(minfo.weight, minfo.nphi)
    icov_t = depth * w / sigma ** 2 * obs
    icov_pix = np.stack([icov_t, icov_t / 2, icov_t / 2])
    sig_pix = np.stack([sigma / np.sqrt(depth), sigma * np.sqrt(2 / depth),
                        sigma * np.sqrt(2 / depth)])
    noise = s_noise.standard_normal((3, minfo.npix)) * sig_pix
    imap = (sig + noise) * obs
    return Problem(lmax, minfo, icov_ell, icov_pix, imap, np.stack([obs] * 3), [0, 2], cov, a)


def car_polarized(ny, nx, lmax, seed=0, sigma=10., ndisk=50, fwhm_arcmin=10., device=None):
    """Inputs of the two-phase CAR filter (filters.CGPixFilter): a full-sky
    Clenshaw-Curtis CAR map of ny rows and nx columns in pixell's convention
    (RA decreasing with x, dec increasing with y), the TQU spectra of
    cov_ell_TQU as theory spectra, a Gaussian beam, the depth and mask models
    of ``polarized`` on the map's rings, beam-convolved signal plus white
    noise. Returns a dict with theory_cls, b_ell, icov_pix (3, ny, nx),
    mask (ny, nx), imap (3, ny, nx), wcs and lmax."""
    wcs = map_utils.CarWCS([-360. / nx, 180. / (ny - 1)], [nx / 2 + 1, (ny + 1) / 2])
    minfo, index = map_utils.unit_stride(map_utils.match_enmap_minfo((ny, nx), wcs, mtype='CC'))
    cov = cov_ell_TQU(lmax)
    cls = {'TT': cov[0, 0].copy(), 'EE': cov[1, 1].copy(), 'BB': cov[2, 2].copy(),
           'TE': cov[0, 1].copy()}
    ell = np.arange(lmax + 1)
    sig_b = np.radians(fwhm_arcmin / 60.) / np.sqrt(8 * np.log(2))
    b_ell = np.exp(-0.5 * ell * (ell + 1) * sig_b ** 2)
    ss = np.random.SeedSequence([4, seed])
    s_alm, s_noise, s_mask, s_depth = [np.random.default_rng(s) for s in ss.spawn(4)]
    a = alm_utils.rand_alm(cov, alm_utils.AlmInfo(lmax), s_alm)
    a = a * b_ell[alm_utils.AlmInfo(lmax).lm_arrays()[0]]
    dev = device or torch.device('cuda', torch.cuda.current_device())
    sig = _synth_device(a, minfo, lmax, [0, 2], dev)
    obs = make_mask(minfo, s_mask, ndisk=ndisk)
    depth = depth_map(minfo, s_depth)
    w = np.repeat(minfo.weight, minfo.nphi)
    icov_t = depth * w / sigma ** 2 * obs
    icov_c = np.stack([icov_t, icov_t / 2, icov_t / 2])
    sig_pix = np.stack([sigma / np.sqrt(depth), sigma * np.sqrt(2 / depth),
                        sigma * np.sqrt(2 / depth)])
    imap_c = (sig + s_noise.standard_normal((3, minfo.npix)) * sig_pix) * obs

    def to_map(c):
        out = np.zeros(c.shape[:-1] + (ny * nx,), dtype=c.dtype)
        out[..., index] = c
        return out.reshape(c.shape[:-1] + (ny, nx))

    return {'theory_cls': cls, 'b_ell': b_ell, 'icov_pix': to_map(icov_c),
            'mask': to_map(obs), 'imap': to_map(imap_c), 'wcs': wcs, 'lmax': lmax}


def batch_data(prob, nbatch, seed=0, synth=None, device=None):
    """(nbatch, npol, npix) data maps of the problem's signal and noise
    model: right-hand side i has its own signal draw from prob.cov_ell and
    white noise of variance w / icov_pix on observed pixels (seed sequence
    [3, seed, i])."""
    npol = prob.icov_pix.shape[0]
    w = np.repeat(prob.minfo.weight, prob.minfo.nphi)
    obs = prob.icov_pix > 0
    sig_pix = np.zeros_like(prob.icov_pix)
    sig_pix[obs] = np.sqrt((w * np.ones((npol, 1)))[obs] / prob.icov_pix[obs])
    out = np.zeros((nbatch, npol, prob.minfo.npix))
    for i in range(nbatch):
        s_alm, s_noise = [np.random.default_rng(s)
                          for s in np.random.SeedSequence([3, seed, i]).spawn(2)]
        a = alm_utils.rand_alm(prob.cov_ell, prob.ainfo, s_alm)
        if synth is None:
            dev = device or torch.device('cuda', torch.cuda.current_device())
            sig = _synth_device(a, prob.minfo, prob.lmax, prob.spin, dev)
        else:
            sig = synth(a, prob.minfo, prob.lmax, prob.spin)
        out[i] = (sig + s_noise.standard_normal((npol, prob.minfo.npix)) * sig_pix) * obs
    return out


def ragged(lmaxs=(255, 511, 1023, 2047), per=16, seed=0, device=None, synth=None):
    """Batched constrained realisations of uneven band limit: for each lmax a
    polarized problem (same spectra, depth and mask construction) and `per`
    data maps (batch_data); the solvers draw S^-1/2 w and N^-1/2 w' per
    right-hand side from the seed sequence [seed, i]. Returns
    [(problem, data (per, 3, npix))]."""
    out = []
    for k, lmax in enumerate(lmaxs):
        prob = polarized(lmax, seed=seed + k, synth=synth, device=device)
        out.append((prob, batch_data(prob, per, seed=seed + k, synth=synth, device=device)))
    return out


def log_bands(lmin, lmax, nband):
    """Edges of nband log-spaced multipole bands covering lmin <= l <= lmax
    (band k: edges[k] <= l < edges[k + 1]): rounded geometric edges, each at
    least one above the previous."""
    edges = np.round(np.geomspace(lmin, lmax + 1, nband + 1)).astype(int)
    for k in range(1, edges.size):
        edges[k] = max(edges[k], edges[k - 1] + 1)
    if edges[-1] != lmax + 1:
        raise ValueError(f'{nband} bands do not fit between l = {lmin} and {lmax}')
    return edges


def needlet_kernels(lmax, nwav=12, lamb=2.0):
    """Cosine-needlet kernels on a log_lamb(l) axis, partition of unity
    sum_j w_j(l)^2 = 1: w_j = cos(pi/2 (t - j)) for |t - j| < 1 with
    t = log_lamb(l); the first kernel is 1 below its peak, the last 1 above."""
    ells = np.arange(lmax + 1, dtype=np.float64)
    t = np.full(lmax + 1, -1e300)
    t[1:] = np.log(ells[1:]) / np.log(lamb)
    w = np.zeros((nwav, lmax + 1))
    for j in range(nwav):
        core = np.where(np.abs(t - j) < 1, np.cos(np.pi / 2 * np.clip(t - j, -1, 1)), 0.0)
        if j == 0:
            core = np.where(t <= 0, 1.0, core)
        if j == nwav - 1:
            core = np.where(t >= j, 1.0, core)
        w[j] = core
    w[np.abs(w) < 1e-14] = 0
    return w


class WaveletProblem:
    """Arrays of the temperature Wiener problem with a wavelet-domain noise
    model (host numpy, data alm on the device)."""

    def __init__(self, **kw):
        self.__dict__.update(kw)


def wavelet_temperature(lmax=4095, nwav=12, lamb=2.0, seed=0, sigma=10., ndisk=200,
                        rdisk_deg=0.1, map_dtype=np.float32, device=None):
    """T-only problem: per-scale inverse noise maps icov_j = depth_j W_j /
    sigma^2 on GL(lmax_j + 1, 2 lmax_j + 1), depth log-normal per pixel
    (sigma 0.3) and 3x the noise variance on the 30 % of the sky with
    cos(theta) > 0.4; |lat| < 10 deg plus `ndisk` holes masked; data =
    mask (signal + noise) in harmonic space."""
    from optweight_amd import wavtrans, operators, alm_utils as au
    if device is None:
        device = torch.device('cuda', torch.cuda.current_device())
    ss = np.random.SeedSequence([4, seed])
    s_alm, s_noise, s_mask, s_depth = [np.random.default_rng(x) for x in ss.spawn(4)]
    minfo = gl_minfo(lmax)
    ainfo = au.AlmInfo(lmax)
    w_ell = needlet_kernels(lmax, nwav, lamb)
    icov_wav = wavtrans.Wav(2, preshape=(1,), dtype=map_dtype)
    cov_wav = wavtrans.Wav(2, preshape=(1,), dtype=map_dtype)
    for j in range(nwav):
        lj = wavtrans.wav_lmax(w_ell, j)
        mi = gl_minfo(lj)
        w = np.repeat(mi.weight, mi.nphi)
        depth = np.exp(0.3 * s_depth.standard_normal(mi.npix))
        th = np.repeat(mi.theta, mi.nphi)
        depth[np.cos(th) > 0.4] /= 3.0
        icov = (depth * w / sigma ** 2)[None, :].astype(map_dtype)
        icov_wav.add((j, j), icov, mi)
        cov_wav.add((j, j), (1.0 / icov.astype(np.float64)).astype(map_dtype), mi)
    mask = make_mask(minfo, s_mask, ndisk=ndisk, rdisk_deg=rdisk_deg).astype(np.float64)[None]
    cov = cov_ell_T(lmax)[None, :]
    icov_ell = np.zeros_like(cov)
    icov_ell[0, 2:] = 1 / cov[0, 2:]
    cdt = np.complex64 if map_dtype == np.float32 else np.complex128
    sig = torch.as_tensor(au.rand_alm(cov, ainfo, s_alm, dtype=cdt), device=device)
    noise = au.rand_alm_wav(cov_wav, ainfo, w_ell, 0, s_noise, adjoint=False,
                            device=device).to(sig.dtype)
    mask_op = operators.PixMatVecAlm(ainfo, mask, minfo, 0, use_weights=True)
    alm_data = mask_op(sig + noise)
    return WaveletProblem(lmax=lmax, minfo=minfo, ainfo=ainfo, w_ell=w_ell, icov_wav=icov_wav,
                          mask=mask, icov_ell=icov_ell, cov_ell=cov, alm_data=alm_data)
