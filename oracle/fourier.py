'''
2-D Fourier-wavelet noise model (oracle; test infrastructure only, numpy,
float64), restating optweight's
  dft.rfft / irfft (dft.py:13-67), fmul_2d (332-365), contract_fxg
      (449-496), slice_fmap (498-533), add_to_fmap (535-560)
  wavtrans.f2wav / wav2f (wavtrans.py:687-745)
  operators.WavMatVecWav (914-989), FMatVecF (259-309), FWavMatVecF
      (311-362), InvFWavMatVecF (364-421), FInvFWavFMatVecMap (423-486),
      InvSqrtFWavMatVecWav (488-560)
  noise_utils.estimate_cov_fwav (79-142) with estimate_cov_pix's flat-sky
      smoothing (144-226) and map_utils.fmul_pix (1751-1793)
  solvers.CGWienerMap.from_arrays_fwav (577-695)
  preconditioners.PseudoInvPreconditionerFWav (397-556)

Transforms are numpy's orthonormal rfft2 / irfft2 (the latter uses only
the real parts of columns lx = 0 and, for even nx, lx = nx / 2 after the
inverse pass along y, as ducc0's c2r does). After the forward transform
the self-conjugate coefficients (ly in {0, ny/2}, lx in {0, nx/2}), which
are real in exact arithmetic, are set exactly real: ducc0's exact
twiddles give exact zeros there for power-of-two sizes, and contract_fxg
decides the parity of nx from them.

A kernel set is a list of (widx, k (ny_j, nx_j), (slice_ypos, slice_yneg),
slice_x); a wavelet vector a dict widx -> (..., ny_j, 2 (nx_j - 1) + 1).
'''
import numpy as np

from oracle import ops, sht as osht
from oracle.cg import CG
from oracle.wiener import WienerMap, ell_op, pix_op, y_op, yt_op, harmonic_matrix


def rfft(imap):
    imap = np.asarray(imap, dtype=np.float64)
    ny, nx = imap.shape[-2:]
    f = np.fft.rfft2(imap, norm='ortho')
    rows = [0] + ([ny // 2] if ny % 2 == 0 else [])
    cols = [0] + ([nx // 2] if nx % 2 == 0 else [])
    for r in rows:
        for c in cols:
            f[..., r, c] = f[..., r, c].real
    return f


def irfft(fmap, nx):
    ny = fmap.shape[-2]
    return np.fft.irfft2(fmap, s=(ny, nx), norm='ortho')


def fmul_2d(fmap, fmat2d):
    fmap = np.asarray(fmap)
    while fmap.ndim < 3:
        fmap = fmap[np.newaxis]
    fmat2d = np.asarray(fmat2d)
    if fmat2d.ndim in (2, 3):
        return fmap * fmat2d
    return np.einsum('ablk, blk -> alk', fmat2d, fmap)


def contract_fxg(fmap, gmap):
    '''2 sum Re(f g*) - column lx = 0 - [nx even] column lx = last, nx even
    when f is exactly real at (0, last) and, for even ny, (ny/2, last).'''
    fmap = np.asarray(fmap)
    gmap = np.conj(np.asarray(gmap))
    if fmap.shape != gmap.shape:
        raise ValueError('shape mismatch')
    total = 2 * complex(np.sum(fmap * gmap))
    total -= np.sum(fmap[..., :, 0] * gmap[..., :, 0])
    ny = fmap.shape[-2]
    even = np.all(np.isreal(fmap[..., 0, -1]))
    if ny % 2 == 0:
        even &= np.all(np.isreal(fmap[..., ny // 2, -1]))
    if even:
        total -= np.sum(fmap[..., :, -1] * gmap[..., :, -1])
    return float(np.real(total))


def slice_fmap(fmap, slices_y, slice_x):
    return np.concatenate((fmap[..., slices_y[0], slice_x], fmap[..., slices_y[1], slice_x]),
                          axis=-2)


def add_to_fmap(full, small, slices_y, slice_x):
    full[..., slices_y[0], slice_x] += small[..., slices_y[0], :]
    full[..., slices_y[1], slice_x] += small[..., slices_y[1], :]
    return full


def f2wav(fmap, kernels):
    out = {}
    for widx, k, sy, sx in kernels:
        nx_j = k.shape[1]
        out[widx] = irfft(slice_fmap(fmap, sy, sx) * k, 2 * (nx_j - 1) + 1)
    return out


def wav2f(wav, kernels, shape_full):
    lead = next(iter(wav.values())).shape[:-2]
    fmap = np.zeros(lead + tuple(shape_full), dtype=np.complex128)
    for widx, k, sy, sx in kernels:
        add_to_fmap(fmap, rfft(wav[widx]) * k, sy, sx)
    return fmap


def wav_mat(m_wav, power=1):
    '''{j: block (j, j)} raised to power (matpow, diagonal form kept).'''
    return {j: (m if power == 1 else ops.matpow(m, power, return_diag=True))
            for j, m in m_wav.items()}


def wav_matvec(mats, wav):
    out = {}
    for j, v in wav.items():
        m = np.asarray(mats[j])
        out[j] = np.einsum('ijk, jk -> ik', m, v) if m.ndim == 3 else m * v
    return out


def _flat(wav_j):
    return wav_j.reshape(wav_j.shape[:-2] + (-1,))


def fwav_op(kernels, shape_full, m_wav, power=1):
    '''K^T F M_wav^p F^T K on 2-D coefficients (npol, ny, nlx); m_wav
    {j: (npol, npol, npix_j) or (npol, npix_j)} on flattened wavelet maps.'''
    mats = wav_mat(m_wav, power)

    def f(fmap):
        wav = f2wav(fmap, kernels)
        shapes = {j: w.shape for j, w in wav.items()}
        prod = wav_matvec(mats, {j: _flat(w) for j, w in wav.items()})
        return wav2f({j: prod[j].reshape(shapes[j]) for j in prod}, kernels, shape_full)
    return f


def inv_fwav_op(kernels, shape_full, m_wav, nsteps=3):
    a_op = fwav_op(kernels, shape_full, m_wav, 1)
    m_op = fwav_op(kernels, shape_full, m_wav, -1)

    def f(fmap):
        cg = CG(a_op, np.array(fmap, copy=True), M=m_op, dot=contract_fxg)
        for _ in range(nsteps):
            cg.step()
        return cg.x
    return f


def x_op(fmat2d, power):
    if power == 0:
        return lambda f: f
    fmat2d = np.asarray(fmat2d, dtype=np.float64)
    if power != 1:
        flat = fmat2d.reshape(fmat2d.shape[:-2] + (-1,))
        fmat2d = ops.matpow(flat, power, return_diag=True).reshape(fmat2d.shape)
    return lambda f: fmul_2d(f, fmat2d)


def finv_fwav_map_op(ny, nx, kernels, m_wav, fmat2d, power_x=1, nsteps=3):
    '''F^T X^q (K^T F M_wav F^T K)^-1 X^q F on maps (npol, ny nx).'''
    inv = inv_fwav_op(kernels, (ny, nx // 2 + 1), m_wav, nsteps)
    xo = x_op(fmat2d, power_x)

    def f(imap):
        imap = np.asarray(imap)
        fm = rfft(imap.reshape(imap.shape[:-1] + (ny, nx)))
        fm = xo(inv(xo(fm)))
        return irfft(fm, nx).reshape(imap.shape)
    f.inv = inv
    f.x = xo
    return f


def inv_sqrt_fwav_op(kernels, shape_full, m_wav, inv):
    sq = wav_mat(m_wav, 0.5)

    def f(wav):
        shapes = {j: w.shape for j, w in wav.items()}
        prod = wav_matvec(sq, {j: _flat(w) for j, w in wav.items()})
        return inv(wav2f({j: prod[j].reshape(shapes[j]) for j in prod}, kernels, shape_full))
    return f


# ------------------------------------------------------------- estimation

def cl2flat(c_ell, ells, modlmap):
    c_ell = np.asarray(c_ell, dtype=np.float64)
    ells = np.asarray(ells, dtype=np.float64)
    lo, hi = modlmap.min(), modlmap.max()
    if ells[0] > lo:
        ells = np.concatenate(([lo], ells))
        c_ell = np.concatenate((c_ell[..., :1], c_ell), axis=-1)
    if ells[-1] < hi:
        ells = np.concatenate((ells, [hi]))
        c_ell = np.concatenate((c_ell, c_ell[..., -1:]), axis=-1)
    return np.interp(modlmap, ells, c_ell)


def estimate_cov_pix_flat(imap2d, fwhm, modlmap, diag=False):
    '''estimate_cov_pix(flatsky=True) of (ncomp, npol, ny, nx) maps.'''
    imap = np.asarray(imap2d, dtype=np.float64)
    while imap.ndim < 4:
        imap = imap[np.newaxis]
    ncomp, npol, ny, nx = imap.shape
    flat = imap.reshape(ncomp, npol, ny * nx)
    if diag:
        cov = np.einsum('abc, dbc -> adbc', flat, flat)
    else:
        cov = np.einsum('abc, dec -> abdec', flat, flat).reshape(ncomp * npol, ncomp * npol, -1)
    lmax_g = int(0.5 * (nx - 1))
    b_ell = ops.band_limit_gauss_beam(lmax_g, fwhm=fwhm)
    fb = cl2flat(b_ell, np.arange(b_ell.size), modlmap)
    for idx in np.ndindex(cov.shape[:-1]):
        if idx[0] <= idx[1]:
            if fwhm > 0:
                m = cov[idx].reshape(ny, nx)
                cov[idx] = irfft(rfft(m) * fb, nx).reshape(-1)
        else:
            cov[idx] = cov[(idx[1], idx[0]) + idx[2:]]
    if not diag:
        cov = ops.eigpow(cov, 1, lim=1e-9).reshape(ncomp, npol, ncomp, npol, -1)
    else:
        cov[cov < 0] = 0
    return cov


def estimate_cov_fwav(fmap, kernels, lmaxs, modlmaps, diag=False, fwhm_fact=2):
    '''{j: cov_pix_j} from noise coefficients (npol, ny, nlx); the monopole
    is left out.'''
    f = np.array(fmap, copy=True)
    f[..., 0, 0] = 0
    wav = f2wav(f, kernels)
    out = {}
    for (widx, k, _, _), lmax_j, modl in zip(kernels, lmaxs, modlmaps):
        fwhm = fwhm_fact * np.pi / lmax_j
        cov = estimate_cov_pix_flat(wav[widx], fwhm, modl, diag=diag)
        out[widx] = cov / np.mean(np.abs(np.asarray(k, dtype=np.float64)) ** 2)
    return out


# ----------------------------------------------------------------- solver

def from_arrays_fwav(imap, geom, lmax, ny, nx, icov_ell, m_wav, kernels, b_ell=None,
                     mask_pix=None, cov_noise_2d=None, spin=None, draw_constr=False, seed=None,
                     wav_draw=None):
    '''CGWienerMap.from_arrays_fwav; the noise draw of a constrained
    realisation uses the given unit-variance wavelet maps wav_draw.'''
    if spin is None:
        spin = osht.default_spin(np.shape(imap))
    imap = np.atleast_2d(imap)
    icov_ell = np.atleast_2d(icov_ell)
    power_x = -0.5 if cov_noise_2d is not None else 0
    noise = finv_fwav_map_op(ny, nx, kernels, m_wav, cov_noise_2d, power_x=power_x, nsteps=3)
    beam = ell_op(b_ell, lmax) if b_ell is not None else None
    mask = pix_op(mask_pix) if mask_pix is not None else None
    if draw_constr:
        rand_isignal = ops.rand_alm(icov_ell, lmax, seed)
        fm = inv_sqrt_fwav_op(kernels, (ny, nx // 2 + 1), m_wav, noise.inv)(wav_draw)
        rand_inoise = irfft(noise.x(fm), nx).reshape(imap.shape)
    else:
        rand_isignal = rand_inoise = None
    sht = (y_op(lmax, geom, spin), yt_op(geom, lmax, spin))
    return WienerMap(imap, ell_op(icov_ell, lmax), noise, sht, beam=beam, mask=mask,
                     rand_isignal=rand_isignal, rand_inoise=rand_inoise)


class PseudoInvPreconditionerFWav:
    '''H [S^-1 + tau B P B tau] H with P = Y^T W M F^T X F M W Y.'''

    def __init__(self, lmax, icov_ell, geom, ny, nx, spin, itau_ell, cov_noise_2d, mask_pix,
                 b_ell=None):
        icov_ell = ops.full_matrix(np.atleast_2d(icov_ell))
        npol, _, nell = icov_ell.shape
        if b_ell is None:
            b_ell = np.ones((npol, nell))
        b = np.asarray(b_ell) * np.eye(npol)[:, :, None]
        self.lmax, self.geom, self.ny, self.nx, self.spin = lmax, geom, ny, nx, spin
        self.H = harmonic_matrix(icov_ell, itau_ell, b_ell)
        self.icov_signal = icov_ell
        self.b = b
        self.itau = np.asarray(itau_ell)
        self.xh = x_op(cov_noise_2d, 0.5)
        self.mask = np.asarray(mask_pix, dtype=np.float64)

    def pcov(self, alm):
        m = np.zeros((alm.shape[0], self.geom.npix))
        osht.alm2map(alm, m, self.lmax, self.geom, self.spin, adjoint=True)
        m = m * self.mask
        f = self.xh(self.xh(rfft(m.reshape(m.shape[:-1] + (self.ny, self.nx)))))
        m = irfft(f, self.nx).reshape(m.shape) * self.mask
        out = np.zeros_like(alm)
        osht.map2alm(m, out, self.geom, self.lmax, self.spin, adjoint=True)
        return out

    def __call__(self, alm):
        alm = ops.lmul(alm, self.H, self.lmax)
        sig = ops.lmul(alm, self.icov_signal, self.lmax)
        noi = ops.lmul(ops.lmul(alm, self.b, self.lmax), self.itau, self.lmax)
        noi = self.pcov(noi)
        noi = ops.lmul(ops.lmul(noi, self.itau, self.lmax), self.b, self.lmax)
        return ops.lmul(noi + sig, self.H, self.lmax)
