'''
Wavelet-domain noise model and the harmonic-space Wiener solver (oracle),
restating:
  alm_utils.alm2wlm_axisym 46-103, wlm2alm_axisym 105-155
  operators.WavMatVecAlm 705-815, EllWavEllMatVecAlm 645-703
  mat_utils.wavmatpow 367-411
  map_utils.gauss2gauss 672-717, get_ivar_ell 1220-1266
  preconditioners.get_itau_ell_harm 916-948,
      PseudoInvPreconditionerWav 251-395 (identity scale filter when sfilt is
      None; the reference fails in that case)
  solvers.CGWiener 697-1139 (a_matrix, get_b_vec, get_b_vec_constr,
      get_wiener, get_icov, get_chisq, get_residual, get_qform, from_arrays,
      from_arrays_wav)

A wavelet block matrix is (maps, geoms): dicts keyed by (j, j') holding
(npol, npix) or (npol, npol, npix) arrays and oracle.geometry.Geom grids.

THIS MODULE IS TEST INFRASTRUCTURE (see oracle/__init__.py).
'''
import numpy as np
from scipy.interpolate import RectBivariateSpline

from oracle import layout, ops, sht as osht
from oracle.cg import CG
from oracle.wiener import ell_op, pix_alm_op, harmonic_matrix


def wav_lmax(w_ell, idx):
    '''Last multipole where kernel idx is positive.'''
    w_ell = np.asarray(w_ell)
    return int(w_ell.shape[-1] - 1 - np.argmax(w_ell[idx, ::-1] > 0))


def grid_lmax(geom):
    '''Band limit of a Gauss-Legendre grid from its ring length.'''
    return (int(geom.nphi[0]) - 1) // 2


def alm2wlm(alm, lmax, w_ell, lmaxs=None):
    '''[(w_ell[j, l] a_lm for l, m <= lmax_j, lmax_j)].'''
    w_ell = np.atleast_2d(w_ell)
    if w_ell.shape[1] != lmax + 1:
        raise ValueError('lmax of alm and w_ell do not match')
    out = []
    for j in range(w_ell.shape[0]):
        lj = lmaxs[j] if lmaxs is not None else wav_lmax(w_ell, j)
        wlm = ops.trunc_alm(alm, lmax, lj)
        out.append((ops.lmul(wlm, w_ell[j, :lj + 1], lj), lj))
    return out


def wlm2alm_add(alm, lmax, wlm, lmax_w, w_j):
    '''alm[l, m] += w_j[l] wlm[l, m] for l, m <= lmax_w.'''
    return ops.wlm2alm(np.asarray(w_j)[:lmax_w + 1], wlm, alm, lmax_w, lmax)


def wavmatpow(maps, power):
    return {k: ops.matpow(np.asarray(v), power) for k, v in maps.items()}


def wav_op(lmax, maps, geoms, w_ell, spin, power=1, adjoint=False):
    '''WavMatVecAlm: sum over blocks (j, j') of
    K_j^T Y^T [W] M_jj' [W] Y K_j' on the grid of block (j', j').'''
    w_ell = np.atleast_2d(w_ell)
    if power != 1:
        maps = wavmatpow({k: ops.full_matrix(np.asarray(v)) for k, v in maps.items()}, power)

    def f(alm):
        out = np.zeros_like(alm)
        wl = {}
        for (j, jp), mat in maps.items():
            g = geoms[(jp, jp)]
            if jp not in wl:
                wl[jp] = alm2wlm(alm, lmax, w_ell[jp:jp + 1], lmaxs=[wav_lmax(w_ell, jp)])[0]
            wlm, lw = wl[jp]
            rdt = np.float32 if alm.dtype == np.complex64 else np.float64
            vmap = np.zeros(alm.shape[:-1] + (g.npix,), dtype=rdt)
            osht.alm2map(wlm, vmap, lw, g, spin, adjoint=adjoint)
            mat = np.asarray(mat)
            prod = mat * vmap if mat.ndim == 2 else np.einsum('ijk,jk->ik', mat, vmap)
            lg = grid_lmax(g)
            res = np.zeros(alm.shape[:-1] + (layout.nelem(lg),), dtype=alm.dtype)
            osht.map2alm(prod, res, g, lg, spin, adjoint=not adjoint)
            wlm2alm_add(out, lmax, res, lg, w_ell[j])
        return out
    return f


def ell_wav_ell_op(lmax, maps, geoms, w_ell, x_ell, spin, power_m=1, power_x=1,
                   adjoint=False):
    x = ell_op(x_ell, lmax, power=power_x)
    mw = wav_op(lmax, maps, geoms, w_ell, spin, power=power_m, adjoint=adjoint)
    return lambda alm: x(mw(x(alm)))


def gauss_coords(geom):
    '''(theta, phi) axes of a constant-nphi ring grid.'''
    n = int(geom.nphi[0])
    phi0 = float(geom.phi0[0])
    return np.asarray(geom.theta), np.mod(phi0 + 2 * np.pi * np.arange(n) / n, 2 * np.pi)


def gauss2gauss(imap, g_in, g_out, order=3, area_pow=0):
    '''Spline interpolation between constant-nphi Gauss-Legendre maps.'''
    th_i, ph_i = gauss_coords(g_in)
    th_o, ph_o = gauss_coords(g_out)
    imap = np.asarray(imap)
    im2 = imap.reshape(imap.shape[:-1] + (len(th_i), len(ph_i)))
    if area_pow != 0:
        im2 = im2 * (np.asarray(g_in.weight) ** -float(area_pow))[:, None]
    out = np.zeros(imap.shape[:-1] + (len(th_o), len(ph_o)), dtype=imap.dtype)
    for idx in np.ndindex(im2.shape[:-2]):
        out[idx] = RectBivariateSpline(th_i, ph_i, im2[idx], kx=order, ky=order)(th_o, ph_o)
    if area_pow != 0:
        out *= (np.asarray(g_out.weight) ** float(area_pow))[:, None]
    return out.reshape(imap.shape[:-1] + (g_out.npix,))


def get_ivar_ell(maps, geoms, w_ell, mask=None, geom_mask=None):
    '''sum_j tau_j w_ell[j]^2 with tau_j the isotropic inverse variance of
    block (j, j) (diagonal only).'''
    w_ell = np.atleast_2d(w_ell)
    m0 = np.asarray(maps[(0, 0)])
    npol = m0.shape[0]
    itaus = np.zeros((npol, npol, w_ell.shape[0]))
    for j in range(w_ell.shape[0]):
        g = geoms[(j, j)]
        mask_j = None
        if mask is not None:
            mask_j = gauss2gauss(np.asarray(mask), geom_mask, g, order=1, area_pow=0)
        itaus[:, :, j] = ops.get_isotropic_ivar(np.asarray(maps[(j, j)]), g, mask=mask_j)
    return np.einsum('ijk,kl->ijl', itaus, w_ell ** 2)


def get_itau_ell_harm(icov_ell, w_ell):
    icov_ell = ops.full_matrix(np.atleast_2d(icov_ell))
    itau = np.zeros_like(icov_ell)
    for j in range(np.atleast_2d(w_ell).shape[0]):
        w2 = np.atleast_2d(w_ell)[j] ** 2
        for p in range(icov_ell.shape[0]):
            itau[p, p] += np.sum(icov_ell[p, p] * w2) / np.sum(w2) * w2
    return itau


class WienerAlm:
    '''Harmonic-space Wiener solver (CGWiener).'''

    def __init__(self, alm_data, icov_signal, icov_noise, beam=None, mask=None,
                 rand_isignal=None, rand_inoise=None, swap_bm=False):
        ident = lambda a: a  # noqa: E731
        self.alm_data = alm_data
        self.icov_signal = icov_signal
        self.icov_noise = icov_noise
        self.beam = beam or ident
        self.mask = mask or ident
        self.rand_isignal = rand_isignal
        self.rand_inoise = rand_inoise
        self.swap_bm = swap_bm
        if rand_isignal is not None and rand_inoise is not None:
            self.b_vec = self.get_b_vec_constr(alm_data)
        else:
            self.b_vec = self.get_b_vec(alm_data)
        self.b0 = self.b_vec.copy()
        self.preconditioner = None

    def a_matrix(self, alm):
        if self.swap_bm:
            n = self.beam(self.mask(alm.copy()))
        else:
            n = self.mask(self.beam(alm.copy()))
        n = self.icov_noise(n)
        n = self.mask(self.beam(n)) if self.swap_bm else self.beam(self.mask(n))
        return self.icov_signal(alm.copy()) + n

    def get_b_vec(self, alm):
        alm = self.icov_noise(alm.copy())
        return self.mask(self.beam(alm)) if self.swap_bm else self.beam(self.mask(alm))

    def get_b_vec_constr(self, alm):
        alm = self.get_b_vec(alm)
        if self.swap_bm:
            alm = alm + self.mask(self.beam(self.rand_inoise.copy()))
        else:
            alm = alm + self.beam(self.mask(self.rand_inoise.copy()))
        return alm + self.rand_isignal

    def add_preconditioner(self, prec, sel=None):
        if self.preconditioner is None:
            if sel is None:
                self.preconditioner = prec
            else:
                def sliced(alm):
                    alm[sel] = prec(alm[sel])
                    return alm
                self.preconditioner = sliced
        else:
            from oracle.wiener import add_operators
            self.preconditioner = add_operators(self.preconditioner, prec, slice_2=sel)

    def init_solver(self, **kwargs):
        kwargs.setdefault('dot', ops.contract_almxblm)
        if self.preconditioner is not None:
            kwargs.setdefault('M', self.preconditioner)
        self.cg = CG(self.a_matrix, self.b_vec, **kwargs)
        return self.cg

    def get_residual(self, x=None):
        x = self.cg.x if x is None else x
        r = self.a_matrix(x) - self.b0
        return np.sqrt(ops.contract_almxblm(r, r))

    def get_chisq(self, x):
        out = ops.contract_almxblm(x, self.icov_signal(x))
        d = self.alm_data - x
        return out + ops.contract_almxblm(self.beam(self.mask(d)),
                                          self.icov_noise(self.beam(self.mask(d))))

    def get_qform(self, x):
        return 0.5 * ops.contract_almxblm(x, self.a_matrix(x)) - ops.contract_almxblm(x, self.b0)

    @classmethod
    def from_arrays(cls, alm_data, lmax, icov_ell, icov_pix, geom, spin, b_ell=None,
                    mask_pix=None, swap_bm=False):
        icov_signal = ell_op(icov_ell, lmax)
        icov_noise = pix_alm_op(icov_pix, lmax, geom, spin)
        beam = ell_op(b_ell, lmax) if b_ell is not None else None
        mask = (pix_alm_op(mask_pix, lmax, geom, spin, use_weights=True)
                if mask_pix is not None else None)
        return cls(alm_data, icov_signal, icov_noise, beam=beam, mask=mask, swap_bm=swap_bm)

    @classmethod
    def from_arrays_wav(cls, alm_data, lmax, icov_ell, maps, geoms, w_ell, spin, b_ell=None,
                        mask_pix=None, geom_mask=None, swap_bm=False, icov_noise_ell=None):
        icov_signal = ell_op(icov_ell, lmax)
        if icov_noise_ell is None:
            icov_noise = wav_op(lmax, maps, geoms, w_ell, spin)
        else:
            icov_noise = ell_wav_ell_op(lmax, maps, geoms, w_ell, icov_noise_ell, spin,
                                        power_x=0.5)
        beam = ell_op(b_ell, lmax) if b_ell is not None else None
        mask = (pix_alm_op(mask_pix, lmax, geom_mask, spin, use_weights=True)
                if mask_pix is not None else None)
        return cls(alm_data, icov_signal, icov_noise, beam=beam, mask=mask, swap_bm=swap_bm)


class PseudoInvPreconditionerWav:
    '''M = H [S^-1 + F B M^-1 tau N_wav^-1' tau M^-1 B F] H (F identity when
    sfilt is None).'''

    def __init__(self, lmax, icov_ell, maps, geoms, w_ell, spin, mask_pix=None, geom_mask=None,
                 itau_ell=None, b_ell=None, icov_noise_ell=None, sfilt=None):
        icov_ell = ops.full_matrix(np.atleast_2d(icov_ell))
        npol, _, nell = icov_ell.shape
        if itau_ell is None:
            if icov_noise_ell is None:
                itau_ell = get_ivar_ell(maps, geoms, w_ell, mask=mask_pix, geom_mask=geom_mask)
            else:
                itau_ell = get_itau_ell_harm(icov_noise_ell, w_ell)
        if b_ell is None:
            b_ell = np.ones((npol, nell))
        b = np.asarray(b_ell) * np.eye(npol)[:, :, None]
        self.lmax = lmax
        ident = lambda a: a  # noqa: E731
        self.imask = (pix_alm_op(mask_pix, lmax, geom_mask, spin, power=-1, use_weights=True)
                      if mask_pix is not None else ident)
        if sfilt is not None:
            sf = ops.full_matrix(np.asarray(sfilt))
            op = np.einsum('ijl,jkl,kol->iol', sf, icov_ell, sf)
            op2 = np.einsum('ijl,jkl,kol->iol', b, itau_ell, b)
            self.H = ops.matpow(op + np.einsum('ijl,jkl,kol->iol', sf, op2, sf), -1)
            self.icov_signal = np.einsum('ijl,jkl,kol->iol', sf, icov_ell, sf)
            self.sfilt = lambda a: ops.lmul(a, sf, lmax)
        else:
            self.H = harmonic_matrix(icov_ell, itau_ell, b_ell)
            self.icov_signal = icov_ell
            self.sfilt = ident
        self.b = b
        self.itau = itau_ell
        if icov_noise_ell is None:
            self.pcov = wav_op(lmax, maps, geoms, w_ell, spin, power=-1, adjoint=True)
        else:
            self.pcov = ell_wav_ell_op(lmax, maps, geoms, w_ell, icov_noise_ell, spin,
                                       power_m=-1, power_x=-0.5, adjoint=True)

    def __call__(self, alm):
        L = self.lmax
        alm = ops.lmul(alm, self.H, L)
        sig = ops.lmul(alm, self.icov_signal, L)
        n = self.sfilt(alm)
        n = ops.lmul(n, self.b, L)
        n = self.imask(n)
        n = ops.lmul(n, self.itau, L)
        n = self.pcov(n)
        n = ops.lmul(n, self.itau, L)
        n = self.imask(n)
        n = ops.lmul(n, self.b, L)
        n = self.sfilt(n)
        return ops.lmul(n + sig, self.H, L)
