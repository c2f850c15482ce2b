'''
Wiener-filter solver, preconditioners and multigrid (oracle), restating:
  solvers.CGWienerMap (__init__, init_solver, add_preconditioner, proj,
      proj_adjoint, a_matrix, get_b_vec, get_b_vec_constr, set_b_vec,
      get_wiener, get_icov, get_chisq, get_residual, get_qform, from_arrays)
      solvers.py:8-456
  operators.EllMatVecAlm 206-257, PixMatVecMap 162-204, YMatVecAlm 49-103,
      YTWMatVecMap 105-160, PixMatVecAlm 562-643, PixEllPixMatVecMap 991-1102,
      add_operators 1458-1489
  preconditioners.HarmonicPreconditioner 8-124 (itau from icov_pix),
      PseudoInvPreconditioner 126-249, MaskedPreconditioner 558-650,
      MaskedPreconditionerCG 652-789, get_2level_prec / get_prec_adef1 /
      get_prec_adef2 791-914
  operators.InvSqrtPixEllPixMatVecMap 1104-1200, InvPixEllPixMatVecMap
      1202-1284; solvers.CGWienerMap.from_arrays_const_cor 459-575
  filters.CGPixFilter 9-301
  multigrid.Level 13-156, get_levels 158-249, get_equal_area_mask_bool
      251-294, restrict/coarsen/interpolate 296-398, v_cycle 400-447
Only the operator variants that the Wiener path uses are restated (no
lensing, no 2D-Fourier filters).
'''
import numpy as np
from scipy import linalg

from oracle import layout, geometry, ops, sht as osht
from oracle.cg import CG


# ----------------------------------------------------------------- operators

def ell_op(lmat, lmax, power=1):
    '''EllMatVecAlm: returns a new array (inplace=False semantics).'''
    m = ops.full_matrix(np.asarray(lmat))
    if power != 1:
        m = ops.matpow(m, power)
    return lambda alm: ops.lmul(alm, m, lmax)


def pix_op(m_pix):
    '''PixMatVecMap(m_pix, power=1).'''
    return lambda imap: ops.matvec(m_pix, imap)


def y_op(lmax, geom, spin, qweight=False):
    '''YMatVecAlm: Y (qweight=False) or W Y.'''
    def f(alm):
        omap = np.zeros(alm.shape[:-1] + (geom.npix,),
                        dtype=np.float32 if alm.dtype == np.complex64 else np.float64)
        osht.alm2map(alm, omap, lmax, geom, spin, adjoint=qweight)
        return omap
    return f


def yt_op(geom, lmax, spin, qweight=False):
    '''YTWMatVecMap: Y^T W (qweight=True) or Y^T.'''
    def f(imap):
        alm = np.zeros(imap.shape[:-1] + (layout.nelem(lmax),),
                       dtype=np.complex64 if imap.dtype == np.float32 else np.complex128)
        osht.map2alm(imap, alm, geom, lmax, spin, adjoint=not qweight)
        return alm
    return f


def pix_alm_op(m_pix, lmax, geom, spin, power=1, adjoint=False, use_weights=False):
    '''PixMatVecAlm: Y^T [W] M [W] Y with the weight placement of
    operators.py:627-641.'''
    m = ops.full_matrix(np.asarray(m_pix))
    if power != 1:
        m = ops.matpow(m, power)

    def f(alm):
        npol = alm.shape[0]
        omap = np.zeros((npol, m.shape[-1]),
                        dtype=np.float32 if alm.dtype == np.complex64 else np.float64)
        osht.alm2map(alm, omap, lmax, geom, spin, adjoint=adjoint)
        imap = np.einsum('ijk,jk->ik', m, omap)
        out = np.zeros_like(alm)
        adj = adjoint if use_weights else not adjoint
        osht.map2alm(imap, out, geom, lmax, spin, adjoint=adj)
        return out
    return f


def pix_ell_pix_op(m_pix, x_ell, geom, spin, lmax=None, power_x=1, adjoint=False,
                   use_weights=False, power_m=1):
    '''PixEllPixMatVecMap (operators.py:991-1102): M Y x_ell^p Y^T M, with
    the quadrature weights W on the analysis when adjoint is use_weights
    (Y^T W) and on the synthesis when adjoint (W Y); no weights for the
    defaults.'''
    if lmax is None:
        lmax = geometry.minfo2lmax(geom)
    if power_x != 1:
        x_ell = ops.matpow(x_ell, power_x, return_diag=True)
    if m_pix is not None and power_m != 1:
        m_pix = ops.matpow(m_pix, power_m, return_diag=True)
    if m_pix is not None and np.asarray(m_pix).ndim == 3:
        mp = np.asarray(m_pix)
        mul = lambda v: np.einsum('abp,bp->ap', mp, v)
    else:
        mul = lambda v: v * m_pix

    def f(imap):
        out = np.array(imap, copy=True)
        if m_pix is not None:
            out = mul(out)
        alm = np.zeros((out.shape[0], layout.nelem(lmax)),
                       dtype=np.complex64 if out.dtype == np.float32 else np.complex128)
        osht.map2alm(out, alm, geom, lmax, spin, adjoint=adjoint is use_weights)
        alm = ops.lmul(alm, x_ell, lmax)
        osht.alm2map(alm, out, lmax, geom, spin, adjoint=adjoint)
        if m_pix is not None:
            out = mul(out)
        return out
    f.m_pix = m_pix
    return f


def inv_pix_ell_pix_op(m_pix, x_ell, geom, spin, power_m=1, power_x=1, mask=None,
                       adjoint=False, lmax=None, nsteps=5):
    '''InvPixEllPixMatVecMap (operators.py:1202-1284): nsteps of CG on
    M' M^p Y X^q Y^T M^p M' from x = 0, preconditioned by the powers -p, -q
    with the weights on the other side; dot = sum(x y).'''
    a_mat = pix_ell_pix_op(m_pix, x_ell, geom, spin, lmax=lmax, power_x=power_x,
                           adjoint=adjoint, power_m=power_m)
    prec = pix_ell_pix_op(m_pix, x_ell, geom, spin, lmax=lmax, power_x=-power_x,
                          adjoint=not adjoint, power_m=-power_m)
    if mask is not None:
        am = lambda v: mask * a_mat(mask * v)
        pm = lambda v: mask * prec(mask * v)
    else:
        am, pm = a_mat, prec

    def f(imap):
        cg = CG(am, np.array(imap, copy=True), M=pm, dot=lambda x, y: np.sum(x * y))
        for _ in range(nsteps):
            cg.step()
        return cg.x
    f.a_mat = a_mat
    return f


def inv_sqrt_pix_ell_pix_op(m_pix, x_ell, geom, spin, power_m=1, power_x=1, mask=None,
                            adjoint=False, lmax=None, inv_op=None, nsteps=5):
    '''InvSqrtPixEllPixMatVecMap (operators.py:1104-1200): inv_op applied to
    M' M^p Y X^(q/2) alm.'''
    if lmax is None:
        lmax = geometry.minfo2lmax(geom)
    if inv_op is None:
        inv_op = inv_pix_ell_pix_op(m_pix, x_ell, geom, spin, power_m=power_m,
                                    power_x=power_x, mask=mask, adjoint=adjoint, lmax=lmax,
                                    nsteps=nsteps)
    xh = ops.matpow(x_ell, power_x / 2, return_diag=True)
    mp = inv_op.a_mat.m_pix

    def f(alm):
        a = ops.lmul(alm, xh, lmax)
        out = np.zeros((a.shape[0], geom.npix))
        osht.alm2map(a, out, lmax, geom, spin, adjoint=adjoint)
        if mp is not None:
            out = np.einsum('abp,bp->ap', mp, out) if np.asarray(mp).ndim == 3 else out * mp
        if mask is not None:
            out = out * mask
        return inv_op(out)
    return f


def add_operators(op1, op2, slice_2=None):
    if slice_2 is None:
        slice_2 = slice(None)

    def add(v):
        o = op1(v.copy())
        o[slice_2] += op2(v[slice_2].copy())
        return o
    return add


# ------------------------------------------------------------------- solver

class WienerMap:
    '''CGWienerMap restated (swap_bm=False and swap_bm=True paths; filt is
    the (F, F^T) pair of map operators, solvers.py:105-111, 177-215).'''

    def __init__(self, imap, icov_signal, icov_noise, sht, beam=None,
                 mask=None, rand_isignal=None, rand_inoise=None,
                 swap_bm=False, scale_filter=None, filt=None):
        self.imap = imap
        self.filt, self.filt_adjoint = filt if filt is not None else ((lambda m: m),
                                                                      (lambda m: m))
        self.icov_signal = icov_signal
        self.icov_noise = icov_noise
        self.beam = beam if beam is not None else (lambda a: a)
        self.sht, self.sht_adjoint = sht
        self.mask = mask if mask is not None else (lambda m: m)
        self.swap_bm = swap_bm
        self.scale_filter = scale_filter if scale_filter is not None else (lambda a: a)
        self.set_b_vec(imap, rand_isignal, rand_inoise)
        self.preconditioner = None

    def proj(self, alm):
        if not self.swap_bm:
            return self.filt(self.mask(self.sht(self.beam(alm))))
        return self.filt(self.sht(self.beam(self.mask(alm))))

    def proj_adjoint(self, imap):
        imap = self.filt_adjoint(imap)
        if not self.swap_bm:
            return self.beam(self.sht_adjoint(self.mask(imap)))
        return self.mask(self.beam(self.sht_adjoint(imap)))

    def a_matrix(self, alm):
        alm = self.scale_filter(alm)
        omap = self.icov_noise(self.proj(alm))
        oalm = self.scale_filter(self.proj_adjoint(omap))
        oalm += self.scale_filter(self.icov_signal(alm))
        return oalm

    def get_b_vec(self, imap):
        return self.scale_filter(self.proj_adjoint(self.icov_noise(imap)))

    def get_b_vec_constr(self, rand_isignal, rand_inoise):
        o = self.scale_filter(self.proj_adjoint(rand_inoise))
        o += self.scale_filter(rand_isignal)
        return o

    def set_b_vec(self, imap, rand_isignal=None, rand_inoise=None):
        self.b_vec = self.get_b_vec(imap)
        if rand_isignal is not None and rand_inoise is not None:
            self.b_vec += self.get_b_vec_constr(rand_isignal, rand_inoise)
        self.b0 = self.b_vec.copy()

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
            self.preconditioner = add_operators(self.preconditioner, prec, slice_2=sel)

    def reset_preconditioner(self):
        self.preconditioner = None

    def init_solver(self, **kwargs):
        kwargs.setdefault('dot', ops.contract_almxblm)
        if self.preconditioner is not None:
            kwargs.setdefault('M', self.preconditioner)
        self.cg = CG(self.a_matrix, self.b_vec, **kwargs)
        return self.cg

    def get_wiener(self):
        return self.scale_filter(self.cg.x.copy())

    def get_icov(self):
        return self.icov_signal(self.get_wiener())

    def get_chisq(self):
        x_w = self.get_wiener()
        out = ops.contract_almxblm(x_w, self.icov_signal(x_w))
        res = self.imap - self.proj(x_w)
        out += np.sum(res * self.icov_noise(res))
        return out

    def get_residual(self):
        r = self.a_matrix(self.cg.x) - self.b0
        return np.sqrt(ops.contract_almxblm(r, r))

    def get_qform(self):
        return (0.5 * ops.contract_almxblm(self.cg.x, self.a_matrix(self.cg.x))
                - ops.contract_almxblm(self.cg.x, self.b0))

    @classmethod
    def from_arrays(cls, imap, geom, lmax, icov_ell, icov_pix, b_ell=None,
                    mask_pix=None, draw_constr=False, spin=None, swap_bm=False,
                    sfilt=None, seed=None, filt=None):
        if spin is None:
            spin = osht.default_spin(imap.shape)
        imap = np.atleast_2d(imap)
        icov_pix = np.atleast_2d(icov_pix)
        icov_ell = np.atleast_2d(icov_ell)
        icov_signal = ell_op(icov_ell, lmax)
        icov_noise = pix_op(icov_pix)
        beam = ell_op(b_ell, lmax) if b_ell is not None else None
        sf = ell_op(sfilt, lmax) if sfilt is not None else None
        if mask_pix is not None:
            if swap_bm:
                mask = pix_alm_op(mask_pix, lmax, geom, spin, use_weights=True)
            else:
                mask = pix_op(mask_pix)
        else:
            mask = None
        if draw_constr:
            rand_isignal = ops.rand_alm(icov_ell, lmax, seed)
            rand_inoise = ops.rand_map_pix(icov_pix, seed)
        else:
            rand_isignal = rand_inoise = None
        sht = (y_op(lmax, geom, spin), yt_op(geom, lmax, spin))
        return cls(imap, icov_signal, icov_noise, sht, beam=beam, mask=mask,
                   rand_isignal=rand_isignal, rand_inoise=rand_inoise,
                   swap_bm=swap_bm, scale_filter=sf, filt=filt)


def from_arrays_const_cor(imap, geom, lmax, icov_ell, icov_pix, icov_noise_ell, nsteps=5,
                          b_ell=None, mask_pix=None, no_masked_noise=False, draw_constr=False,
                          spin=None, seed=None):
    '''CGWienerMap.from_arrays_const_cor (solvers.py:459-575).'''
    if spin is None:
        spin = osht.default_spin(np.shape(imap))
    imap = np.atleast_2d(imap)
    icov_pix = np.atleast_2d(icov_pix)
    icov_ell = np.atleast_2d(icov_ell)
    icov_noise_ell = np.atleast_2d(icov_noise_ell)
    nmask = None if no_masked_noise else mask_pix
    icov_noise = inv_pix_ell_pix_op(icov_pix, icov_noise_ell, geom, spin, power_m=-0.5,
                                    power_x=-1, mask=nmask, nsteps=nsteps)
    beam = ell_op(b_ell, lmax) if b_ell is not None else None
    mask = pix_op(mask_pix) if mask_pix is not None else None
    if draw_constr:
        rand_isignal = ops.rand_alm(icov_ell, lmax, seed)
        rand_inoise = inv_sqrt_pix_ell_pix_op(icov_pix, icov_noise_ell, geom, spin,
                                              power_m=-0.5, power_x=-1, mask=nmask,
                                              inv_op=icov_noise)(
            ops.unit_var_alm(lmax, imap.shape[0], seed))
    else:
        rand_isignal = rand_inoise = None
    sht = (y_op(lmax, geom, spin), yt_op(geom, lmax, spin))
    return WienerMap(imap, ell_op(icov_ell, lmax), icov_noise, sht, beam=beam, mask=mask,
                     rand_isignal=rand_isignal, rand_inoise=rand_inoise)


# ------------------------------------------------------------ preconditioners

def harmonic_matrix(icov_ell, itau, b_ell=None):
    '''(S^-1 + B itau B)^-1 per l (HarmonicPreconditioner, no sfilt), as a
    full (npol, npol, nell) matrix via eigpow(lim=1e-9).'''
    icov_ell = ops.full_matrix(np.asarray(icov_ell))
    npol, _, nell = icov_ell.shape
    itau = np.asarray(itau)
    if itau.ndim == 1:
        itau = itau * np.eye(npol)
    if itau.ndim == 2:
        itau = itau[:, :, None]
    if b_ell is None:
        b_ell = np.ones((npol, nell))
    b = np.asarray(b_ell) * np.eye(npol)[:, :, None]
    op = icov_ell + np.einsum('ijl,jkl,kol->iol', b, itau, b)
    return ops.matpow(op, -1)


class HarmonicPreconditioner:

    def __init__(self, lmax, icov_ell, itau=None, icov_pix=None, geom=None,
                 b_ell=None):
        if itau is None:
            itau = ops.get_isotropic_ivar(np.atleast_2d(icov_pix), geom)
        self.lmax = lmax
        self.mat = harmonic_matrix(icov_ell, itau, b_ell)

    def __call__(self, alm):
        return ops.lmul(alm, self.mat, self.lmax)


class PseudoInvPreconditioner:
    '''M = H [S^-1 + (tau B) Y^T W N W Y (tau B)] H, H = (S^-1 + B tau B)^-1.'''

    def __init__(self, lmax, icov_ell, icov_pix, geom, spin, itau=None,
                 b_ell=None, cov_pix=None):
        icov_ell = ops.full_matrix(np.atleast_2d(icov_ell))
        npol, _, nell = icov_ell.shape
        if itau is None:
            itau = ops.get_isotropic_ivar(np.atleast_2d(icov_pix), geom)
        if itau.ndim == 2:
            itau = itau[:, :, None]
        if b_ell is None:
            b_ell = np.ones((npol, nell))
        b = np.asarray(b_ell) * np.eye(npol)[:, :, None]
        if cov_pix is None:
            cov_pix = ops.matpow(np.atleast_2d(icov_pix), -1, return_diag=True)
        self.lmax = lmax
        self.H = harmonic_matrix(icov_ell, itau, b_ell)
        self.icov_signal = icov_ell
        self.tau_b = itau * b
        self.pcov = pix_alm_op(cov_pix, lmax, geom, spin, adjoint=True)

    def __call__(self, alm):
        alm = ops.lmul(alm, self.H, self.lmax)
        sig = ops.lmul(alm, self.icov_signal, self.lmax)
        noi = ops.lmul(alm, self.tau_b, self.lmax)
        noi = self.pcov(noi)
        noi = ops.lmul(noi, self.tau_b, self.lmax)
        return ops.lmul(noi + sig, self.H, self.lmax)


# --------------------------------------------------------------- multigrid

class Level:
    '''multigrid.Level: G = M_u Y d_ell Y^T M_u on an equal-area grid.
    The sparse smoother is 0.2 / diag(G) with diag(G) evaluated at one
    unobserved pixel per component in float64 and rounded to float32 (the
    reference probes with a float32 transform, multigrid.py:82-113).'''

    def __init__(self, mask_obs, geom, d_ell, spin, dense=False):
        if mask_obs.ndim == 1:
            mask_obs = mask_obs[None]
        self.mask_unobs = ~mask_obs
        self.geom = geom
        self.spin = spin
        self.d_ell = d_ell
        self.lmax = geometry.minfo2lmax(geom)
        self.npol = self.mask_unobs.shape[0]
        self.g_op = pix_ell_pix_op(self.mask_unobs, d_ell, geom, spin)
        self.pinv_g = None
        if dense:
            self.pinv_g = linalg.pinvh(self.dense_g())
            self.smoother = self._dense_smoother
        else:
            self.m_prec = self.probe_diag()
            self.smoother = lambda x: self.m_prec[:, None] * x

    def probe_diag(self):
        m_prec = np.zeros(self.npol)
        idx = np.argmax(self.mask_unobs, axis=-1)
        test = np.zeros((self.npol, self.geom.npix))
        for p in range(self.npol):
            test[p, idx[p]] = 1
            o = self.g_op(test)
            m_prec[p] = 1 / np.float32(o[p, idx[p]])
            test[p, idx[p]] = 0
        return 0.2 * m_prec

    def dense_g(self):
        nun = int(np.sum(self.mask_unobs))
        mat = np.zeros((nun, nun))
        u = np.zeros(nun)
        for i in range(nun):
            u[i] = 1
            full = np.zeros(self.mask_unobs.shape)
            full[self.mask_unobs] = u
            mat[:, i] = self.g_op(full)[self.mask_unobs]
            u[i] = 0
        return mat

    def _dense_smoother(self, imap):
        o = np.zeros(imap.shape, imap.dtype)
        o[self.mask_unobs] = np.dot(self.pinv_g, imap[self.mask_unobs])
        return o


def equal_area_mask(mask_bool, geom, lmax, geom_fullsky=None):
    if geom_fullsky is None:
        geom_fullsky = geometry.gauss_legendre(lmax + 1, 2 * lmax + 1)
    geom_out = geometry.equal_area(geom_fullsky, orig_band=0.4, ratio_pow=1)
    m = geometry.gauss2map(mask_bool.astype(np.float32), geom, geom_out, order=1)
    m[(m > 0) & (m < 1)] = 1.
    return m.astype(bool), geom_out


def get_levels(mask, geom, icov_ell, spin, min_pix=1000, lmax_r_ell=6000):
    icov_ell = np.asarray(icov_ell)
    if icov_ell.ndim == 1:
        icov_ell = icov_ell[None]
    if icov_ell.ndim == 2:
        npol, nell = icov_ell.shape
        icov_ell = icov_ell * np.eye(npol)[:, :, None]
    elif icov_ell.ndim == 3:
        npol, _, nell = icov_ell.shape
    else:
        raise ValueError('icov_ell shape not supported')
    if mask.ndim == 1:
        mask = np.ones((npol, mask.size), dtype=mask.dtype) * mask
    if mask.dtype != bool:
        raise ValueError('mask must be bool')
    lmax = geometry.minfo2lmax(geom)
    levels = []
    for idx in range(16):
        lmax_level = lmax // 2 ** idx
        mask_level, geom_level = equal_area_mask(
            mask, geom, lmax_level,
            geom_fullsky=geometry.fullsky_geometry(geom) if idx == 0 else None)
        nmasked = np.sum(~mask_level, axis=1)
        last = bool(np.any(nmasked < min_pix))
        if np.any(nmasked == 0):
            raise ValueError('no masked pixels left')
        if idx == 0:
            r_ell = ops.lowpass_filter(lmax_r_ell, lmax=lmax_level)
            d_ell = icov_ell[:, :, :lmax_level + 1] * r_ell ** 2
        else:
            r_ell = ops.band_limit_gauss_beam(lmax_level, fwhm=2 * np.pi / lmax_level)
            d_ell = levels[idx - 1].d_ell[:, :, :lmax_level + 1] * r_ell ** 2
        levels.append(Level(mask_level, geom_level, d_ell, spin, dense=last))
        if last:
            break
    return levels


def restrict(imap, lin, lout, spin, adjoint=False):
    lmax_in, lmax_out = lin.lmax, lout.lmax
    imap = imap * lin.mask_unobs
    lmax_t = lmax_out if lmax_out < lmax_in else lmax_in
    alm = np.zeros((imap.shape[0], layout.nelem(lmax_t)), dtype=np.complex128)
    osht.map2alm(imap, alm, lin.geom, lmax_t, spin, adjoint=adjoint)
    if not adjoint:
        r_ell = ops.band_limit_gauss_beam(lmax_out, fwhm=2 * np.pi / lmax_out)
    else:
        r_ell = ops.band_limit_gauss_beam(lmax_in, fwhm=2 * np.pi / lmax_in)
    alm = ops.lmul(alm, r_ell, lmax_t)
    omap = np.zeros((imap.shape[0], lout.geom.npix))
    osht.alm2map(alm, omap, lmax_t, lout.geom, spin, adjoint=adjoint)
    return omap * lout.mask_unobs


def v_cycle(levels, imap, spin, idx=0, n_jacobi=1, x0=None):
    if idx == len(levels) - 1:
        return levels[idx].smoother(imap)
    lev = levels[idx]
    if x0 is not None:
        x = x0.copy()
        x += lev.smoother(imap - lev.g_op(x))
    else:
        x = np.zeros_like(imap)
        x += lev.smoother(imap)
    for _ in range(n_jacobi - 1):
        x += lev.smoother(imap - lev.g_op(x))
    r_h = imap - lev.g_op(x)
    r_H = restrict(r_h, lev, levels[idx + 1], spin, adjoint=False)
    c_H = v_cycle(levels, r_H, spin, idx=idx + 1, n_jacobi=n_jacobi)
    x += restrict(c_H, levels[idx + 1], lev, spin, adjoint=True)
    for _ in range(n_jacobi):
        x += lev.smoother(imap - lev.g_op(x))
    return x


class MaskedPreconditioner:

    def __init__(self, lmax, icov_ell, spin, mask_bool, geom, min_pix=1000,
                 n_jacobi=1, lmax_r_ell=6000, nsteps=1, levels=None):
        icov_ell = np.asarray(icov_ell)
        if icov_ell.ndim == 1:
            icov_ell = icov_ell[None]
        self.npol = icov_ell.shape[0]
        self.spin = spin
        self.lmax = lmax
        self.n_jacobi = n_jacobi
        self.nsteps = nsteps
        if levels is None:
            levels = get_levels(mask_bool, geom, icov_ell, spin,
                                min_pix=min_pix, lmax_r_ell=lmax_r_ell)
        self.levels = levels
        self.r_ell = ops.lowpass_filter(lmax_r_ell, lmax=lmax)
        self.mask_unobs = levels[0].mask_unobs
        self.geom = levels[0].geom

    def __call__(self, alm):
        alm = ops.lmul(alm, self.r_ell, self.lmax)
        imap = np.zeros((self.npol, self.geom.npix))
        osht.alm2map(alm, imap, self.lmax, self.geom, self.spin)
        imap *= self.mask_unobs
        omap = None
        for _ in range(self.nsteps):
            omap = v_cycle(self.levels, imap, self.spin, n_jacobi=self.n_jacobi, x0=omap)
        omap *= self.mask_unobs
        out = np.zeros_like(alm)
        osht.map2alm(omap, out, self.geom, self.lmax, self.spin, adjoint=True)
        return ops.lmul(out, self.r_ell, self.lmax)


class MaskedPreconditionerCG:
    '''MaskedPreconditionerCG (preconditioners.py:652-789): nsteps of CG on
    the masked-pixel system G x = M_u Y r_ell alm_lowres with
    G = M_u Y d_ell Y^T M_u (d_ell = icov_ell r_ell^2, or icov_ell without
    r_ell) and inner preconditioner M_u W Y icov_ell^-1 Y^T W M_u, on the
    equal-area geometry at the preconditioner's lmax; the result
    r_ell Y^T M_u x is placed in the l, m <= lmax part of a zero alm.'''

    def __init__(self, lmax_in, icov_ell, spin, mask_bool, geom, lmax_r_ell=None, lmax=None,
                 nsteps=15, sfilt=None):
        icov_ell = np.asarray(icov_ell)
        self.spin = spin
        self.npol = icov_ell.shape[0]
        self.lmax_in = lmax_in
        self.nsteps = nsteps
        if sfilt is not None:
            sf = ops.full_matrix(np.asarray(sfilt))
            ic = ops.full_matrix(icov_ell)
            icov_ell = np.einsum('ijl, jkl, kol -> iol', sf, ic, sf)
        if lmax is not None:
            lmax = min(lmax, lmax_in)
            icov_ell = icov_ell[..., :lmax + 1].copy()
        else:
            lmax = lmax_in
        self.lmax = lmax
        if lmax_r_ell is not None:
            self.r_ell = ops.lowpass_filter(lmax_r_ell, lmax=lmax_in)
            d_ell = icov_ell * self.r_ell[:lmax + 1] ** 2
        else:
            self.r_ell = None
            d_ell = icov_ell.copy()
        mask_bool = np.asarray(mask_bool)
        if mask_bool.ndim == 1:
            mask_bool = np.ones((self.npol, mask_bool.size), dtype=mask_bool.dtype) * mask_bool
        elif mask_bool.shape[0] != self.npol:
            raise ValueError(f'shape[0] mask != npol {mask_bool.shape[0]} != {self.npol}')
        if mask_bool.dtype != bool:
            raise ValueError(f'Input mask must be bool, got {mask_bool.dtype}')
        mask_bool, geom = equal_area_mask(mask_bool, geom, lmax)
        self.mask_unobs = ~mask_bool
        self.geom = geom
        self.g_op = pix_ell_pix_op(self.mask_unobs, d_ell, geom, spin, lmax=lmax)
        self.prec = pix_ell_pix_op(self.mask_unobs, icov_ell, geom, spin, lmax=lmax,
                                   power_x=-1, adjoint=True)
        self.dot = lambda x, y: np.dot(x.reshape(-1), y.reshape(-1))

    def __call__(self, alm):
        imap = np.zeros((self.npol, self.geom.npix),
                        dtype=np.float32 if alm.dtype == np.complex64 else np.float64)
        alm_lo = ops.trunc_alm(alm, self.lmax_in, self.lmax)
        if self.r_ell is not None:
            alm_lo = ops.lmul(alm_lo, self.r_ell[:self.lmax + 1], self.lmax)
        osht.alm2map(alm_lo, imap, self.lmax, self.geom, self.spin)
        imap *= self.mask_unobs
        cg = CG(self.g_op, imap, dot=self.dot, M=self.prec)
        for _ in range(self.nsteps):
            cg.step()
        omap = cg.x * self.mask_unobs
        osht.map2alm(omap, alm_lo, self.geom, self.lmax, self.spin, adjoint=True)
        if self.r_ell is not None:
            alm_lo = ops.lmul(alm_lo, self.r_ell[:self.lmax + 1], self.lmax)
        out = np.zeros_like(alm)
        return ops.add_to_alm(out, alm_lo, self.lmax_in, self.lmax)


def prec_adef1(prec_main, prec_masked, a_op, sel_masked=None):
    '''ADEF-1 (preconditioners.py:824-865): M2 b = M (b - A Q b) + Q b.'''
    def f(alm):
        if sel_masked is not None:
            q = np.zeros_like(alm)
            q[sel_masked] = prec_masked(alm[sel_masked].copy())
        else:
            q = prec_masked(alm)
        out = prec_main(alm - a_op(q))
        return out + q
    return f


def prec_adef2(prec_main, prec_masked, a_op, sel_masked=None):
    '''ADEF-2 (preconditioners.py:867-914): M2 b = M b - Q A M b + Q b.'''
    def f(alm):
        m = prec_main(alm)
        qam = a_op(m)
        if sel_masked is not None:
            tmp = qam[sel_masked].copy()
            qam = np.zeros_like(qam)
            qam[sel_masked] = prec_masked(tmp)
        else:
            qam = prec_masked(qam)
        m = m - qam
        if sel_masked is not None:
            q = np.zeros_like(alm)
            q[sel_masked] = prec_masked(alm[sel_masked].copy())
        else:
            q = prec_masked(alm)
        return m + q
    return f


def get_2level_prec(prec_main, prec_masked, a_op, method, sel_masked=None):
    '''get_2level_prec (preconditioners.py:791-822).'''
    if method == 'ADEF-1':
        return prec_adef1(prec_main, prec_masked, a_op, sel_masked=sel_masked)
    if method == 'ADEF-2':
        return prec_adef2(prec_main, prec_masked, a_op, sel_masked=sel_masked)
    raise ValueError(f'method : {method} not supported.')


class CGPixFilter:
    '''filters.CGPixFilter (filters.py:9-301) on a CAR map given by its WCS:
    TT / TE / EE / BB signal, beam, pixel inverse covariance and mask;
    phase 1 with the pseudo-inverse plus MaskedPreconditionerCG for
    niter_masked_cg steps, then the pseudo-inverse plus the T multigrid
    restarted from the phase-1 solution. The pixels are gathered once into
    the unit-stride ring order of the map's geometry. scale_a is not
    restated.'''

    def __init__(self, theory_cls, b_ell, icov_pix, mask_bool, wcs, include_te=True, q_low=0,
                 q_high=1, swap_bm=False, lmax=None, lmax_prec_cg=None,
                 no_masked_prec_mg=False, no_masked_prec_cg=False):
        b_ell = np.asarray(b_ell)
        icov_pix = np.asarray(icov_pix)
        if not np.all(np.isfinite(b_ell)) or not np.all(np.isfinite(icov_pix)):
            raise ValueError('non-finite input')
        self.shape_in = icov_pix.shape[-2:]
        icov_pix = icov_pix.reshape((-1,) * (3 - icov_pix.ndim) + icov_pix.shape) \
            if icov_pix.ndim < 3 else icov_pix
        mask_bool = np.asarray(mask_bool).astype(bool)
        mask_bool = mask_bool.reshape((1,) * (3 - mask_bool.ndim) + mask_bool.shape)
        ncomp = icov_pix.shape[0]
        if mask_bool.shape[0] == 1:
            mask_bool = (np.ones(ncomp)[:, None, None] * mask_bool).astype(bool)
        for mtype in ('CC', 'fejer1'):
            try:
                _, geom, index = geometry.match_car(icov_pix.shape, wcs, mtype=mtype)
            except ValueError:
                continue
            break
        self.geom, self.index = geom, index
        if lmax is None:
            lmax = geometry.minfo2lmax(geom)
        icov_pix = icov_pix.reshape(icov_pix.shape[:-2] + (-1,))[..., index]
        mask_bool = mask_bool.reshape(mask_bool.shape[:-2] + (-1,))[..., index]
        if q_low != 0 or q_high != 1:
            icov_pix = ops.threshold_icov(icov_pix, q_low=q_low, q_high=q_high)
        if not theory_cls['TT'].size - 1 >= lmax:
            raise ValueError('theory_cls shorter than lmax')
        cov_ell = np.zeros((ncomp, ncomp, lmax + 1))
        cov_ell[0, 0] = theory_cls['TT'][:lmax + 1]
        if ncomp > 1:
            if include_te:
                cov_ell[0, 1] = theory_cls['TE'][:lmax + 1]
                cov_ell[1, 0] = theory_cls['TE'][:lmax + 1]
            cov_ell[1, 1] = theory_cls['EE'][:lmax + 1]
            cov_ell[2, 2] = theory_cls['BB'][:lmax + 1]
        icov_ell = np.zeros_like(cov_ell)
        for l in range(icov_ell.shape[-1]):
            icov_ell[:, :, l] = np.linalg.pinv(cov_ell[:, :, l])
        if b_ell.ndim == 1:
            b_ell = b_ell[None] * np.asarray((1, 1, 1)[:ncomp])[:, None]
        elif b_ell.ndim != 2 or b_ell.shape[0] != ncomp:
            raise ValueError('b_ell shape does not match ncomp')
        b_ell = np.ascontiguousarray(b_ell[:, :lmax + 1])
        spin = 0 if ncomp == 1 else [0, 2]
        self.prec_pinv = PseudoInvPreconditioner(lmax, icov_ell, icov_pix, geom, spin,
                                                 b_ell=b_ell)
        self.prec_masked_cg = None if no_masked_prec_cg else MaskedPreconditionerCG(
            lmax, icov_ell, spin, mask_bool[0], geom,
            lmax=lmax_prec_cg if lmax_prec_cg else lmax, nsteps=15, lmax_r_ell=None)
        self.prec_masked_mg = None if no_masked_prec_mg else MaskedPreconditioner(
            lmax, icov_ell[0:1, 0:1], 0, mask_bool[0], geom, min_pix=1000, n_jacobi=1,
            lmax_r_ell=6000)
        self.icov_ell, self.icov_pix, self.mask_bool = icov_ell, icov_pix, mask_bool
        self.b_ell, self.ncomp, self.swap_bm, self.lmax, self.spin = \
            b_ell, ncomp, swap_bm, lmax, spin

    def filter(self, imap, niter=15, niter_masked_cg=5, err_tol=1e-15):
        imap = np.asarray(imap, dtype=self.icov_pix.dtype)
        if imap.shape != (self.ncomp,) + tuple(self.shape_in):
            raise ValueError('imap shape does not match the filter geometry')
        imap = imap.reshape(self.ncomp, -1)[..., self.index]
        sol = WienerMap.from_arrays(imap, self.geom, self.lmax, self.icov_ell, self.icov_pix,
                                    b_ell=self.b_ell, mask_pix=self.mask_bool,
                                    swap_bm=self.swap_bm, spin=self.spin)
        sol.add_preconditioner(self.prec_pinv)
        if self.prec_masked_cg is not None:
            sol.add_preconditioner(self.prec_masked_cg)
        cg = sol.init_solver()
        errors = [np.nan]
        for idx in range(niter_masked_cg + niter):
            if idx == niter_masked_cg:
                sol.reset_preconditioner()
                sol.add_preconditioner(self.prec_pinv)
                if self.prec_masked_mg is not None:
                    sol.add_preconditioner(self.prec_masked_mg, sel=np.s_[0])
                sol.b_vec = sol.b0
                cg = sol.init_solver(x0=cg.x)
            cg.step()
            if idx >= niter_masked_cg and niter_masked_cg > 0:
                errors.append(cg.err * errors[niter_masked_cg - 1])
            else:
                errors.append(cg.err)
            if cg.err < err_tol:
                break
        return {'walm': sol.get_wiener(), 'ialm': sol.get_icov(), 'errors': errors}
