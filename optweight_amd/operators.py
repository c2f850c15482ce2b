'''
Linear operators as callables on device tensors, mirroring the Wiener-path
subset of optweight's operators module: MatVecAlm / MatVecMap ABCs (9-47),
YMatVecAlm (49-103), YTWMatVecMap (105-160), PixMatVecMap (162-204),
EllMatVecAlm (206-257), PixMatVecAlm (562-643), PixEllPixMatVecMap
(991-1102), InvSqrtPixEllPixMatVecMap (1104-1200),
InvPixEllPixMatVecMap (1202-1284), op2mat (1419-1456), add_operators
(1458-1489), and the 2-D Fourier-wavelet noise operators: FMatVecF
(259-309), FWavMatVecF (311-362), InvFWavMatVecF (364-421),
FInvFWavFMatVecMap (423-486), InvSqrtFWavMatVecWav (488-560),
WavMatVecWav (914-989).

Host-side matrices (numpy) are moved to the device once per (device,
dtype). When a pixel matrix is diagonal, PixMatVecAlm runs the fused
synthesis -> pixel multiply -> adjoint kernel (sht.pixel_sandwich) instead
of materialising the map.
'''
from abc import ABC, abstractmethod

import numpy as np
import torch

from optweight_amd import sht, alm_c_utils, mat_utils, map_utils, alm_utils, dft, wavtrans


class MatVecAlm(ABC):
    def __call__(self, alm, *args, **kwargs):
        return self.call(alm, *args, **kwargs)

    @abstractmethod
    def call(self, alm, *args, **kwargs):
        pass


class MatVecMap(ABC):
    def __call__(self, imap, *args, **kwargs):
        return self.call(imap, *args, **kwargs)

    @abstractmethod
    def call(self, imap, *args, **kwargs):
        pass


def _real_dtype(dtype):
    return {torch.complex128: torch.float64, torch.complex64: torch.float32,
            torch.float64: torch.float64, torch.float32: torch.float32}[dtype]


def _complex_dtype(dtype):
    return {torch.float64: torch.complex128, torch.float32: torch.complex64,
            torch.complex128: torch.complex128, torch.complex64: torch.complex64}[dtype]


class _DeviceCache:
    '''Device copies of a host array, per (device, dtype).'''

    def __init__(self, arr):
        self.host = arr
        self._cache = {}

    def get(self, device, dtype):
        key = (str(device), dtype)
        if key not in self._cache:
            a = self.host
            if isinstance(a, torch.Tensor):
                t = a.to(device=device, dtype=dtype)
            else:
                t = torch.as_tensor(np.asarray(a), device=device).to(dtype)
            self._cache[key] = t.contiguous()
        return self._cache[key]


class YMatVecAlm(MatVecAlm):
    '''Y a (or W Y a with qweight=True).'''

    def __init__(self, ainfo, minfo, spin, qweight=False):
        self.ainfo = ainfo
        self.minfo = minfo
        self.spin = spin
        self.qweight = qweight

    def call(self, alm, omap=None):
        if omap is None:
            omap = torch.zeros(tuple(alm.shape[:-1]) + (self.minfo.npix,),
                               dtype=_real_dtype(alm.dtype), device=alm.device)
        sht.alm2map(alm, omap, self.ainfo, self.minfo, self.spin, adjoint=self.qweight)
        return omap


class YTWMatVecMap(MatVecMap):
    '''Y^T W m (or Y^T m with qweight=False).'''

    def __init__(self, minfo, ainfo, spin, qweight=True):
        self.minfo = minfo
        self.ainfo = ainfo
        self.spin = spin
        self.qweight = qweight

    def call(self, imap, oalm=None):
        if oalm is None:
            oalm = torch.zeros(tuple(imap.shape[:-1]) + (self.ainfo.nelem,),
                               dtype=_complex_dtype(imap.dtype), device=imap.device)
        sht.map2alm(imap, oalm, self.minfo, self.ainfo, self.spin, adjoint=not self.qweight)
        return oalm


class PixMatVecMap(MatVecMap):
    '''M^p m for M diagonal in pixels: (npol, npol, npix) or (npol, npix).'''

    def __init__(self, m_pix, power, inplace=False):
        if power != 1:
            m_pix = mat_utils.matpow(m_pix, power, return_diag=True)
        self.m_pix = m_pix
        self._dev = _DeviceCache(m_pix)
        self.inplace = inplace

    def call(self, imap):
        if not isinstance(imap, torch.Tensor):
            return mat_utils.matvec(self.m_pix, imap, inplace=self.inplace)
        m = self._dev.get(imap.device, imap.dtype)
        return mat_utils.matvec(m, imap, inplace=self.inplace)

    def diagonal(self, device, dtype):
        '''(npol, npix) device diagonal, or None if the matrix has
        off-diagonal terms.'''
        m = np.asarray(self.m_pix)
        if m.ndim == 2:
            return self._dev.get(device, dtype)
        npol = m.shape[0]
        off = m * (1 - np.eye(npol))[:, :, None]
        if np.any(off):
            return None
        key = ('diag', str(device), dtype)
        if key not in self._dev._cache:
            d = np.ascontiguousarray(np.einsum('iip->ip', m))
            self._dev._cache[key] = torch.as_tensor(d, device=device).to(dtype)
        return self._dev._cache[key]


class EllMatVecAlm(MatVecAlm):
    '''M^p alm for M diagonal in l, (npol, npol, nell) or (npol, nell).'''

    def __init__(self, ainfo, m_ell, power=1, inplace=False):
        m_ell = mat_utils.full_matrix(np.asarray(m_ell))
        if power != 1:
            m_ell = mat_utils.matpow(m_ell, power)
        self.m_ell = m_ell
        self._dev = _DeviceCache(m_ell)
        self.ainfo = ainfo
        self.inplace = inplace

    def device_matrix(self, device, dtype):
        return self._dev.get(device, _real_dtype(dtype))

    def call(self, alm):
        m = self.device_matrix(alm.device, alm.dtype)
        if self.inplace:
            return alm_c_utils.lmul(alm, m, self.ainfo, inplace=True)
        return alm_c_utils.lmul(alm, m, self.ainfo, alm_out=torch.empty_like(alm))

    def apply_copy(self, alm):
        '''M^p alm into a new array, whatever ``inplace`` says.'''
        m = self.device_matrix(alm.device, alm.dtype)
        return alm_c_utils.lmul(alm, m, self.ainfo, alm_out=torch.empty_like(alm))


def _pix_alm_weights(adjoint, use_weights):
    '''(weight after synthesis, weight before adjoint) of PixMatVecAlm.'''
    a2 = adjoint if use_weights else not adjoint
    return bool(adjoint), not a2


class PixMatVecAlm(MatVecAlm):
    '''Y^T [W] M^p [W] Y alm with the weight placement of
    operators.py:627-641 (adjoint -> W after Y; use_weights selects the
    weight before Y^T).'''

    def __init__(self, ainfo, m_pix, minfo, spin, power=1, inplace=False, adjoint=False,
                 use_weights=False):
        m_pix = mat_utils.full_matrix(np.asarray(m_pix))
        if power != 1:
            m_pix = mat_utils.matpow(m_pix, power)
        self.m_pix = m_pix
        self._dev = _DeviceCache(m_pix)
        self.ainfo = ainfo
        self.minfo = minfo
        self.spin = spin
        self.inplace = inplace
        self.adjoint = adjoint
        self.use_weights = use_weights
        npol = m_pix.shape[0]
        off = m_pix * (1 - np.eye(npol))[:, :, None]
        self._is_diag = not np.any(off)
        self._diag_cache = {}

    def fused_diagonal(self, device, dtype):
        '''Per-component pixel diagonal including the quadrature weights,
        for the fused kernel; None when the matrix is not diagonal.'''
        if not self._is_diag:
            return None
        key = (str(device), dtype)
        if key not in self._diag_cache:
            d = np.ascontiguousarray(np.einsum('iip->ip', self.m_pix)).astype(np.float64)
            w_after, w_before = _pix_alm_weights(self.adjoint, self.use_weights)
            w = np.ones(self.minfo.nrow)
            if w_after:
                w = w * self.minfo.weight
            if w_before:
                w = w * self.minfo.weight
            if w_after or w_before:
                d = map_utils._ringweight_np(d.copy(), self.minfo, w)
            self._diag_cache[key] = torch.as_tensor(d, device=device).to(dtype)
        return self._diag_cache[key]

    def call(self, alm):
        out = alm if self.inplace else torch.empty_like(alm)
        rdt = _real_dtype(alm.dtype)
        d = self.fused_diagonal(alm.device, rdt)
        if d is not None:
            if self.inplace:
                res = torch.empty_like(alm)
                sht.pixel_sandwich(alm, res, self.ainfo, self.minfo, self.spin, d)
                alm.copy_(res)
                return alm
            return sht.pixel_sandwich(alm, out, self.ainfo, self.minfo, self.spin, d)
        omap = torch.zeros(tuple(alm.shape[:-1]) + (self.minfo.npix,), dtype=rdt,
                           device=alm.device)
        sht.alm2map(alm, omap, self.ainfo, self.minfo, self.spin, adjoint=self.adjoint)
        imap = mat_utils.matvec(self._dev.get(alm.device, rdt), omap)
        a2 = self.adjoint if self.use_weights else not self.adjoint
        sht.map2alm(imap, out, self.minfo, self.ainfo, self.spin, adjoint=a2)
        return out


def _diag_of(mat):
    """(npol, npix) diagonal of a pixel matrix (npol, npix) or
    (npol, npol, npix), or None if the latter has off-diagonal terms."""
    mat = np.asarray(mat)
    if mat.ndim == 2:
        return mat
    npol = mat.shape[0]
    if np.any(mat * (1 - np.eye(npol))[:, :, None]):
        return None
    return np.ascontiguousarray(np.einsum('iip->ip', mat))


class WavMatVecAlm(MatVecAlm):
    """K^T Y^T M Y K alm (adjoint: K^T Y^T W M W Y K) for a wavelet block
    matrix M (wavtrans.Wav) and axisymmetric kernels K = w_ell
    (operators.py:705-815). Scale j' of the input is synthesised on the
    Gauss-Legendre grid of block (j', j'); block (j, j') multiplies it and
    the result goes back through K_j. A diagonal block with a per-component
    pixel diagonal runs as one fused synthesis -> multiply -> adjoint
    transform (sht.pixel_sandwich)."""

    def __init__(self, ainfo, m_wav, w_ell, spin, power=1, adjoint=False):
        from optweight_amd import wavtrans
        self.ainfo = ainfo
        self.w_ell = np.asarray(w_ell)
        self.spin = spin
        self.adjoint = adjoint
        if power != 1:
            m_wav = mat_utils.wavmatpow(m_wav, power)
        self.m_wav = m_wav
        self.npol = wavtrans.preshape2npol(m_wav.preshape)
        self.blocks = []
        for key in m_wav.maps:
            j, jp = key
            minfo = m_wav.minfos[(jp, jp)]
            lmax_g = (int(minfo.nphi[0]) - 1) // 2
            lmax_w = wavtrans.wav_lmax(self.w_ell, jp)
            mat = m_wav.maps[key]
            mat = mat.detach().cpu().numpy() if isinstance(mat, torch.Tensor) else np.asarray(mat)
            diag = _diag_of(mat) if j == jp else None
            fused = diag is not None and lmax_w <= lmax_g <= ainfo.lmax
            if fused:
                d = np.array(diag, dtype=np.float64)
                if adjoint:
                    w = np.asarray(minfo.weight, dtype=np.float64) ** 2
                    d = map_utils._ringweight_np(d, minfo, w)
                cache = _DeviceCache(d)
            else:
                cache = _DeviceCache(mat)
            self.blocks.append({'j': j, 'jp': jp, 'minfo': minfo, 'lmax_g': lmax_g,
                                'lmax_w': lmax_w, 'fused': fused, 'mat': cache})

    def call(self, alm):
        out = torch.zeros_like(alm)
        rdt = _real_dtype(alm.dtype)
        wlm_cache = {}
        for b in self.blocks:
            j, jp, minfo = b['j'], b['jp'], b['minfo']
            m = b['mat'].get(alm.device, rdt)
            if b['fused']:
                L = b['lmax_g']
                winfo = alm_utils.AlmInfo(L)
                wlm, _ = alm_utils.trunc_alm(alm, self.ainfo, L)
                alm_c_utils.lmul(wlm, self.w_ell[jp, :L + 1], winfo, inplace=True)
                res = torch.empty_like(wlm)
                sht.pixel_sandwich(wlm, res, winfo, minfo, self.spin, m)
            else:
                if jp not in wlm_cache:
                    wlms, winfos = alm_utils.alm2wlm_axisym(
                        alm, self.ainfo, self.w_ell[jp:jp + 1], lmaxs=[b['lmax_w']])
                    wlm_cache[jp] = (wlms[0], winfos[0])
                wlm, winf = wlm_cache[jp]
                vmap = torch.zeros(tuple(alm.shape[:-1]) + (minfo.npix,), dtype=rdt,
                                   device=alm.device)
                sht.alm2map(wlm, vmap, winf, minfo, self.spin, adjoint=self.adjoint)
                prod = mat_utils.matvec(m, vmap)
                winfo = alm_utils.AlmInfo(b['lmax_g'])
                res = torch.zeros(tuple(alm.shape[:-1]) + (winfo.nelem,), dtype=alm.dtype,
                                  device=alm.device)
                sht.map2alm(prod, res, minfo, winfo, self.spin, adjoint=not self.adjoint)
            alm_utils.wlm2alm_axisym([res], [winfo], self.w_ell[j:j + 1], alm=out,
                                     ainfo=self.ainfo)
        return out


class EllWavEllMatVecAlm(MatVecAlm):
    """X_ell^q M_wav^p X_ell^q alm (operators.py:645-703)."""

    def __init__(self, ainfo, m_wav, w_ell, x_ell, spin, power_m=1, power_x=1, adjoint=False):
        self.x_ell_op = EllMatVecAlm(ainfo, x_ell, power=power_x, inplace=False)
        self.m_wav_op = WavMatVecAlm(ainfo, m_wav, w_ell, spin, power=power_m, adjoint=adjoint)

    def call(self, alm):
        alm = self.x_ell_op(alm)
        alm = self.m_wav_op(alm)
        return self.x_ell_op(alm)


class PixEllPixMatVecMap(MatVecMap):
    '''M_pix^p Y X_ell^q Y^T M_pix^p m (no quadrature weights unless
    adjoint / use_weights say so), maps (npol, npix).'''

    def __init__(self, m_pix, x_ell, minfo, spin, power_m=1, power_x=1, inplace=False,
                 adjoint=False, use_weights=False, lmax=None):
        if m_pix is not None and power_m != 1:
            m_pix = mat_utils.matpow(m_pix, power_m, return_diag=True)
        if power_x != 1:
            x_ell = mat_utils.matpow(x_ell, power_x, return_diag=True)
        self.m_pix = m_pix
        self.x_ell = x_ell
        self._m = _DeviceCache(m_pix) if m_pix is not None else None
        self._x = _DeviceCache(x_ell)
        self.npol = np.asarray(x_ell).shape[0]
        if lmax is None:
            lmax = map_utils.minfo2lmax(minfo)
        self.ainfo = alm_utils.AlmInfo(lmax)
        self.minfo = minfo
        self.spin = spin
        self.inplace = inplace
        self.adjoint = adjoint
        self.use_weights = use_weights
        # M vanishes off these rings: the transforms skip the others exactly
        # (None: M touches every ring, the plan of the whole geometry)
        self.rings = (map_utils.ring_support(m_pix, minfo) if m_pix is not None
                      and np.asarray(m_pix).shape[-1] == minfo.npix else None)
        if self.rings is not None and self.rings.size == minfo.nrow:
            self.rings = None

    def _apply_m(self, out):
        if self._m is None:
            return out
        m = self._m.get(out.device, out.dtype)
        if m.ndim == 3:
            return mat_utils.matvec(m, out, inplace=True)
        if m.dtype == torch.bool:
            return out.mul_(m)
        return mat_utils.matvec(m, out, inplace=True)

    def call(self, imap):
        assert tuple(imap.shape[-2:]) == (self.npol, self.minfo.npix), (
            f'Wrong imap shape: {tuple(imap.shape)} != {(self.npol, self.minfo.npix)}')
        m = self._m.get(imap.device, imap.dtype) if self._m is not None else None
        if not self.inplace and m is not None and m.ndim == 2:
            out = mat_utils.matvec(m, imap.contiguous())  # multiplied copy in one pass
        else:
            out = imap if self.inplace else imap.clone()
            self._apply_m(out)
        # the analysis overwrites every coefficient
        alm = torch.empty(tuple(imap.shape[:-1]) + (self.ainfo.nelem,),
                          dtype=_complex_dtype(imap.dtype), device=imap.device)
        sht.map2alm(out, alm, self.minfo, self.ainfo, self.spin,
                    adjoint=self.adjoint is self.use_weights, rings=self.rings)
        alm_c_utils.lmul(alm, self._x.get(imap.device, _real_dtype(imap.dtype)), self.ainfo,
                         inplace=True)
        # pixels off self.rings keep their masked (zero) input values
        sht.alm2map(alm, out, self.ainfo, self.minfo, self.spin, adjoint=self.adjoint,
                    rings=self.rings)
        self._apply_m(out)
        return out


def map_dot(x, y):
    '''<x, y> of real maps in float64: a 0-d tensor for one map set
    (npol, npix), one value per right-hand side (nb,) for a batch.'''
    if x.ndim >= 3:
        return (x * y).reshape(x.shape[0], -1).sum(dim=1, dtype=torch.float64)
    return (x * y).sum(dtype=torch.float64)


def _ratio(num, den):
    '''num / den per right-hand side, 0 where den is 0 (a right-hand side
    whose input is zero stays zero instead of turning into NaN).'''
    return torch.where(den != 0, num / torch.where(den != 0, den, torch.ones_like(den)),
                       torch.zeros_like(num))


def inner_cg(a_op, m_op, b, nsteps, dot=map_dot, verbose=False, batched=None):
    '''nsteps of preconditioned CG from x = 0 on maps (the recurrence and
    order of operations of pixell.utils.CG), with per-right-hand-side
    scalars for a batch (nb, npol, npix) (batched None: when b.ndim >= 3)
    and zero-safe divisions.'''
    if batched is None:
        batched = b.ndim >= 3
    bshape = (-1,) + (1,) * (b.ndim - 1) if batched else ()
    x = torch.zeros_like(b)
    r = b
    z = m_op(r)
    rz = dot(r, z)
    p = z
    if b.is_complex() and b.is_cuda and b.is_contiguous():
        return _inner_cg_complex(a_op, m_op, x, r, p, rz, nsteps, dot, verbose,
                                 b.shape[0] if batched else 1)
    for idx in range(nsteps):
        Ap = a_op(p)
        alpha = _ratio(rz, dot(p, Ap)).to(b.dtype).reshape(bshape)
        x += alpha * p
        r = r - alpha * Ap
        del Ap
        z = m_op(r)
        rz_next = dot(r, z)
        beta = _ratio(rz_next, rz).to(b.dtype).reshape(bshape)
        rz = rz_next
        p = z + beta * p
        if verbose:
            print(idx, rz.cpu().numpy())
    return x


def _inner_cg_complex(a_op, m_op, x, r, p, rz, nsteps, dot, verbose, nb):
    '''The loop of inner_cg on complex device vectors with the CG update
    kernels: x += alpha p, r -= alpha A p and p = z + beta p in place, alpha
    and beta formed on the device (0 for a zero denominator). r is the
    caller's right-hand side array, updated in place.'''
    from optweight_amd import _lib
    code = _lib.dtype_code(x)
    n = x.numel() // nb
    st = _lib.stream_ptr(x.device)
    if p.data_ptr() == r.data_ptr():
        p = p.clone()
    for idx in range(nsteps):
        ap = a_op(p).contiguous()
        pap = dot(p, ap)
        _lib.call('owmi_cg_xr_batch', code, _lib.ptr(x), code, _lib.ptr(r), _lib.ptr(p),
                  _lib.ptr(ap), _lib.ptr(rz), _lib.ptr(pap), n, nb, st)
        del ap
        z = m_op(r).contiguous()
        rz_next = dot(r, z)
        if z.data_ptr() == p.data_ptr():
            z = z.clone()
        _lib.call('owmi_cg_p_batch', code, _lib.ptr(p), _lib.ptr(z), _lib.ptr(rz_next),
                  _lib.ptr(rz), n, nb, st)
        rz = rz_next
        if verbose:
            print(idx, rz.cpu().numpy())
    return x


class InvPixEllPixMatVecMap(MatVecMap):
    '''(M' M_pix^p Y X_ell^q Y^T M_pix^p M')^-1 m by nsteps of CG
    (operators.py:1202-1284), preconditioned by the same form with powers
    -p, -q and the quadrature weights on the other side; M' an optional
    binary mask applied on both sides of both operators.'''

    def __init__(self, m_pix, x_ell, minfo, spin, power_m=1, power_x=1, mask=None,
                 adjoint=False, lmax=None, nsteps=5, verbose=False):
        self._a_mat = PixEllPixMatVecMap(m_pix, x_ell, minfo, spin, power_m=power_m,
                                         power_x=power_x, inplace=False, adjoint=adjoint,
                                         lmax=lmax)
        self._prec = PixEllPixMatVecMap(m_pix, x_ell, minfo, spin, power_m=-power_m,
                                        power_x=-power_x, inplace=False, adjoint=not adjoint,
                                        lmax=lmax)
        self.mask = mask
        self._mask = _DeviceCache(np.asarray(mask)) if mask is not None else None
        self.nsteps = nsteps
        self.verbose = verbose

    def _masked(self, op):
        if self._mask is None:
            return op

        def f(imap):
            m = self._mask.get(imap.device, imap.dtype)
            return m * op(m * imap)
        return f

    def call(self, imap):
        return inner_cg(self._masked(self._a_mat), self._masked(self._prec), imap.clone(),
                        self.nsteps, verbose=self.verbose)


class InvSqrtPixEllPixMatVecMap(MatVecMap):
    '''Inverse square root of a PixEllPixMatVecMap noise model applied to
    alm (operators.py:1104-1200): N^-1 (M' M_pix^p Y X_ell^(q/2) alm), the
    inverse by the CG of InvPixEllPixMatVecMap (inv_op, or a new one).'''

    def __init__(self, m_pix, x_ell, minfo, spin, power_m=1, power_x=1, mask=None,
                 adjoint=False, lmax=None, inv_op=None, verbose=False):
        if inv_op:
            self.inv_op = inv_op
        else:
            self.inv_op = InvPixEllPixMatVecMap(m_pix, x_ell, minfo, spin, power_m=power_m,
                                                power_x=power_x, mask=mask, adjoint=adjoint,
                                                lmax=lmax, verbose=verbose)
        self.npol = np.asarray(x_ell).shape[0]
        self.minfo = minfo
        self.spin = spin
        self.adjoint = adjoint
        self.x_ell_halfq = mat_utils.matpow(x_ell, power_x / 2, return_diag=True)
        self._xh = _DeviceCache(self.x_ell_halfq)
        self.mask = mask
        self._mask = _DeviceCache(np.asarray(mask)) if mask is not None else None

    def _sqrt_n(self, ialm):
        ainfo = self.inv_op._a_mat.ainfo
        alm = alm_c_utils.lmul(ialm, self._xh.get(ialm.device, _real_dtype(ialm.dtype)), ainfo)
        out = torch.zeros(tuple(ialm.shape[:-1]) + (self.minfo.npix,),
                          dtype=_real_dtype(ialm.dtype), device=ialm.device)
        sht.alm2map(alm, out, ainfo, self.minfo, self.spin, adjoint=self.adjoint)
        out = self.inv_op._a_mat._apply_m(out)
        if self._mask is not None:
            out *= self._mask.get(out.device, out.dtype)
        return out

    def call(self, ialm):
        return self.inv_op(self._sqrt_n(ialm))


def op2mat(op, nrow, dtype, ncol=None, input_shape=None, device=None):
    '''Matrix of a linear operator from its action on unit vectors (host
    numpy operator, or device operator when device is given).'''
    ncol = nrow if ncol is None else ncol
    if device is None:
        uvec = np.zeros(ncol, dtype=dtype)
        shaped = uvec.reshape(input_shape) if input_shape is not None else uvec
        mat = np.zeros((nrow, ncol), dtype=dtype)
        for idx in range(ncol):
            uvec[idx] = 1
            mat[:, idx] = np.asarray(op(shaped)).reshape(-1)
            uvec[:] = 0
        return mat
    tdt = dtype if isinstance(dtype, torch.dtype) else torch.float64
    mat = torch.zeros((nrow, ncol), dtype=tdt, device=device)
    uvec = torch.zeros(ncol, dtype=tdt, device=device)
    for idx in range(ncol):
        uvec[idx] = 1
        shaped = uvec.reshape(input_shape) if input_shape is not None else uvec
        mat[:, idx] = op(shaped).reshape(-1)
        uvec[idx] = 0
    return mat


def component_index(vec, sel):
    """Index taking components ``sel`` of ``vec``: ``vec[sel]`` for one
    right-hand side (ncomp, nelem), as the reference indexes; for a batch
    (nbatch, ncomp, nelem) the same components of every right-hand side,
    keeping the component axis (an integer selects a length-1 slice)."""
    if sel is None or vec.ndim <= 2:
        return sel
    if vec.ndim != 3:
        raise ValueError(f'expected (nbatch, ncomp, nelem), got {tuple(vec.shape)}')
    if isinstance(sel, (int, np.integer)):
        return (slice(None), slice(int(sel), int(sel) + 1))
    if isinstance(sel, slice):
        return (slice(None), sel)
    raise NotImplementedError(f'component selection {sel!r} on a batch of right-hand sides')


def add_operators(op_1, op_2, slice_2=None):
    """Callable acting as op_1(v) + op_2(v[slice_2]) scattered into slice_2
    (per right-hand side for a batch, see component_index)."""

    def add(ivec):
        # operators that never write to their input (preserves_input) get
        # the input itself; others a copy, as the reference gives them
        if getattr(op_1, 'preserves_input', False):
            ovec = op_1(ivec)
        else:
            ovec = op_1(ivec.clone() if isinstance(ivec, torch.Tensor) else ivec.copy())
        idx = component_index(ivec, slice_2)
        if idx is None:
            idx = slice(None)
        sub = ivec[idx]
        sub = (sub.clone(memory_format=torch.contiguous_format)
               if isinstance(sub, torch.Tensor) else sub.copy())
        ovec[idx] += op_2(sub)
        return ovec

    return add


class MatVecF(ABC):
    def __call__(self, fmap, *args, **kwargs):
        return self.call(fmap, *args, **kwargs)

    @abstractmethod
    def call(self, fmap, *args, **kwargs):
        pass


class MatVecWav(ABC):
    def __call__(self, wav, *args, **kwargs):
        return self.call(wav, *args, **kwargs)

    @abstractmethod
    def call(self, wav, *args, **kwargs):
        pass


class WavMatVecWav(MatVecWav):
    """Wavelet block matrix times a wavelet vector, in place: map j' of the
    vector is multiplied by block (j, j') for every block present, in key
    order (operators.py:914-989); blocks are (npol, npol, npix) matrices or
    (npol, npix) diagonals, raised to power at construction."""

    def __init__(self, m_wav, power=1, inplace=False, op=None):
        if m_wav.ndim != 2:
            raise ValueError(f'Expected m_wav.ndim = 2, got : {m_wav.ndim}')
        if power != 1:
            m_wav = mat_utils.wavmatpow(m_wav, power, return_diag=True)
        self.m_wav = m_wav
        self.inplace = inplace
        if not self.inplace:
            raise NotImplementedError()
        self.op = op
        self._dev = {key: _DeviceCache(m) for key, m in m_wav.maps.items()}

    def call(self, wav):
        if wav.ndim != 1:
            raise ValueError(f'Expected wav.ndim = 1, got : {wav.ndim}')
        nrow, ncol = self.m_wav.shape
        for jidx in range(nrow):
            for jpidx in range(ncol):
                if (jidx, jpidx) not in self._dev:
                    continue
                vec = wav.maps[jpidx]
                mat = self._dev[(jidx, jpidx)].get(vec.device, vec.dtype)
                if self.op is None:
                    mat_utils.matvec(mat, vec, inplace=True)
                else:
                    vec.copy_(torch.einsum(self.op, mat, vec))
        return wav


class FMatVecF(MatVecF):
    """M^p f for M diagonal in the 2-D Fourier plane: (npol, npol, nly, nlx)
    matrices or (npol, nly, nlx) diagonals (operators.py:259-309)."""

    def __init__(self, fmat2d, power=1, inplace=False):
        fmat2d = np.asarray(fmat2d)
        if power != 1:
            flat = fmat2d.reshape(fmat2d.shape[:-2] + (-1,))
            fmat2d = mat_utils.matpow(flat, power, return_diag=True).reshape(fmat2d.shape)
        self.fmat2d = fmat2d
        self._dev = _DeviceCache(fmat2d)
        self.inplace = inplace

    def call(self, fmap):
        m = self._dev.get(fmap.device, _real_dtype(fmap.dtype))
        return dft.fmul(fmap, fmat2d=m, out=fmap if self.inplace else None)


class FWavMatVecF(MatVecF):
    """K^T F M_wav^p F^T K f: 2-D coefficients to Fourier-wavelet maps
    (f2wav), the wavelet block matrix, and back (wav2f)
    (operators.py:311-362). The wavelet maps live in a template vector per
    (device, dtype), shared with other operators through wav_template."""

    def __init__(self, m_wav, fkernelset, power=1, wav_template=None):
        self.m_wav_op = WavMatVecWav(m_wav, power=power, inplace=True)
        self.fkernelset = fkernelset
        self._preshape = tuple(m_wav.preshape[-1:])
        if isinstance(wav_template, dict):
            self.wav_template = wav_template
        else:
            self.wav_template = {}
            if wav_template is not None:
                self.wav_template['given'] = wav_template

    def _template(self, device, dtype):
        if 'given' in self.wav_template:
            return self.wav_template['given']
        key = (str(device), dtype)
        if key not in self.wav_template:
            npdt = np.float64 if dtype == torch.float64 else np.float32
            self.wav_template[key] = self.fkernelset.get_wav_vec(self._preshape, dtype=npdt,
                                                                 device=device)
        return self.wav_template[key]

    def call(self, fmap):
        tmpl = self._template(fmap.device, _real_dtype(fmap.dtype))
        wavtrans.f2wav(fmap, tmpl, self.fkernelset)
        self.m_wav_op(tmpl)
        out = torch.empty_like(fmap)
        wavtrans.wav2f(tmpl, out, self.fkernelset)
        return out


class InvFWavMatVecF(MatVecF):
    """(K^T F M_wav F^T K)^-1 f by nsteps of CG from zero, preconditioned by
    the same operator with M_wav^-1, under the real-field inner product of
    2-D coefficients (operators.py:364-421)."""

    def __init__(self, m_wav, fkernelset, nsteps=3):
        self.fkernelset = fkernelset
        self.nsteps = nsteps
        self._a_mat = FWavMatVecF(m_wav, fkernelset, power=1)
        self._prec = FWavMatVecF(m_wav, fkernelset, power=-1,
                                 wav_template=self._a_mat.wav_template)

    def call(self, fmap, verbose=False):
        return inner_cg(self._a_mat, self._prec, fmap.clone(), self.nsteps,
                        dot=dft.contract_fxg_dev, verbose=verbose, batched=False)


class FInvFWavFMatVecMap(MatVecMap):
    """F^T X^q (K^T F M_wav F^T K)^-1 X^q F m of a map on a rectangular grid
    (operators.py:423-486): X a 2-D Fourier-diagonal matrix raised to
    power_x (identity for power_x 0), the inverse by InvFWavMatVecF.
    Leading dimensions beyond (npol, npix) are independent maps."""

    def __init__(self, minfo, m_wav, fkernelset, fmat2d, power_x=1, nsteps=3):
        self.minfo = minfo
        self.inv_op = InvFWavMatVecF(m_wav, fkernelset, nsteps=nsteps)
        if power_x != 0:
            self.x_op = FMatVecF(fmat2d, power=power_x, inplace=True)
        else:
            self.x_op = lambda x: x

    def _one(self, imap, verbose):
        imap2 = map_utils.view_2d(imap, self.minfo)
        fmap = dft.allocate_fmap(tuple(imap2.shape), imap2.dtype, device=imap.device)
        omap = torch.zeros_like(imap2)
        dft.rfft(imap2, fmap)
        self.x_op(fmap)
        fmap = self.inv_op(fmap, verbose=verbose)
        self.x_op(fmap)
        dft.irfft(fmap, omap)
        return map_utils.view_1d(omap, self.minfo)

    def call(self, imap, verbose=False):
        imap = imap.contiguous()
        if imap.ndim >= 3:
            flat = imap.reshape((-1,) + tuple(imap.shape[-2:]))
            return torch.stack([self._one(m, verbose) for m in flat]).reshape(imap.shape)
        return self._one(imap, verbose)


class InvSqrtFWavMatVecWav(MatVecWav):
    """(K^T F M_wav F^T K)^-1 K^T F M_wav^1/2 w of a wavelet vector w
    (operators.py:488-560); M_wav^1/2 is applied to w in place."""

    def __init__(self, m_wav, fkernelset, inv_op=None):
        self.sqrt_cov_wav_op = WavMatVecWav(m_wav, power=0.5, inplace=True)
        self.fkernelset = fkernelset
        self.inv_op = InvFWavMatVecF(m_wav, fkernelset) if inv_op is None else inv_op
        self.oshape = tuple(m_wav.preshape[-1:]) + tuple(fkernelset.shape_full)

    def _sqrt_n(self, wav):
        first = wav.maps[wav.get_indices_diag()[0]]
        out = torch.zeros(self.oshape, dtype=_complex_dtype(first.dtype), device=first.device)
        self.sqrt_cov_wav_op(wav)
        wavtrans.wav2f(wav, out, self.fkernelset)
        return out

    def call(self, wav, **kwargs):
        return self.inv_op(self._sqrt_n(wav), **kwargs)
