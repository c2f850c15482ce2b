'''
Conjugate-gradient Wiener filter on device alm, mirroring optweight's
solvers.CGWienerMap (solvers.py:8-456) and the pixell.utils.CG recurrence
it inherits (init / step as used at solvers.py:127-141):

  init : x = 0 or x0;  r = b (aliased) or b - A x0;  z = M r;  rz = <r,z>;
         rz0 = rz;  p = z;  i = 0;  err = inf
  step : Ap = A p;  alpha = rz/<p,Ap>;  x += alpha p;  r -= alpha Ap;
         z = M r;  rz' = <r,z>;  err = rz'/rz0;  beta = rz'/rz;  rz = rz';
         p = z + beta p;  i += 1

The CG scalars stay on the device (float64) and the vector updates are the
HIP kernels owmi_cg_xr / owmi_cg_p: one iteration needs no host
synchronisation; reading ``err`` synchronises.

A batch of right-hand sides, b of shape (nbatch, ncomp, nelem), runs as
nbatch independent CG recurrences sharing every operator application: the
default inner product is then per right-hand side (a device vector), the
updates are owmi_cg_xr_batch / owmi_cg_p_batch, and ``err`` is an array.

When the data model is pixel-diagonal (mask and inverse noise diagonal per
component, no swap_bm / lensing / filter / scale filter) the A operator
runs Y^T (M N^-1 M) Y through the fused synthesis -> pixel multiply ->
adjoint kernel (sht.pixel_sandwich).
'''
import numpy as np
import torch

from optweight_amd import (operators, alm_utils, map_utils, mat_utils, sht as sht_tools,
                           alm_c_utils, _lib)


def _is_fast(t):
    return (isinstance(t, torch.Tensor) and t.is_cuda and t.is_contiguous()
            and t.dtype in (torch.complex64, torch.complex128))


class CG:
    '''Preconditioned conjugate gradients with pixell.utils.CG semantics.'''

    def __init__(self, A, b, x0=None, M=lambda x: x, dot=None):
        if dot is None:
            dot = _default_dot(b)
        self.A = A
        self.b = b
        self.M = M
        self.dot = dot
        if x0 is None:
            self.x = b * 0
            self.r = b
        else:
            self.x = x0.clone() if isinstance(x0, torch.Tensor) else x0.copy()
            self.r = b - self.A(self.x)
        z = self.M(self.r)
        self.rz = self.dot(self.r, z)
        self.rz0 = _host(self.rz)
        self.p = z
        self.i = 0
        self._err = np.inf

    @property
    def err(self):
        e = self._err
        if isinstance(e, torch.Tensor):
            e = _host(e)
            self._err = e
        return e

    @err.setter
    def err(self, value):
        self._err = value

    def step(self):
        Ap = self.A(self.p)
        pAp = self.dot(self.p, Ap)
        dev_scalars = isinstance(self.rz, torch.Tensor) and isinstance(pAp, torch.Tensor)
        fast = (dev_scalars and _is_fast(self.x) and _is_fast(self.r) and _is_fast(self.p)
                and _is_fast(Ap))
        if fast and self.rz.ndim == 1:
            self._step_batch(Ap, pAp)
            return
        if fast:
            _lib.call('owmi_cg_xr', _lib.dtype_code(self.x), _lib.ptr(self.x), _lib.ptr(self.r),
                      _lib.ptr(self.p), _lib.ptr(Ap), _lib.ptr(self.rz), _lib.ptr(pAp),
                      self.x.numel(), _lib.stream_ptr(self.x.device))
        else:
            alpha = self.rz / pAp
            self.x += alpha * self.p
            self.r -= alpha * Ap
        del Ap
        z = self.M(self.r)
        next_rz = self.dot(self.r, z)
        self._err = next_rz / self.rz0
        if dev_scalars and isinstance(next_rz, torch.Tensor) and _is_fast(z) \
                and _is_fast(self.p) and z.shape == self.p.shape:
            self._unalias_p(z)
            _lib.call('owmi_cg_p', _lib.dtype_code(self.p), _lib.ptr(self.p), _lib.ptr(z),
                      _lib.ptr(next_rz), _lib.ptr(self.rz), self.p.numel(),
                      _lib.stream_ptr(self.p.device))
        else:
            beta = next_rz / self.rz
            self.p = z + beta * self.p
        self.rz = next_rz
        self.i += 1


    def _unalias_p(self, z):
        '''The in-place direction update p <- z + beta p must not write
        through p into r or z: with a preconditioner that returns its input
        (the identity), p, z and r start as one array, and the reference's
        out-of-place update leaves r alone.'''
        ptr = self.p.data_ptr()
        if ptr == z.data_ptr() or ptr == self.r.data_ptr():
            self.p = self.p.clone()

    def _step_batch(self, Ap, pAp):
        nb = self.rz.numel()
        n = self.x.numel() // nb
        st = _lib.stream_ptr(self.x.device)
        code = _lib.dtype_code(self.x)
        _lib.call('owmi_cg_xr_batch', code, _lib.ptr(self.x), code, _lib.ptr(self.r),
                  _lib.ptr(self.p), _lib.ptr(Ap), _lib.ptr(self.rz), _lib.ptr(pAp), n, nb, st)
        del Ap
        z = self.M(self.r)
        next_rz = self.dot(self.r, z)
        if getattr(self, '_rz0_dev', None) is None:
            self._rz0_dev = torch.as_tensor(self.rz0, device=next_rz.device)
        self._err = next_rz / self._rz0_dev
        self._unalias_p(z)
        _lib.call('owmi_cg_p_batch', code, _lib.ptr(self.p), _lib.ptr(z), _lib.ptr(next_rz),
                  _lib.ptr(self.rz), n, nb, st)
        self.rz = next_rz
        self.i += 1


def _host(v):
    '''float of a 0-d device scalar, numpy array of a vector of them.'''
    if isinstance(v, torch.Tensor):
        v = v.detach().cpu().numpy()
        return float(v) if v.ndim == 0 else v
    return v


def _default_dot(b):
    if isinstance(b, torch.Tensor):
        if b.ndim == 3:
            return alm_utils.contract_almxblm_batch
        return alm_utils.contract_almxblm_dev
    return alm_utils.contract_almxblm


class CGWienerMap(CG):
    '''Solve (S^-1 + P^T N^-1 P) x = P^T N^-1 d (+ constrained-realisation
    terms) with P = M Y B for device alm (npol, nelem) and maps
    (npol, npix). Same constructor, attributes and methods as optweight's
    CGWienerMap (lensing is not part of this path). ``filt`` is a pair of
    map-space callables (F, F^T) applied after the mask in the projection
    and before it in the adjoint (solvers.py:105-111, 177-215); the fused
    A operator is used only without one.'''

    def __init__(self, imap, icov_signal, icov_noise, sht, beam=None, lens=None, filt=None,
                 mask=None, rand_isignal=None, rand_inoise=None, swap_bm=False,
                 scale_filter=None):
        if lens is not None:
            raise NotImplementedError('lensing is not supported')
        if filt is not None and not (len(filt) == 2 and callable(filt[0])
                                     and callable(filt[1])):
            raise ValueError('filt must be a pair of callables (filter, filter_adjoint)')
        self.imap = imap
        self.icov_signal = icov_signal
        self.icov_noise = icov_noise
        self._beam_op = beam
        self.beam = beam if beam is not None else (lambda alm: alm)
        self.sht = sht[0]
        self.sht_adjoint = sht[1]
        self.lens = self.lens_adjoint = lambda alm: alm
        self._filt_op = filt
        if filt is None:
            self.filt = self.filt_adjoint = lambda omap: omap
        else:
            self.filt, self.filt_adjoint = filt
        self._mask_op = mask
        self.mask = mask if mask is not None else (lambda imap: imap)
        self.swap_bm = swap_bm
        self._sfilt_op = scale_filter
        self.scale_filter = scale_filter if scale_filter is not None else (lambda alm: alm)
        self._fast = None
        self.set_b_vec(self.imap, rand_isignal=rand_isignal, rand_inoise=rand_inoise)
        self.preconditioner = None

    # --------------------------------------------------------- set-up
    def init_solver(self, **kwargs):
        kwargs.setdefault('dot', _default_dot(self.b_vec))
        if self.preconditioner is not None:
            kwargs.setdefault('M', self.preconditioner)
        super().__init__(self.a_matrix, self.b_vec, **kwargs)

    def add_preconditioner(self, prec, sel=None):
        if self.preconditioner is None:
            if sel is None:
                self.preconditioner = prec
            else:
                def sliced_prec(alm):
                    idx = operators.component_index(alm, sel)
                    sub = alm[idx]
                    if isinstance(sub, torch.Tensor) and not sub.is_contiguous():
                        sub = sub.contiguous()
                    alm[idx] = prec(sub)
                    return alm
                self.preconditioner = sliced_prec
        else:
            self.preconditioner = operators.add_operators(self.preconditioner, prec,
                                                          slice_2=sel)

    def reset_preconditioner(self):
        self.preconditioner = None

    # -------------------------------------------------------- operators
    def proj(self, alm):
        if not self.swap_bm:
            alm = self.lens(alm)
            alm = self.beam(alm)
            omap = self.sht(alm)
            omap = self.mask(omap)
            return self.filt(omap)
        alm = self.mask(alm)
        alm = self.lens(alm)
        alm = self.beam(alm)
        omap = self.sht(alm)
        return self.filt(omap)

    def proj_adjoint(self, imap):
        if not self.swap_bm:
            omap = self.filt_adjoint(imap)
            omap = self.mask(omap)
            alm = self.sht_adjoint(omap)
            alm = self.beam(alm)
            return self.lens_adjoint(alm)
        omap = self.filt_adjoint(imap)
        alm = self.sht_adjoint(omap)
        alm = self.beam(alm)
        alm = self.lens_adjoint(alm)
        return self.mask(alm)

    def _fast_plan(self, alm):
        '''(pixel diagonal, beam matrix, S^-1 matrix, Y op) when A can run
        fused, else None.'''
        key = (str(alm.device), alm.dtype)
        if self._fast is None:
            self._fast = {}
        if key in self._fast:
            return self._fast[key]
        plan = None
        y, yt = self.sht, self.sht_adjoint
        ok = (not self.swap_bm and self._sfilt_op is None and self._filt_op is None
              and isinstance(y, operators.YMatVecAlm) and not y.qweight
              and isinstance(yt, operators.YTWMatVecMap) and not yt.qweight
              and isinstance(self.icov_noise, operators.PixMatVecMap)
              and isinstance(self.icov_signal, operators.EllMatVecAlm)
              and (self._beam_op is None or isinstance(self._beam_op, operators.EllMatVecAlm))
              and (self._mask_op is None or isinstance(self._mask_op, operators.PixMatVecMap)))
        if ok:
            rdt = operators._real_dtype(alm.dtype)
            # the pixel diagonal depends on the noise and mask operators only:
            # it is kept across resets of the plan (a new S^-1, as the
            # differentiable solve sets per evaluation), and with it the
            # spectral-sandwich multipliers cached on it
            # one entry per (device, dtype) and operator pair, so that the
            # float32 and float64 operators of a mixed solve do not evict each
            # other (a recording HIP graph finds its entry made eagerly)
            okey = (id(self.icov_noise), id(self._mask_op))
            cached = getattr(self, '_fast_diag', None)
            if cached is None or cached[0] != okey:
                cached = (okey, {}, self.icov_noise, self._mask_op)
                self._fast_diag = cached
            diags = cached[1]
            if key not in diags:
                d = self.icov_noise.diagonal(alm.device, rdt)
                if d is not None and self._mask_op is not None:
                    md = self._mask_op.diagonal(alm.device, rdt)
                    d = None if md is None else (md * d * md).contiguous()
                diags[key] = d
            d = diags[key]
            if d is not None:
                beam = (self._beam_op.device_matrix(alm.device, alm.dtype)
                        if self._beam_op is not None else None)
                plan = (d, beam, self.icov_signal.device_matrix(alm.device, alm.dtype), y)
        self._fast[key] = plan
        return plan

    def a_matrix(self, alm, dot_out=None):
        '''A alm. dot_out (batched alm, device tensors): float64 (nbatch,)
        that receives the real-field <alm, A alm> per right-hand side (on
        the fused path in the operator's last pass).'''
        fast = self._fast_plan(alm) if isinstance(alm, torch.Tensor) else None
        if fast is not None:
            d, beam, icov, y = fast
            ainfo = y.ainfo
            t = alm_c_utils.lmul(alm, beam, ainfo) if beam is not None else alm
            u = torch.empty_like(alm)
            sht_tools.pixel_sandwich(t, u, ainfo, y.minfo, y.spin, d)
            if alm.shape[-2] <= 4:
                # B u + S^-1 alm in one pass (and <alm, A alm> with it)
                return alm_c_utils.lmul_add(u, beam, alm, icov, ainfo, out=u, dot_out=dot_out)
            if beam is not None:
                alm_c_utils.lmul(u, beam, ainfo, inplace=True)
            u += alm_c_utils.lmul(alm, icov, ainfo)
            oalm = u
        else:
            a0 = alm
            alm = self.scale_filter(alm)
            omap = self.proj(alm)
            omap = self.icov_noise(omap)
            oalm = self.proj_adjoint(omap)
            oalm = self.scale_filter(oalm)
            oalm += self.scale_filter(self.icov_signal(alm))
            alm = a0
        if dot_out is not None:
            alm_utils.contract_almxblm_batch(alm, oalm, out=dot_out)
        return oalm

    def get_b_vec(self, imap):
        omap = self.icov_noise(imap)
        oalm = self.proj_adjoint(omap)
        return self.scale_filter(oalm)

    def get_b_vec_constr(self, rand_isignal, rand_inoise):
        oalm = self.scale_filter(self.proj_adjoint(rand_inoise))
        oalm += self.scale_filter(rand_isignal)
        return oalm

    def set_b_vec(self, imap, rand_isignal=None, rand_inoise=None):
        self.b_vec = self.get_b_vec(imap)
        if rand_isignal is not None and rand_inoise is not None:
            self.b_vec += self.get_b_vec_constr(rand_isignal, rand_inoise)
        self.b0 = self.b_vec.clone() if isinstance(self.b_vec, torch.Tensor) \
            else self.b_vec.copy()

    # ----------------------------------------------------------- output
    def get_wiener(self):
        return self.scale_filter(self.x.clone() if isinstance(self.x, torch.Tensor)
                                 else self.x.copy())

    def get_icov(self):
        return self.icov_signal(self.get_wiener())

    def _solver_dot(self):
        '''The inner product the solver was initialised with (pixell CG's
        ``dot``), the default one before init_solver.'''
        dot = getattr(self, 'dot', None)
        return dot if dot is not None else _default_dot(self.b0)

    def get_chisq(self):
        '''x^T S^-1 x + (d - P x)^T N^-1 (d - P x) of the scaled solution
        (solvers.py:331-338), through the solver's inner product; per
        right-hand side (an array) for a batch.'''
        dot = self._solver_dot()
        x_w = self.get_wiener()
        out = _host(dot(x_w, self.icov_signal(x_w)))
        res = self.imap - self.proj(x_w)
        prod = res * self.icov_noise(res)
        if isinstance(prod, torch.Tensor):
            batched = self.b0.ndim == 3
            pix = (prod.reshape(prod.shape[0], -1).sum(dim=1) if batched else prod.sum())
            pix = _host(pix.to(torch.float64))
        else:
            pix = np.sum(prod)
        return out + pix

    def get_residual(self):
        '''sqrt((A x - b)^T (A x - b)) under the solver's inner product
        (solvers.py:340-344).'''
        r = self.A(self.x) - self.b0
        return np.sqrt(_host(self._solver_dot()(r, r)))

    def get_qform(self):
        '''x^T A x / 2 - x^T b under the solver's inner product
        (solvers.py:346-349).'''
        dot = self._solver_dot()
        return 0.5 * _host(dot(self.x, self.A(self.x))) - _host(dot(self.x, self.b0))

    @classmethod
    def from_arrays(cls, imap, minfo, ainfo, icov_ell, icov_pix, *extra_args, b_ell=None,
                    mask_pix=None, minfo_mask=None, draw_constr=False, spin=None,
                    swap_bm=False, sfilt=None, lensop=None, seed=None, device=None):
        '''Build the solver from arrays; host inputs are moved to ``device``
        (default: current CUDA device), the alm precision follows imap.'''
        if lensop:
            raise NotImplementedError('lensing is not supported')
        if device is None:
            device = torch.device('cuda', torch.cuda.current_device())
        imap_np = imap.detach().cpu().numpy() if isinstance(imap, torch.Tensor) else np.asarray(imap)
        # (npol, npix) for one right-hand side, (nbatch, npol, npix) for a batch
        batched = imap_np.ndim == 3
        if spin is None:
            spin = sht_tools.default_spin(imap_np.shape[-2:] if batched else imap_np.shape)
        imap_np = mat_utils.atleast_nd(imap_np, 2)
        icov_pix = mat_utils.atleast_nd(np.asarray(icov_pix), 2)
        icov_ell = mat_utils.atleast_nd(np.asarray(icov_ell), 2)
        if icov_pix.shape[-2:] != imap_np.shape[-2:]:
            raise ValueError(f'Mismatch {icov_pix.shape=} and {imap_np.shape=}')
        if icov_ell.shape[0] != imap_np.shape[-2]:
            raise ValueError(f'Mismatch {icov_ell.shape=} and {imap_np.shape=}')
        icov_signal = operators.EllMatVecAlm(ainfo, icov_ell)
        icov_noise = operators.PixMatVecMap(icov_pix, 1, inplace=False)
        beam = operators.EllMatVecAlm(ainfo, b_ell) if b_ell is not None else None
        sf = operators.EllMatVecAlm(ainfo, sfilt) if sfilt is not None else None
        if mask_pix is not None:
            if minfo_mask is None:
                minfo_mask = minfo
            if swap_bm:
                mask = operators.PixMatVecAlm(ainfo, mask_pix, minfo_mask, spin,
                                              use_weights=True)
            else:
                mask = operators.PixMatVecMap(mask_pix, 1, inplace=False)
        else:
            mask = None
        rdt = torch.float64 if imap_np.dtype == np.float64 else torch.float32
        cdt = torch.complex128 if rdt == torch.float64 else torch.complex64
        if draw_constr:
            # a batch draws right-hand side i from the seed sequence [seed, i]
            seeds = ([None if seed is None else [seed, i] for i in range(imap_np.shape[0])]
                     if batched else [seed])
            cnp = np.complex128 if cdt == torch.complex128 else np.complex64
            rs = [alm_utils.rand_alm(icov_ell, ainfo, sd, dtype=cnp) for sd in seeds]
            rn = [map_utils.rand_map_pix(icov_pix, sd) for sd in seeds]
            rand_isignal = torch.as_tensor(np.stack(rs) if batched else rs[0], device=device)
            rand_inoise = torch.as_tensor(np.stack(rn) if batched else rn[0],
                                          device=device).to(rdt)
        else:
            rand_isignal = rand_inoise = None
        sht = (operators.YMatVecAlm(ainfo, minfo, spin, qweight=False),
               operators.YTWMatVecMap(minfo, ainfo, spin, qweight=False))
        imap_dev = torch.as_tensor(np.ascontiguousarray(imap_np), device=device).to(rdt)
        return cls(imap_dev, icov_signal, icov_noise, sht, beam=beam, mask=mask,
                   rand_isignal=rand_isignal, rand_inoise=rand_inoise, swap_bm=swap_bm,
                   scale_filter=sf)


def _from_arrays_const_cor(cls, imap, minfo, ainfo, icov_ell, icov_pix, icov_noise_ell,
                           nsteps=5, b_ell=None, mask_pix=None, minfo_mask=None,
                           no_masked_noise=False, draw_constr=False, spin=None, sfilt=None,
                           lensop=None, seed=None, verbose=False, device=None):
    '''CGWienerMap with the constant-correlation noise model
    N = M N_pix^1/2 Y C_ell Y^T N_pix^1/2 M (solvers.py:459-575): N^-1 is
    applied by nsteps of CG (InvPixEllPixMatVecMap with icov_pix^-1/2 and
    icov_noise_ell^-1), the constrained-realisation noise draw by its
    inverse square root on unit-variance alm.'''
    if lensop:
        raise NotImplementedError('lensing is not supported')
    if device is None:
        device = torch.device('cuda', torch.cuda.current_device())
    imap_np = imap.detach().cpu().numpy() if isinstance(imap, torch.Tensor) else np.asarray(imap)
    batched = imap_np.ndim == 3
    if spin is None:
        spin = sht_tools.default_spin(imap_np.shape[-2:] if batched else imap_np.shape)
    imap_np = mat_utils.atleast_nd(imap_np, 2)
    icov_pix = mat_utils.atleast_nd(np.asarray(icov_pix), 2)
    icov_ell = mat_utils.atleast_nd(np.asarray(icov_ell), 2)
    icov_noise_ell = mat_utils.atleast_nd(np.asarray(icov_noise_ell), 2)
    if icov_pix.shape[-2:] != imap_np.shape[-2:]:
        raise ValueError(f'Mismatch {icov_pix.shape=} and {imap_np.shape=}')
    if icov_ell.shape[0] != imap_np.shape[-2]:
        raise ValueError(f'Mismatch {icov_ell.shape=} and {imap_np.shape=}')
    icov_signal = operators.EllMatVecAlm(ainfo, icov_ell)
    noise_mask = None if no_masked_noise else mask_pix
    icov_noise = operators.InvPixEllPixMatVecMap(icov_pix, icov_noise_ell, minfo, spin,
                                                 power_m=-0.5, power_x=-1, mask=noise_mask,
                                                 nsteps=nsteps, verbose=verbose)
    beam = operators.EllMatVecAlm(ainfo, b_ell) if b_ell is not None else None
    sf = operators.EllMatVecAlm(ainfo, sfilt) if sfilt is not None else None
    mask = operators.PixMatVecMap(mask_pix, 1, inplace=False) if mask_pix is not None else None
    rdt = torch.float64 if imap_np.dtype == np.float64 else torch.float32
    cdt = torch.complex128 if rdt == torch.float64 else torch.complex64
    if draw_constr:
        seeds = ([None if seed is None else [seed, i] for i in range(imap_np.shape[0])]
                 if batched else [seed])
        cnp = np.complex128 if cdt == torch.complex128 else np.complex64
        rs = [alm_utils.rand_alm(icov_ell, ainfo, sd, dtype=cnp) for sd in seeds]
        rn = [alm_utils.unit_var_alm(ainfo, imap_np.shape[-2:-1], sd, dtype=cnp) for sd in seeds]
        rand_isignal = torch.as_tensor(np.stack(rs) if batched else rs[0], device=device)
        sqrt_op = operators.InvSqrtPixEllPixMatVecMap(icov_pix, icov_noise_ell, minfo, spin,
                                                      power_m=-0.5, power_x=-1, mask=noise_mask,
                                                      inv_op=icov_noise, verbose=verbose)
        rand_inoise = sqrt_op(torch.as_tensor(np.stack(rn) if batched else rn[0], device=device))
    else:
        rand_isignal = rand_inoise = None
    sht = (operators.YMatVecAlm(ainfo, minfo, spin, qweight=False),
           operators.YTWMatVecMap(minfo, ainfo, spin, qweight=False))
    imap_dev = torch.as_tensor(np.ascontiguousarray(imap_np), device=device).to(rdt)
    return cls(imap_dev, icov_signal, icov_noise, sht, beam=beam, mask=mask,
               rand_isignal=rand_isignal, rand_inoise=rand_inoise, scale_filter=sf)


CGWienerMap.from_arrays_const_cor = classmethod(_from_arrays_const_cor)


def _from_arrays_fwav(cls, imap, minfo, ainfo, icov_ell, cov_wav, fkernelset, *extra_args,
                      b_ell=None, mask_pix=None, minfo_mask=None, draw_constr=False, spin=None,
                      swap_bm=False, sfilt=None, cov_noise_2d=None, lensop=None, seed=None,
                      device=None):
    '''CGWienerMap with the 2-D Fourier-wavelet noise model
    (solvers.py:577-695): N^-1 = F^T X^-1/2 (K^T F M_wav F^T K)^-1 X^-1/2 F
    on maps of a rectangular grid (operators.FInvFWavFMatVecMap, 3 CG steps;
    X = cov_noise_2d, the identity when absent). The constrained-realisation
    noise draw maps unit-variance wavelet maps (unseeded, as in the
    reference) through InvSqrtFWavMatVecWav and X^-1/2.'''
    from optweight_amd import dft, noise_utils
    if lensop:
        raise NotImplementedError('lensing is not supported')
    if device is None:
        device = torch.device('cuda', torch.cuda.current_device())
    imap_np = imap.detach().cpu().numpy() if isinstance(imap, torch.Tensor) else np.asarray(imap)
    batched = imap_np.ndim == 3
    if spin is None:
        spin = sht_tools.default_spin(imap_np.shape[-2:] if batched else imap_np.shape)
    imap_np = mat_utils.atleast_nd(imap_np, 2)
    icov_ell = mat_utils.atleast_nd(np.asarray(icov_ell), 2)
    if icov_ell.shape[0] != imap_np.shape[-2]:
        raise ValueError(f'Mismatch {icov_ell.shape=} and {imap_np.shape=}')
    icov_signal = operators.EllMatVecAlm(ainfo, icov_ell)
    power_x = -0.5 if cov_noise_2d is not None else 0
    icov_noise = operators.FInvFWavFMatVecMap(minfo, cov_wav, fkernelset, cov_noise_2d,
                                              power_x=power_x, nsteps=3)
    beam = operators.EllMatVecAlm(ainfo, b_ell) if b_ell is not None else None
    sf = operators.EllMatVecAlm(ainfo, sfilt) if sfilt is not None else None
    if mask_pix is not None:
        if minfo_mask is None:
            minfo_mask = minfo
        if swap_bm:
            mask = operators.PixMatVecAlm(ainfo, mask_pix, minfo_mask, spin, use_weights=True)
        else:
            mask = operators.PixMatVecMap(mask_pix, 1, inplace=False)
    else:
        mask = None
    rdt = torch.float64 if imap_np.dtype == np.float64 else torch.float32
    cdt = torch.complex128 if rdt == torch.float64 else torch.complex64
    imap_dev = torch.as_tensor(np.ascontiguousarray(imap_np), device=device).to(rdt)
    if draw_constr:
        if batched:
            raise NotImplementedError('constrained realisations of a batch need one solver '
                                      'per right-hand side with this noise model')
        cnp = np.complex128 if cdt == torch.complex128 else np.complex64
        rand_isignal = torch.as_tensor(alm_utils.rand_alm(icov_ell, ainfo, seed, dtype=cnp),
                                       device=device)
        wav_draw = noise_utils.unit_var_wav(cov_wav.get_minfos_diag(), imap_np.shape[:-1],
                                            cov_wav.dtype, seed=None).to(device)
        sqrt_op = operators.InvSqrtFWavMatVecWav(cov_wav, fkernelset,
                                                 inv_op=icov_noise.inv_op)
        fmap_draw = icov_noise.x_op(sqrt_op(wav_draw)).to(cdt)
        rand_inoise = torch.zeros_like(imap_dev)
        dft.irfft(fmap_draw, map_utils.view_2d(rand_inoise, minfo))
    else:
        rand_isignal = rand_inoise = None
    sht = (operators.YMatVecAlm(ainfo, minfo, spin, qweight=False),
           operators.YTWMatVecMap(minfo, ainfo, spin, qweight=False))
    return cls(imap_dev, icov_signal, icov_noise, sht, beam=beam, mask=mask,
               rand_isignal=rand_isignal, rand_inoise=rand_inoise, swap_bm=swap_bm,
               scale_filter=sf)


CGWienerMap.from_arrays_fwav = classmethod(_from_arrays_fwav)


class CGWiener(CG):
    """Solve (S^-1 + B F^T M N^-1 M F B) x = B F^T M N^-1 a (+ w_s + B F^T M
    w_n for constrained realisations) entirely in harmonic space, with the
    constructor, attributes and methods of optweight's CGWiener
    (solvers.py:697-1139). Operators are callables on device alm
    (..., npol, nelem); a leading batch axis solves independent right-hand
    sides together. Map-space filters (2D Fourier) are not part of this
    path."""

    def __init__(self, alm_data, icov_signal, icov_noise, beam=None, filt=None, mask=None,
                 rand_isignal=None, rand_inoise=None, swap_bm=False):
        self.alm_data = alm_data
        self.icov_signal = icov_signal
        self.icov_noise = icov_noise
        self.beam = beam if beam is not None else (lambda alm: alm)
        if filt is None:
            self.filt = self.filt_adjoint = lambda alm: alm
        else:
            self.filt, self.filt_adjoint = filt
        self.mask = mask if mask is not None else (lambda alm: alm)
        self.rand_isignal = rand_isignal
        self.rand_inoise = rand_inoise
        self.swap_bm = swap_bm
        if self.rand_isignal is not None and self.rand_inoise is not None:
            self.b_vec = self.get_b_vec_constr(self.alm_data)
        else:
            self.b_vec = self.get_b_vec(self.alm_data)
        self.b0 = self.b_vec.clone()
        self.preconditioner = None

    def init_solver(self, **kwargs):
        kwargs.setdefault('dot', _default_dot(self.b_vec))
        if self.preconditioner is not None:
            kwargs.setdefault('M', self.preconditioner)
        CG.__init__(self, self.a_matrix, self.b_vec, **kwargs)

    add_preconditioner = CGWienerMap.add_preconditioner
    reset_preconditioner = CGWienerMap.reset_preconditioner

    def a_matrix(self, alm):
        if self.swap_bm:
            n = self.filt(self.beam(self.mask(alm.clone())))
        else:
            n = self.mask(self.filt(self.beam(alm.clone())))
        n = self.icov_noise(n)
        if self.swap_bm:
            n = self.mask(self.beam(self.filt_adjoint(n)))
        else:
            n = self.beam(self.filt_adjoint(self.mask(n)))
        return self.icov_signal(alm.clone()) + n

    def get_b_vec(self, alm):
        alm = self.icov_noise(alm.clone())
        if self.swap_bm:
            return self.mask(self.beam(self.filt_adjoint(alm)))
        return self.beam(self.filt_adjoint(self.mask(alm)))

    def get_b_vec_constr(self, alm):
        alm = self.get_b_vec(alm)
        if self.swap_bm:
            alm += self.mask(self.beam(self.filt_adjoint(self.rand_inoise.clone())))
        else:
            alm += self.beam(self.filt_adjoint(self.mask(self.rand_inoise.clone())))
        alm += self.rand_isignal
        return alm

    def get_wiener(self):
        return self.x.clone()

    def get_icov(self):
        return self.icov_signal(self.x.clone())

    def get_chisq(self):
        x_w = self.get_wiener()
        out = _host(self.dot(x_w, self.icov_signal(x_w)))
        d = self.alm_data - x_w
        out += _host(self.dot(self.beam(self.filt_adjoint(self.mask(d))),
                              self.icov_noise(self.beam(self.filt(self.mask(d))))))
        return out

    def get_residual(self):
        r = self.A(self.x) - self.b0
        return np.sqrt(_host(self.dot(r, r)))

    def get_qform(self):
        return (0.5 * _host(self.dot(self.x, self.A(self.x)))
                - _host(self.dot(self.x, self.b0)))

    @classmethod
    def from_arrays(cls, alm_data, ainfo, icov_ell, icov_pix, minfo, *extra_args, b_ell=None,
                    mask_pix=None, draw_constr=False, spin=None, icov_noise_flat_ell=None,
                    swap_bm=False, kfilt=None, minfo_kfilt=None, seed=None, device=None):
        """Harmonic-space solver with a pixel-diagonal noise model
        (solvers.py:958-1046): N^-1 = Y^T N^-1_pix Y, mask = Y^T W M Y."""
        if icov_noise_flat_ell is not None:
            raise NotImplementedError('icov_noise_flat_ell is not supported')
        if kfilt is not None:
            raise NotImplementedError('Fourier filters are not part of this path')
        alm_data = _to_device_alm(alm_data, device)
        if spin is None:
            spin = sht_tools.default_spin(tuple(alm_data.shape[-2:]) if alm_data.ndim == 3
                                          else tuple(alm_data.shape))
        icov_signal = operators.EllMatVecAlm(ainfo, icov_ell)
        icov_noise = operators.PixMatVecAlm(ainfo, icov_pix, minfo, spin)
        beam = operators.EllMatVecAlm(ainfo, b_ell) if b_ell is not None else None
        mask = (operators.PixMatVecAlm(ainfo, mask_pix, minfo, spin, use_weights=True)
                if mask_pix is not None else None)
        if draw_constr:
            cnp = np.complex128 if alm_data.dtype == torch.complex128 else np.complex64
            rand_isignal = torch.as_tensor(alm_utils.rand_alm(icov_ell, ainfo, seed, dtype=cnp),
                                           device=alm_data.device)
            rand_inoise = alm_utils.rand_alm_pix(np.asarray(icov_pix), ainfo, minfo, spin, seed,
                                                 adjoint=True, device=alm_data.device)
            rand_inoise = rand_inoise.to(alm_data.dtype)
        else:
            rand_isignal = rand_inoise = None
        return cls(alm_data, icov_signal, icov_noise, *extra_args, beam=beam, mask=mask,
                   rand_isignal=rand_isignal, rand_inoise=rand_inoise, swap_bm=swap_bm)

    @classmethod
    def from_arrays_wav(cls, alm_data, ainfo, icov_ell, icov_wav, w_ell, *extra_args,
                        b_ell=None, mask_pix=None, minfo_mask=None, draw_constr=False,
                        spin=None, swap_bm=False, icov_noise_ell=None, kfilt=None,
                        minfo_kfilt=None, seed=None, device=None):
        """Harmonic-space solver with a wavelet-domain noise model
        (solvers.py:1048-1139): N^-1 = K^T Y^T N^-1_wav Y K (WavMatVecAlm),
        optionally flattened by icov_noise_ell^(1/2) on both sides."""
        if kfilt is not None:
            raise NotImplementedError('Fourier filters are not part of this path')
        alm_data = _to_device_alm(alm_data, device)
        if spin is None:
            spin = sht_tools.default_spin(tuple(alm_data.shape[-2:]) if alm_data.ndim == 3
                                          else tuple(alm_data.shape))
        icov_signal = operators.EllMatVecAlm(ainfo, icov_ell)
        if icov_noise_ell is None:
            icov_noise = operators.WavMatVecAlm(ainfo, icov_wav, w_ell, spin)
        else:
            icov_noise = operators.EllWavEllMatVecAlm(ainfo, icov_wav, w_ell, icov_noise_ell,
                                                      spin, power_x=0.5)
        beam = operators.EllMatVecAlm(ainfo, b_ell) if b_ell is not None else None
        mask = (operators.PixMatVecAlm(ainfo, mask_pix, minfo_mask, spin, use_weights=True)
                if mask_pix is not None else None)
        if draw_constr:
            cnp = np.complex128 if alm_data.dtype == torch.complex128 else np.complex64
            rand_isignal = torch.as_tensor(alm_utils.rand_alm(icov_ell, ainfo, seed, dtype=cnp),
                                           device=alm_data.device)
            rand_inoise = alm_utils.rand_alm_wav(icov_wav, ainfo, w_ell, spin, seed,
                                                 adjoint=True, device=alm_data.device)
            rand_inoise = rand_inoise.to(alm_data.dtype)
            if icov_noise_ell is not None:
                rand_inoise = operators.EllMatVecAlm(ainfo, icov_noise_ell, power=0.5)(
                    rand_inoise)
        else:
            rand_isignal = rand_inoise = None
        return cls(alm_data, icov_signal, icov_noise, *extra_args, beam=beam, mask=mask,
                   rand_isignal=rand_isignal, rand_inoise=rand_inoise, swap_bm=swap_bm)


def _to_device_alm(alm, device):
    if isinstance(alm, torch.Tensor):
        return alm if device is None else alm.to(device)
    if device is None:
        device = torch.device('cuda', torch.cuda.current_device())
    return torch.as_tensor(np.ascontiguousarray(alm), device=device)
