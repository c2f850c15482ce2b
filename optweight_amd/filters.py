'''
The two-phase Wiener filter of CAR maps, mirroring optweight's
filters.CGPixFilter (filters.py:9-301): a harmonic signal model from theory
spectra and a pixel inverse-covariance noise model; phase 1 runs CG with the
pseudo-inverse plus MaskedPreconditionerCG, phase 2 restarts from its
solution with the pseudo-inverse plus the multigrid preconditioner on T.

The map's pixels are gathered once into the unit-stride ring order of its
geometry (match_enmap_minfo, map_utils.unit_stride), so every transform of
the solve runs on a unit-stride plan; the outputs are alm.
'''
from timeit import default_timer as timer
import warnings

import numpy as np
import torch

from optweight_amd import map_utils, mat_utils, solvers, preconditioners, alm_utils


def alm2cl(alm, blm, ainfo):
    '''Cross spectra C_l[i, j] = (Re a_l0 b*_l0 + 2 sum_{m>0} Re a_lm b*_lm) /
    (2l + 1) of device alm (npol, nelem) (host numpy (npol, npol, lmax+1)).'''
    a = alm.detach().cpu().numpy() if isinstance(alm, torch.Tensor) else np.asarray(alm)
    b = blm.detach().cpu().numpy() if isinstance(blm, torch.Tensor) else np.asarray(blm)
    ells, ms = ainfo.lm_arrays()
    fac = np.where(ms == 0, 1.0, 2.0)
    prod = np.real(a[:, None, :] * np.conj(b[None, :, :])) * fac
    out = np.zeros(prod.shape[:2] + (ainfo.lmax + 1,))
    for i in range(out.shape[0]):
        for j in range(out.shape[1]):
            out[i, j] = np.bincount(ells, weights=prod[i, j], minlength=ainfo.lmax + 1)
    return out / (2 * np.arange(ainfo.lmax + 1) + 1)


class CGPixFilter:
    '''Wiener / inverse-variance filter of CAR maps (filters.py:9-167).

    theory_cls: dict of TT (and TE, EE, BB) spectra from ell = 0; b_ell:
    (nell,) or (ncomp, nell) beam; icov_pix: (ncomp, Ny, Nx) or
    (ncomp, ncomp, Ny, Nx) inverse pixel covariance; mask_bool: (ncomp, Ny, Nx)
    or (Ny, Nx), True where observed. The CAR geometry comes from
    ``icov_pix.wcs`` (a pixell ndmap) or the ``wcs`` argument (any object
    with astropy's ``wcs.wcs.cdelt / crpix / crval``, e.g.
    map_utils.CarWCS). The remaining arguments follow the reference.'''

    def __init__(self, theory_cls, b_ell, icov_pix, mask_bool, include_te=True, q_low=0,
                 q_high=1, swap_bm=False, scale_a=False, lmax=None, lmax_prec_cg=None,
                 no_masked_prec_mg=False, no_masked_prec_cg=False, wcs=None, device=None):
        if wcs is None:
            wcs = getattr(icov_pix, 'wcs', None)
        if wcs is None:
            raise ValueError('icov_pix carries no wcs: pass wcs=')
        b_ell = np.asarray(b_ell)
        icov_pix = np.asarray(icov_pix)
        if np.any(np.logical_not(np.isfinite(b_ell))):
            raise ValueError('b_ell is not finite')
        if np.any(np.logical_not(np.isfinite(icov_pix))):
            raise ValueError('icov_pix is not finite')
        if device is None:
            device = torch.device('cuda', torch.cuda.current_device())
        self.device = device
        shape_in = icov_pix.shape[-2:]
        icov_pix = mat_utils.atleast_nd(icov_pix, 3)
        mask_bool = mat_utils.atleast_nd(np.asarray(mask_bool).astype(bool, copy=False), 3)
        ncomp = icov_pix.shape[0]
        if mask_bool.shape[0] == 1:
            mask_bool = (np.ones(ncomp)[:, np.newaxis, np.newaxis] * mask_bool).astype(bool)
        minfo_map = None
        for mtype in ('CC', 'fejer1'):
            try:
                minfo_map = map_utils.match_enmap_minfo(icov_pix.shape, wcs, mtype=mtype)
            except ValueError:
                continue
            break
        if minfo_map is None:
            raise ValueError('map geometry is neither CC nor fejer1 CAR')
        minfo, index = map_utils.unit_stride(minfo_map)
        if lmax is None:
            lmax = map_utils.minfo2lmax(minfo)
        icov_pix = icov_pix.reshape(icov_pix.shape[:-2] + (-1,))[..., index]
        mask_bool = mask_bool.reshape(mask_bool.shape[:-2] + (-1,))[..., index]
        if q_low != 0 or q_high != 1:
            icov_pix = map_utils.threshold_icov(icov_pix, q_low=q_low, q_high=q_high)
        tlmax = theory_cls['TT'].size - 1
        if not tlmax >= lmax:
            raise ValueError(f'theory_cls reach lmax {tlmax} < {lmax}')
        cov_ell = np.zeros((ncomp, ncomp, lmax + 1))
        cov_ell[0, 0] = theory_cls['TT'][:lmax + 1]
        if ncomp > 1:
            if include_te:
                cov_ell[0, 1] = theory_cls['TE'][:lmax + 1]
                cov_ell[1, 0] = theory_cls['TE'][:lmax + 1]
            cov_ell[1, 1] = theory_cls['EE'][:lmax + 1]
            cov_ell[2, 2] = theory_cls['BB'][:lmax + 1]
        icov_ell = np.zeros_like(cov_ell)
        for lidx in range(icov_ell.shape[-1]):
            icov_ell[:, :, lidx] = np.linalg.pinv(cov_ell[:, :, lidx])
        if b_ell.ndim == 1:
            b_ell = b_ell[np.newaxis] * np.asarray((1, 1, 1)[:ncomp])[:, np.newaxis]
        elif b_ell.ndim != 2 or b_ell.shape[0] != ncomp:
            raise ValueError(f'b_ell shape {b_ell.shape} does not match ncomp {ncomp}')
        b_ell = np.ascontiguousarray(b_ell[:, :lmax + 1])
        if scale_a:
            sfilt = mat_utils.matpow(b_ell, -0.5)
            lmax_mg = 3000
        else:
            sfilt = None
            lmax_mg = 6000
        ainfo = alm_utils.AlmInfo(lmax)
        spin = 0 if ncomp == 1 else [0, 2]
        self.prec_pinv = preconditioners.PseudoInvPreconditioner(
            ainfo, icov_ell, icov_pix, minfo, spin, b_ell=b_ell, sfilt=sfilt)
        if no_masked_prec_cg:
            self.prec_masked_cg = None
        else:
            self.prec_masked_cg = preconditioners.MaskedPreconditionerCG(
                ainfo, icov_ell, spin, mask_bool[0].astype(bool), minfo,
                lmax=lmax_prec_cg if lmax_prec_cg else lmax, nsteps=15, lmax_r_ell=None,
                sfilt=sfilt)
        if no_masked_prec_mg:
            self.prec_masked_mg = None
        else:
            self.prec_masked_mg = preconditioners.MaskedPreconditioner(
                ainfo, icov_ell[0:1, 0:1], 0, mask_bool[0], minfo, min_pix=1000, n_jacobi=1,
                lmax_r_ell=lmax_mg, sfilt=None if sfilt is None else sfilt[0:1, 0:1],
                device=device)
        self.shape_in = shape_in
        self.icov_ell = icov_ell
        self.icov_pix = icov_pix
        self.mask_bool = mask_bool
        self.minfo = minfo
        self.minfo_map = minfo_map
        self.index = index
        self.b_ell = b_ell
        self.sfilt = sfilt
        self.ncomp = ncomp
        self.swap_bm = swap_bm
        self.lmax = lmax
        self.ainfo = ainfo
        self.spin = spin

    def filter(self, imap, niter=None, niter_masked_cg=5, benchmark=False, verbose=True,
               err_tol=1e-15, draw_constr=False, seed=None):
        '''Run niter_masked_cg phase-1 steps and niter phase-2 steps
        (filters.py:169-301); returns a dict with 'walm' and 'ialm' (device
        alm), 'solver', 'errors' and, with benchmark = k, the statistics at
        every k-th step.'''
        imap = np.asarray(imap)
        if imap.shape != (self.ncomp,) + tuple(self.shape_in):
            raise ValueError(f'imap shape {imap.shape} != {(self.ncomp,) + tuple(self.shape_in)}')
        imap = imap.astype(self.icov_pix.dtype, copy=False).reshape(self.ncomp, -1)[
            ..., self.index]
        solver = solvers.CGWienerMap.from_arrays(
            imap, self.minfo, self.ainfo, self.icov_ell, self.icov_pix, b_ell=self.b_ell,
            draw_constr=draw_constr, mask_pix=self.mask_bool, swap_bm=self.swap_bm,
            spin=self.spin, sfilt=self.sfilt, seed=seed, device=self.device)
        solver.add_preconditioner(self.prec_pinv)
        if self.prec_masked_cg is not None:
            solver.add_preconditioner(self.prec_masked_cg)
        solver.init_solver()
        errors = [np.nan]
        chisqs, residuals, qforms, ps_c_ells, itnums = [], [], [], [], []
        if benchmark:
            warnings.warn('optweight: Benchmarking is turned on. '
                          'This significantly slows down the filtering.')
            chisqs.append(solver.get_chisq())
            residuals.append(solver.get_residual())
            itnums.append(0)
            if verbose:
                print('|b| :', np.sqrt(alm_utils.contract_almxblm(solver.b0, solver.b0)))
        if niter is None:
            niter = 15
            warnings.warn(f'optweight: Using the default number of iterations :'
                          f'{niter_masked_cg=} + {niter=}.')
        for idx in range(niter_masked_cg + niter):
            if idx == niter_masked_cg:
                solver.reset_preconditioner()
                solver.add_preconditioner(self.prec_pinv)
                if self.prec_masked_mg is not None:
                    solver.add_preconditioner(self.prec_masked_mg, sel=np.s_[0])
                solver.b_vec = solver.b0
                solver.init_solver(x0=solver.x)
            t_start = timer()
            solver.step()
            err = solver.err
            t_eval = timer() - t_start
            if idx >= niter_masked_cg and niter_masked_cg > 0:
                errors.append(err * errors[niter_masked_cg - 1])
            else:
                errors.append(err)
            if benchmark and (idx + 1) % benchmark == 0:
                chisqs.append(solver.get_chisq())
                residuals.append(solver.get_residual())
                qforms.append(solver.get_qform())
                w = solver.get_wiener()
                ps_c_ells.append(alm2cl(w, w, self.ainfo))
                itnums.append(idx)
                print(f'optweight benchmark: \t chisq : {chisqs[-1]:.2f} \t '
                      f'residual : {residuals[-1]:.2f} \t qform : {qforms[-1]:.2f}')
            if verbose:
                print(f'optweight step {idx + 1} / {niter_masked_cg + niter},'
                      f', error {errors[-1]:.2e}, time {t_eval:.3f} s')
            if err < err_tol:
                warnings.warn(f'Stopping early because solver.err={err} is below {err_tol=}')
                break
        output = {'walm': solver.get_wiener(), 'ialm': solver.get_icov(), 'solver': solver,
                  'errors': errors}
        if benchmark:
            output.update(chisqs=chisqs, residuals=residuals, qforms=qforms, ps=ps_c_ells,
                          itnums=itnums)
        return output
