'''
The two-phase Wiener filter of CAR maps with the interface of optweight's
filters.CGPixFilter (filters.py:9-301): a harmonic signal model built from
theory spectra, a pixel inverse-covariance noise model, and a solve in two
phases on one CGWienerMap:

  phase 1  (niter_masked_cg steps)  pseudo-inverse + MaskedPreconditionerCG
  phase 2  (niter steps)            pseudo-inverse + multigrid on T, CG
                                    restarted from phase 1's solution

The map's pixels are gathered once into the unit-stride ring order of its
geometry (map_utils.match_enmap_minfo, map_utils.unit_stride), so every
transform of the solve runs on a unit-stride plan; the outputs are device
alm.
'''
from timeit import default_timer as timer
import warnings

import numpy as np
import torch

from optweight_amd import map_utils, mat_utils, solvers, preconditioners, alm_utils

# multigrid low-pass scale of phase 2, without and with the beam rescaling
_LMAX_MG = {False: 6000, True: 3000}


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


def car_geometry(shape, wcs):
    '''(map geometry, its unit-stride form, gather index) of a CAR map of
    ``shape`` and ``wcs``: Clenshaw-Curtis if the pixels sit on the poles,
    else Fejer's first rule; ValueError if neither fits.'''
    errors = []
    for rule in ('CC', 'fejer1'):
        try:
            minfo_map = map_utils.match_enmap_minfo(shape, wcs, mtype=rule)
        except ValueError as err:
            errors.append(f'{rule}: {err}')
            continue
        minfo, index = map_utils.unit_stride(minfo_map)
        return minfo_map, minfo, index
    raise ValueError('the map is neither a Clenshaw-Curtis nor a Fejer-1 CAR map ('
                     + '; '.join(errors) + ')')


def signal_icov_ell(theory_cls, ncomp, lmax, include_te=True):
    '''Inverse signal covariance (ncomp, ncomp, lmax+1): per l the
    pseudo-inverse of the TT (and TE, EE, BB) spectra matrix.'''
    have = theory_cls['TT'].size - 1
    if have < lmax:
        raise ValueError(f'theory_cls reach l = {have}, below lmax = {lmax}')
    spectra = {(0, 0): 'TT'}
    if ncomp > 1:
        spectra.update({(1, 1): 'EE', (2, 2): 'BB'})
        if include_te:
            spectra.update({(0, 1): 'TE', (1, 0): 'TE'})
    cov = np.zeros((lmax + 1, ncomp, ncomp))
    for (i, j), name in spectra.items():
        cov[:, i, j] = np.asarray(theory_cls[name])[:lmax + 1]
    return np.ascontiguousarray(np.linalg.pinv(cov).transpose(1, 2, 0))


def beam_per_component(b_ell, ncomp, lmax):
    '''(ncomp, lmax+1) beam from a shared (nell,) beam or one per component.'''
    b_ell = np.asarray(b_ell)
    if b_ell.ndim == 1:
        b_ell = np.broadcast_to(b_ell, (ncomp,) + b_ell.shape)
    elif b_ell.ndim != 2 or b_ell.shape[0] != ncomp:
        raise ValueError(f'b_ell of shape {b_ell.shape} is neither (nell,) nor ({ncomp}, nell)')
    return np.ascontiguousarray(b_ell[:, :lmax + 1], dtype=np.float64)


class CGPixFilter:
    '''Wiener / inverse-variance filter of CAR maps (filters.py:9-167).

    theory_cls: dict of TT (and TE, EE, BB) spectra from l = 0; b_ell:
    (nell,) or (ncomp, nell) beam; icov_pix: (ncomp, Ny, Nx) or
    (ncomp, ncomp, Ny, Nx) inverse pixel covariance; mask_bool: (ncomp, Ny, Nx)
    or (Ny, Nx), True where observed; ncomp is 1 (T) or 3 (T, Q, U). The CAR
    geometry comes from ``icov_pix.wcs`` (a pixell ndmap) or the ``wcs``
    argument (any object with astropy's ``wcs.wcs.cdelt / crpix / crval``,
    e.g. map_utils.CarWCS). q_low / q_high threshold icov_pix by quantiles;
    scale_a rescales the signal by b_ell^-1/2; lmax_prec_cg caps the band
    limit of phase 1's masked CG; no_masked_prec_cg / no_masked_prec_mg drop
    the phase-1 / phase-2 masked preconditioner.'''

    def __init__(self, theory_cls, b_ell, icov_pix, mask_bool, include_te=True, q_low=0,
                 q_high=1, swap_bm=False, scale_a=False, lmax=None, lmax_prec_cg=None,
                 no_masked_prec_mg=False, no_masked_prec_cg=False, wcs=None, device=None):
        wcs = getattr(icov_pix, 'wcs', None) if wcs is None else wcs
        if wcs is None:
            raise ValueError('icov_pix carries no wcs: pass wcs=')
        icov_pix = np.asarray(icov_pix)
        for name, arr in (('b_ell', b_ell), ('icov_pix', icov_pix)):
            if not np.all(np.isfinite(arr)):
                raise ValueError(f'{name} holds non-finite values')
        self.device = (torch.device('cuda', torch.cuda.current_device()) if device is None
                       else device)
        self.shape_in = tuple(icov_pix.shape[-2:])
        icov_pix = mat_utils.atleast_nd(icov_pix, 3)
        self.ncomp = ncomp = icov_pix.shape[0]
        if ncomp not in (1, 3):
            raise ValueError(f'{ncomp} components: the filter takes T (1) or T, Q, U (3)')
        self.spin = 0 if ncomp == 1 else [0, 2]

        self.minfo_map, self.minfo, self.index = car_geometry(icov_pix.shape, wcs)
        self.lmax = lmax = map_utils.minfo2lmax(self.minfo) if lmax is None else lmax
        self.ainfo = alm_utils.AlmInfo(lmax)

        observed = mat_utils.atleast_nd(np.asarray(mask_bool).astype(bool, copy=False), 3)
        observed = np.broadcast_to(observed, (ncomp,) + observed.shape[1:])
        self.mask_bool = self._gather(observed)
        self.icov_pix = self._gather(icov_pix)
        if q_low != 0 or q_high != 1:
            self.icov_pix = map_utils.threshold_icov(self.icov_pix, q_low=q_low, q_high=q_high)

        self.icov_ell = signal_icov_ell(theory_cls, ncomp, lmax, include_te=include_te)
        self.b_ell = beam_per_component(b_ell, ncomp, lmax)
        self.sfilt = mat_utils.matpow(self.b_ell, -0.5) if scale_a else None
        self.swap_bm = swap_bm

        self.prec_pinv = preconditioners.PseudoInvPreconditioner(
            self.ainfo, self.icov_ell, self.icov_pix, self.minfo, self.spin, b_ell=self.b_ell,
            sfilt=self.sfilt)
        self.prec_masked_cg = None if no_masked_prec_cg else \
            preconditioners.MaskedPreconditionerCG(
                self.ainfo, self.icov_ell, self.spin, self.mask_bool[0], self.minfo,
                lmax=lmax_prec_cg or lmax, nsteps=15, lmax_r_ell=None, sfilt=self.sfilt)
        # phase 2's multigrid acts on T only
        self.prec_masked_mg = None if no_masked_prec_mg else \
            preconditioners.MaskedPreconditioner(
                self.ainfo, self.icov_ell[:1, :1], 0, self.mask_bool[0], self.minfo,
                min_pix=1000, n_jacobi=1, lmax_r_ell=_LMAX_MG[bool(scale_a)],
                sfilt=None if self.sfilt is None else self.sfilt[:1, :1], device=self.device)

    def _gather(self, arr):
        '''(..., Ny, Nx) map -> (..., npix) in the unit-stride ring order.'''
        arr = np.asarray(arr)
        return np.ascontiguousarray(arr.reshape(arr.shape[:-2] + (-1,))[..., self.index])

    def _phase_preconditioners(self, phase):
        precs = [(self.prec_pinv, None)]
        masked = (self.prec_masked_cg, None) if phase == 1 else (self.prec_masked_mg, np.s_[0])
        if masked[0] is not None:
            precs.append(masked)
        return precs

    def filter(self, imap, niter=None, niter_masked_cg=5, benchmark=False, verbose=True,
               err_tol=1e-15, draw_constr=False, seed=None):
        '''niter_masked_cg phase-1 steps, then niter phase-2 steps (default
        15, with a warning), stopping early once the solver's error drops
        below err_tol (filters.py:169-301). Returns {'walm', 'ialm' (device
        alm), 'solver', 'errors'}; with benchmark = k also 'chisqs',
        'residuals', 'qforms', 'ps' (host spectra of the Wiener alm) and
        'itnums' every k-th step (step 0 holds chisq and residual only).

        'errors' starts with nan, then one entry per step. A phase-2 entry is
        the solver's error (relative to the restart) times entry
        niter_masked_cg - 1 of the list, the reference's scaling kept as it
        is: that entry is the error after the second-to-last phase-1 step
        (nan for niter_masked_cg = 1).'''
        expect = (self.ncomp,) + self.shape_in
        if np.shape(imap) != expect:
            raise ValueError(f'imap has shape {np.shape(imap)}, expected {expect}')
        data = self._gather(np.asarray(imap).astype(self.icov_pix.dtype, copy=False))
        solver = solvers.CGWienerMap.from_arrays(
            data, self.minfo, self.ainfo, self.icov_ell, self.icov_pix, b_ell=self.b_ell,
            draw_constr=draw_constr, mask_pix=self.mask_bool, swap_bm=self.swap_bm,
            spin=self.spin, sfilt=self.sfilt, seed=seed, device=self.device)
        for prec, sel in self._phase_preconditioners(1):
            solver.add_preconditioner(prec, sel=sel)
        solver.init_solver()

        stats = {'chisqs': [], 'residuals': [], 'qforms': [], 'ps': [], 'itnums': []}
        if benchmark:
            warnings.warn('optweight: benchmark statistics are on; each one costs extra '
                          'operator applications and slows the filter down.')
            stats['chisqs'].append(solver.get_chisq())
            stats['residuals'].append(solver.get_residual())
            stats['itnums'].append(0)
            if verbose:
                print(f'|b| = {np.sqrt(solvers._host(solver.dot(solver.b0, solver.b0))):.6e}')
        if niter is None:
            niter = 15
            warnings.warn(f'optweight: niter not given, running niter_masked_cg = '
                          f'{niter_masked_cg} + niter = {niter} steps.')

        errors = [np.nan]
        total = niter_masked_cg + niter
        for idx in range(total):
            phase2 = idx >= niter_masked_cg
            if idx == niter_masked_cg:
                solver.reset_preconditioner()
                for prec, sel in self._phase_preconditioners(2):
                    solver.add_preconditioner(prec, sel=sel)
                solver.b_vec = solver.b0
                solver.init_solver(x0=solver.x)
            t0 = timer()
            solver.step()
            seconds = timer() - t0
            scale = errors[niter_masked_cg - 1] if phase2 and niter_masked_cg > 0 else 1.0
            errors.append(solver.err * scale)
            if benchmark and (idx + 1) % benchmark == 0:
                self._record(solver, stats, idx)
            if verbose:
                print(f'optweight step {idx + 1}/{total} (phase {2 if phase2 else 1}): '
                      f'error {errors[-1]:.2e}, {seconds:.3f} s')
            if solver.err < err_tol:
                warnings.warn(f'optweight: stopped after step {idx + 1}, solver error '
                              f'{solver.err:.3e} < err_tol {err_tol:.3e}')
                break

        out = {'walm': solver.get_wiener(), 'ialm': solver.get_icov(), 'solver': solver,
               'errors': errors}
        if benchmark:
            out.update(stats)
        return out

    def _record(self, solver, stats, idx):
        stats['chisqs'].append(solver.get_chisq())
        stats['residuals'].append(solver.get_residual())
        stats['qforms'].append(solver.get_qform())
        wiener = solver.get_wiener()
        stats['ps'].append(alm2cl(wiener, wiener, self.ainfo))
        stats['itnums'].append(idx)
        print(f'optweight benchmark step {idx + 1}: chisq {stats["chisqs"][-1]:.2f}, '
              f'residual {stats["residuals"][-1]:.2f}, qform {stats["qforms"][-1]:.2f}')
