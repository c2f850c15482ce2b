'''
Scale-discrete wavelet kernels (Leistedt et al. 2013, arXiv:1211.1680,
eqs. 10-16), mirroring optweight's wlm_utils (wlm_utils.py:8-287): the
harmonic kernels w_ell behind the wavelet noise models and, sampled on the
flat-sky wavenumber plane, the 2-D Fourier kernels of fkernel.

Definitions (host, float64):
  Schwarz function  s(t) = exp(-1 / (1 - t^2)) on |t| < 1, else 0;
  s_lambda(t)       = s(2 lambda (t - 1/lambda) / (lambda - 1) - 1);
  k_lambda(t)       = int_t^1 s_lambda^2(u) / u du / int_{1/lambda}^1 (same),
                      1 below 1/lambda and 0 above 1, both integrals by the
                      trapezoid rule on 1000 equally spaced points;
  kappa_lambda(t)   = sqrt(k_lambda(t / lambda) - k_lambda(t)) (negative
                      rounding clipped to 0), eta_lambda = sqrt(k_lambda),
                      zeta_lambda = sqrt(1 - k_lambda);
  psi_j(ell)        = kappa_lambda(ell / lambda^j), phi_j0 = eta_lambda(ell /
                      lambda^j0), omega_J = zeta_lambda(ell / lambda^J),
so that the kernels of one set add up in quadrature to one.
'''
import numpy as np


def schwarz(ts):
    '''exp(-1 / (1 - t^2)) inside (-1, 1), zero outside.'''
    ts = np.asarray(ts, dtype=np.float64)
    out = np.zeros(ts.size)
    inside = (ts > -1) & (ts < 1)
    out[inside] = np.exp(-1.0 / (1.0 - ts[inside] ** 2))
    return out


def s_lambda(ts, lamb):
    if lamb <= 0:
        raise ValueError('lambda parameter needs to be positive.')
    ts = np.asarray(ts, dtype=np.float64)
    return schwarz(2 * lamb * (ts - 1 / lamb) / (lamb - 1) - 1)


def _trapezoid(y, x):
    return float(np.sum((y[1:] + y[:-1]) * np.diff(x)) / 2)


def k_lambda(ts, lamb):
    ts = np.atleast_1d(np.asarray(ts, dtype=np.float64))
    nint = 1000
    out = np.zeros(ts.size)
    u_den = np.linspace(1 / lamb, 1, num=nint, endpoint=True)
    den = _trapezoid(s_lambda(u_den, lamb) ** 2 / u_den, u_den)
    for i, t in enumerate(ts):
        if t < 1 / lamb:
            out[i] = 1
        elif t > 1:
            out[i] = 0
        else:
            u = np.linspace(t, 1, num=nint, endpoint=True)
            out[i] = _trapezoid(s_lambda(u, lamb) ** 2 / u, u) / den
    return out


def kappa_lambda(ts, lamb):
    ts = np.asarray(ts, dtype=np.float64)
    diff = k_lambda(ts / lamb, lamb) - k_lambda(ts, lamb)
    return np.sqrt(np.maximum(diff, 0))


def eta_lambda(ts, lamb):
    return np.sqrt(k_lambda(ts, lamb))


def zeta_lambda(ts, lamb):
    arg = 1 - k_lambda(ts, lamb)
    return np.sqrt(np.where(arg >= 0, arg, 0))


def psi_ell(ells, lamb, j_scale):
    return kappa_lambda(np.asarray(ells) / lamb ** j_scale, lamb)


def phi_ell(ells, lamb, j0_scale):
    return eta_lambda(np.asarray(ells) / lamb ** j0_scale, lamb)


def omega_ell(ells, lamb, j_scale):
    return zeta_lambda(np.asarray(ells) / lamb ** j_scale, lamb)


def j_scale_to_lmax(j_scale, lamb):
    '''Band limit of scale j: ceil(lambda^(j + 1)).'''
    return int(np.ceil(lamb ** (j_scale + 1)))


def j_scale_to_lmin(j_scale, lamb):
    return max(0, int(lamb ** (j_scale - 1)))


def lmax_to_j_scale(lmax, lamb):
    '''Smallest J with lambda^J >= lmax.'''
    if lamb <= 1:
        raise ValueError(f'lambda needs to be > 1, got {lamb}')
    return int(np.ceil(np.log10(lmax) / np.log10(lamb)))


def _scale_range(lamb, lmax, j0, lmin, jmax, lmax_j):
    '''(first, last) wavelet scale of a kernel set. The last scale is the
    one reaching lmax, or lower if jmax (or lmax_j, the band limit that
    sets it) asks; the first is j0, or the scale whose phi ends near lmin,
    else 0. Only nonzero j0 / lmin (jmax / lmax_j) count as both given.'''
    if j0 and lmin:
        raise ValueError(f'give j0 ({j0}) or lmin ({lmin}), not both')
    if jmax and lmax_j:
        raise ValueError(f'give jmax ({jmax}) or lmax_j ({lmax_j}), not both')
    top = lmax_to_j_scale(lmax, lamb)
    if lmax_j:
        jmax = lmax_to_j_scale(lmax_j, lamb)
    last = top if jmax is None else min(jmax, top)
    if j0 is not None:
        first = j0
    elif lmin is not None:
        if lmin < 0:
            raise ValueError(f'lmin must be >= 0, got {lmin}')
        first = max(0, lmax_to_j_scale(lmin, lamb) - 1)
    else:
        first = 0
    if first < 0:
        raise ValueError(f'j0 must be >= 0, got {first}')
    if first > last:
        raise ValueError(f'first scale {first} lies above the last scale {last} that lmax '
                         f'{lmax} (and jmax / lmax_j) allow: lower j0 or lmin, or raise lmax')
    return first, last


def get_sd_kernels(lamb, lmax, j0=None, lmin=None, jmax=None, lmax_j=None, return_j=False):
    '''Scale-discrete kernels w_ell (nwav, lmax + 1) with their band limits
    (and scales, with return_j) for the scales j0..J (wlm_utils.py:8-109):
    the first kernel is phi at scale j0 + 1 (1 from l = 0), the last (for
    two or more scales) omega at J (1 from lambda^(J+1) on), psi_j between.
    Band limits: ceil(lambda^(j+1)) capped at lmax, lmax for the last.'''
    first, last = _scale_range(lamb, lmax, j0, lmin, jmax, lmax_j)
    js = np.arange(first, last + 1)
    ells = np.arange(lmax + 1)
    kernels = [phi_ell(ells, lamb, first + 1)]
    kernels += [psi_ell(ells, lamb, j) for j in js[1:-1]]
    if js.size > 1:
        kernels.append(omega_ell(ells, lamb, last))
    w_ell = np.stack(kernels)
    lmaxs = np.array([min(lmax, j_scale_to_lmax(j, lamb)) for j in js], dtype=int)
    lmaxs[-1] = lmax
    return (w_ell, lmaxs, js) if return_j else (w_ell, lmaxs)


def get_lmax_array(w_ell, eps=1e-10):
    '''Band limit of each kernel (wlm_utils.py:111-142): one past the last
    multipole where |w_ell| > eps, at most lmax, 0 for a kernel that never
    exceeds eps.'''
    above = np.atleast_2d(np.abs(w_ell)) > eps
    nell = above.shape[-1]
    last = nell - 1 - np.argmax(above[:, ::-1], axis=-1)
    return np.where(above.any(axis=-1), np.minimum(last + 1, nell - 1), 0)
