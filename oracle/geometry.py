'''
Ring geometry (oracle). Restates map_utils.MapInfo (map_utils.py:19-297),
map_info_gauss_legendre (150-185), get_mtype (299-347), get_fullsky_geometry
(349-413), get_minfo/get_gauss_minfo (819-941), get_equal_area_minfo
(943-992), gauss2map (719-785), minfo2lmax (1795-1811),
map_info_clenshaw_curtis / map_info_fejer1 (187-257), match_enmap_minfo
(1490-1591), with the CAR pixel-to-sky relation of pixell's enmap.pix2sky
for cylindrical plate carree (pixell is not vendored) and the Clenshaw-Curtis
and Fejer-1 ring weights that ducc0 (>= 0.31, not importable here) computes
with get_gridweights, from Waldvogel's closed forms (BIT Numer. Math. 46,
2006).
'''
import numpy as np
from scipy.special import roots_legendre
from scipy.interpolate import RectBivariateSpline


class Geom:
    '''theta (colatitude, ascending), weight (per pixel), nphi, phi0,
    offsets (first pixel of each ring), stride (always 1 here).'''

    def __init__(self, theta, weight, nphi, phi0=0.0, offsets=None):
        self.theta = np.asarray(theta, dtype=np.float64)
        self.nrow = self.theta.size
        self.weight = np.asarray(weight, dtype=np.float64).reshape(self.nrow)
        nphi = np.asarray(nphi, dtype=np.int64)
        self.nphi = np.full(self.nrow, nphi) if nphi.ndim == 0 else nphi.copy()
        phi0 = np.asarray(phi0, dtype=np.float64)
        self.phi0 = np.full(self.nrow, phi0) if phi0.ndim == 0 else phi0.copy()
        if offsets is None:
            offsets = np.concatenate([[0], np.cumsum(self.nphi[:-1])])
        self.offsets = np.asarray(offsets, dtype=np.int64)
        self.npix = int(np.sum(self.nphi))


def mirror_pairs(geom, tol=1e-12):
    '''Pair ring i (theta < pi/2) with ring j when |theta_i + theta_j - pi| <
    tol; returns (north, south) index arrays, south = -1 for unpaired rings.
    The equator ring of an odd grid is its own mirror and stays unpaired.'''
    th = geom.theta
    used = np.zeros(geom.nrow, dtype=bool)
    north, south = [], []
    order = np.argsort(th, kind='stable')
    for i in order:
        if used[i]:
            continue
        used[i] = True
        j = -1
        if th[i] < np.pi / 2:
            cand = np.nonzero((np.abs(th + th[i] - np.pi) < tol) & ~used)[0]
            if cand.size:
                j = int(cand[0])
                used[j] = True
        north.append(int(i))
        south.append(j)
    return np.asarray(north), np.asarray(south)


def ring_x(geom):
    '''Per-ring (theta_eff, x) in longdouble as the transforms evaluate them:
    x of a north ring is cos(theta) rounded to float64; its mirror ring uses
    exactly -x and theta_eff = pi - theta_north.'''
    LD = np.longdouble
    north, south = mirror_pairs(geom)
    th = np.asarray(geom.theta, dtype=LD)
    x = np.cos(th).astype(np.float64).astype(LD)
    th_eff = th.copy()
    for n, s in zip(north, south):
        if s >= 0:
            x[s] = -x[n]
            th_eff[s] = LD(np.pi) - th[n]
    return th_eff, x


def gauss_legendre(nrings, nphi, phi0=0.0):
    '''map_info_gauss_legendre: theta = arccos(x) ascending,
    weight = w_gl * 2 pi / nphi.'''
    x, w = roots_legendre(nrings)
    theta = np.ascontiguousarray(np.arccos(x)[::-1])
    weight = np.ascontiguousarray(w[::-1] * 2 * np.pi) / nphi
    return Geom(theta, weight, nphi, phi0=phi0)


def gauss_minfo(lmax):
    '''get_gauss_minfo(lmax): floor(lmax/2)+1 rings, lmax+1 pixels.'''
    return gauss_legendre(lmax // 2 + 1, lmax + 1)


def minfo2lmax(geom):
    return int(0.5 * (np.max(geom.nphi) - 1))


def get_mtype(geom):
    if geom.theta.size < 3:
        raise ValueError('need at least 3 rings')
    diff = np.diff(geom.theta)
    if not np.allclose(diff, diff[0]):
        return 'GL'
    dtheta = diff[0]
    remain = np.round(np.remainder((np.pi - geom.theta.max()) / dtheta, 1), 1)
    if remain == 0.0:
        return 'CC'
    if remain == 0.5:
        return 'F1'
    raise ValueError('cannot determine mtype')


def fullsky_geometry(geom):
    '''get_fullsky_geometry for GL, CC and F1 inputs.'''
    nphi = geom.nphi[0]
    if not np.all(geom.nphi == nphi):
        raise ValueError('varying nphi not supported')
    mtype = get_mtype(geom)
    diff = np.diff(geom.theta)
    if mtype == 'GL':
        return gauss_legendre(int(np.round(np.min(np.pi / np.abs(diff) - 0.5))), nphi)
    if mtype == 'CC':
        return clenshaw_curtis(int(np.round(np.pi / np.abs(diff[0]) + 1)), nphi)
    return fejer1(int(np.round(np.pi / np.abs(diff[0]))), nphi)


def equal_area(geom, orig_band=np.pi / 8, ratio_pow=1.0):
    '''get_equal_area_minfo: nphi_r = max(2, round(nphi * (w_r/w_eq)^pow))
    outside |theta - pi/2| < orig_band; weight_r = w_r nphi / nphi_r.'''
    delta_theta = np.min(np.abs(np.diff(geom.theta)))
    if not np.any(np.abs(geom.theta - np.pi / 2) < 2 * delta_theta):
        raise ValueError('geometry must include equator')
    idx_ref = np.searchsorted(geom.theta, np.pi / 2, side='left')
    area_ref = geom.weight[idx_ref]
    nphi_r = np.zeros(geom.nrow, dtype=np.int64)
    w_r = np.zeros(geom.nrow)
    for t in range(geom.nrow):
        area = geom.weight[t]
        if (np.pi / 2 - orig_band) < geom.theta[t] < (np.pi / 2 + orig_band):
            ratio = 1
        else:
            ratio = (area / area_ref) ** ratio_pow
        nphi_r[t] = max(2, int(np.round(geom.nphi[t] * ratio)))
        w_r[t] = area * geom.nphi[t] / nphi_r[t]
    return Geom(geom.theta, w_r, nphi_r)


def view_2d(imap, geom):
    nphi = geom.nphi[0]
    return imap.reshape(imap.shape[:-1] + (geom.nrow, nphi))


def gauss2map(imap, geom_in, geom_out, order=3):
    '''Spline interpolation of a constant-nphi GL map onto a ring geometry
    (gauss2map with area_pow=0).'''
    thetas = geom_in.theta
    nphi = geom_in.nphi[0]
    phis = np.mod(np.linspace(geom_in.phi0[0], geom_in.phi0[0] + 2 * np.pi,
                              nphi, endpoint=False), 2 * np.pi)
    tidx = np.argsort(thetas)
    pidx = np.argsort(phis)
    omap = np.zeros(imap.shape[:-1] + (geom_out.npix,), dtype=imap.dtype)
    imap2 = view_2d(imap, geom_in)
    for index in np.ndindex(imap.shape[:-1]):
        m2 = imap2[index][tidx, :][:, pidx]
        rbs = RectBivariateSpline(thetas[tidx], phis[pidx], m2, kx=order,
                                  ky=order)
        for t in range(geom_out.nrow):
            n = geom_out.nphi[t]
            p0 = geom_out.phi0[t]
            ph = np.linspace(p0, 2 * np.pi + p0, n, endpoint=False)
            sl = slice(geom_out.offsets[t], geom_out.offsets[t] + n)
            omap[index][sl] = rbs(geom_out.theta[t], ph)
    return omap


# ----------------------------------------------------------- CAR grids

def gridweights(kind, ntheta):
    '''Ring weights summing to 4 pi: 'CC' (Clenshaw-Curtis, rings at
    theta_k = k pi / (ntheta - 1)) or 'F1' (Fejer's first rule, theta_k =
    (k + 1/2) pi / ntheta). Weights on x = cos(theta) times 2 pi.'''
    k = np.arange(ntheta)
    if kind == 'CC':
        n = ntheta - 1
        if n < 1:
            raise ValueError('CC needs at least 2 rings')
        w = np.ones(ntheta)
        for j in range(1, n // 2 + 1):
            b = 1.0 if 2 * j == n else 2.0
            w -= b / (4.0 * j * j - 1) * np.cos(2 * j * k * np.pi / n)
        c = np.full(ntheta, 2.0)
        c[0] = c[-1] = 1.0
        return 2 * np.pi * c / n * w
    if kind == 'F1':
        th = (2 * k + 1) * np.pi / (2 * ntheta)
        w = np.ones(ntheta)
        for j in range(1, ntheta // 2 + 1):
            w -= 2 * np.cos(2 * j * th) / (4.0 * j * j - 1)
        return 2 * np.pi * 2.0 / ntheta * w
    raise ValueError(f'unknown quadrature {kind}')


def clenshaw_curtis(nrings, nphi, phi0=0.0):
    theta = np.linspace(0, np.pi, num=nrings, endpoint=True)
    return Geom(theta, gridweights('CC', nrings) / nphi, nphi, phi0=phi0)


def fejer1(nrings, nphi, phi0=0.0):
    theta = np.linspace(0, np.pi, num=nrings + 1, endpoint=True)[:-1] + np.pi / nrings / 2
    return Geom(theta, gridweights('F1', nrings) / nphi, nphi, phi0=phi0)


def car_pix2sky(wcs, y, x):
    '''(dec, ra) in radians of pixel (y, x) of a cylindrical CAR map
    (reference value crval[1] = 0): ra = crval0 + (x + 1 - crpix0) cdelt0,
    dec = crval1 + (y + 1 - crpix1) cdelt1 (degrees, FITS 1-based crpix).'''
    w = wcs.wcs
    ra = w.crval[0] + (np.asarray(x, dtype=np.float64) + 1 - w.crpix[0]) * w.cdelt[0]
    dec = w.crval[1] + (np.asarray(y, dtype=np.float64) + 1 - w.crpix[1]) * w.cdelt[1]
    return np.radians(dec), np.radians(ra)


def match_car(shape, wcs, mtype='CC'):
    '''match_enmap_minfo: the ring geometry of a CAR map of shape (..., ny,
    nx) in its own row order. Returns (theta, weight, nphi, phi0, offsets,
    stride) with the reference's meaning (pixel j of ring r stored at
    offsets[r] + j stride[r] of the flattened map, at longitude
    phi0 + 2 pi j / nphi), and the equivalent unit-stride Geom plus the index
    that gathers a flattened map into it (canonical = flat[..., index]).'''
    if mtype not in ('CC', 'fejer1'):
        raise NotImplementedError('Only "CC" or "fejer1" `mtype` implemented for now.')
    yo = 1 if mtype == 'CC' else 0
    cd = wcs.wcs.cdelt
    nx = int(np.round(np.abs(360 / cd[0])))
    if nx != shape[-1]:
        raise ValueError(f'Input needs to span full range in phi, {shape[-1]}, {nx} ')
    ny = int(np.round(abs(180 / cd[1]) + yo))
    phi0 = float(car_pix2sky(wcs, 0, 0)[1])
    full = clenshaw_curtis(ny, nx) if mtype == 'CC' else fejer1(ny, nx)
    ntheta = shape[-2]
    theta = np.pi / 2 - car_pix2sky(wcs, np.arange(ntheta), np.zeros(ntheta))[0]
    i0 = np.flatnonzero(np.abs(theta[0] - full.theta) < 1e-6)
    i1 = np.flatnonzero(np.abs(theta[-1] - full.theta) < 1e-6)
    if i0.size != 1 or i1.size != 1:
        raise ValueError(f'Input geometry is not {mtype} at atol = 1e-6')
    i0, i1 = i0[0], i1[0]
    keep = slice(min(i0, i1), max(i0, i1) + 1)
    th = full.theta[keep]
    wt = full.weight[keep]
    offsets = np.arange(th.size) * nx
    stride = 1 if cd[0] > 0 else -1
    if cd[0] < 0:
        offsets = offsets + nx - 1
        phi0 -= np.radians(cd[0])
    if cd[1] > 0:
        th = th[::-1].copy()
        wt = wt[::-1].copy()
    geom = Geom(th, wt, nx, phi0=phi0)
    j = np.arange(nx)
    index = np.concatenate([offsets[r] + stride * j for r in range(th.size)])
    return (th, wt, nx, phi0, offsets, np.full(th.size, stride)), geom, index
