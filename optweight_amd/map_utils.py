'''
Ring geometry and map helpers (host-side setup), mirroring optweight's
map_utils: MapInfo (map_utils.py:19-297), map_info_gauss_legendre (150-185),
get_mtype (299-347), get_fullsky_geometry (349-413), view_2d/view_1d
(415-475), gauss2map (719-785), get_minfo/get_gauss_minfo (819-941),
get_equal_area_minfo (943-992), rand_map_pix (1084-1122), inv_qweight_map
(1124-1170), get_isotropic_ivar (1172-1218), minfo2lmax (1795-1811),
map_info_clenshaw_curtis / map_info_fejer1 (187-257), threshold_icov
(1297-1363), match_enmap_minfo (1490-1591).

The ring weights of the Clenshaw-Curtis and Fejer-1 grids are those that
ducc0's get_gridweights gives optweight (ducc0 >= 0.31; not used here),
computed from Waldvogel's closed forms (BIT Numer. Math. 46, 2006). CAR
geometries are read from any object with astropy's ``wcs.wcs.cdelt / crpix /
crval`` attributes (a pixell ndmap's ``.wcs``, or CarWCS below).

Geometry and setup quantities are numpy on the host; map-space operations
on device tensors go through the HIP library.
'''
import numpy as np
import torch
from scipy.special import roots_legendre
from scipy.interpolate import RectBivariateSpline

from optweight_amd import mat_utils


class MapInfo:
    '''Ring geometry: theta (colatitude), weight (quadrature weight per
    pixel of each ring), nphi, phi0, offsets (first pixel of each ring) and
    stride per ring, with the attribute names of optweight's MapInfo.'''

    def __init__(self, theta, weight, nphi=0, phi0=0, offsets=None, stride=None):
        self.theta = np.array(theta, dtype=np.float64)
        if self.theta.ndim != 1:
            raise ValueError(f'theta array must be 1d, got {self.theta.ndim}')
        self.nrow = self.theta.size
        self.weight = np.array(weight, dtype=np.float64)
        if self.weight.shape != (self.nrow,):
            raise ValueError(f'weight must be ({self.nrow},) got {self.weight.shape}')
        self._nphi = np.array(nphi, dtype=np.uint64)
        if self._nphi.shape not in ((), (self.nrow,)):
            raise ValueError(f'nphi must be 0d or ({self.nrow},) got {self._nphi.shape}')
        if self._nphi.ndim == 0:
            self._nphi = np.full(self.nrow, self._nphi, dtype=self._nphi.dtype)
        self.phi0 = np.array(phi0, dtype=np.float64)
        if self.phi0.shape not in ((), (self.nrow,)):
            raise ValueError(f'phi0 must be 0d or ({self.nrow},) got {self.phi0.shape}')
        if self.phi0.ndim == 0:
            self.phi0 = np.full(self.nrow, self.phi0, dtype=self.phi0.dtype)
        if offsets is None:
            self._offsets = np.concatenate([[0], np.cumsum(self.nphi[:-1])]).astype(np.uint64)
        else:
            self._offsets = np.array(offsets, dtype=np.uint64)
        if self._offsets.shape != (self.nrow,):
            raise ValueError(f'offsets must be ({self.nrow},) got {self._offsets.shape}')
        if stride is None:
            self.stride = np.ones(self.nrow, np.int64)
        else:
            self.stride = np.array(stride, dtype=np.int64)
        if self.stride.shape not in ((), (self.nrow,)):
            raise ValueError(f'stride must be 0d or ({self.nrow},) got {self.stride.shape}')
        if self.stride.ndim == 0:
            self.stride = np.full(self.nrow, self.stride, dtype=self.stride.dtype)
        self.npix = int(np.sum(self.nphi))
        self._device = {}

    @property
    def nphi(self):
        return self._nphi.astype(np.int64)

    @property
    def offsets(self):
        return self._offsets.astype(np.int64)

    def device_arrays(self, device):
        '''weight, nphi, offsets, stride as device tensors (cached).'''
        key = str(device)
        if key not in self._device:
            self._device[key] = (
                torch.as_tensor(self.weight, dtype=torch.float64, device=device),
                torch.as_tensor(self.nphi, dtype=torch.int64, device=device),
                torch.as_tensor(self.offsets, dtype=torch.int64, device=device),
                torch.as_tensor(self.stride, dtype=torch.int64, device=device))
        return self._device[key]

    @classmethod
    def map_info_gauss_legendre(cls, nrings, nphi, phi0=0, stride_lon=None, stride_lat=None):
        x, w = roots_legendre(nrings)
        theta = np.ascontiguousarray(np.arccos(x)[::-1])
        weight = np.ascontiguousarray(w[::-1] * 2 * np.pi)
        weight /= nphi
        offsets = np.arange(theta.size) * stride_lat if stride_lat is not None else None
        return cls(theta, weight, nphi=nphi, phi0=phi0, offsets=offsets, stride=stride_lon)

    @classmethod
    def map_info_clenshaw_curtis(cls, nrings, nphi, phi0=0, stride_lon=None, stride_lat=None):
        theta = np.linspace(0, np.pi, num=nrings, endpoint=True)
        # one ring (a one-row rectangular grid, e.g. a Fourier-wavelet map
        # of the lowest scale): the whole sphere's weight on it
        weight = (get_gridweights('CC', nrings) if nrings > 1 else np.full(1, 4 * np.pi)) / nphi
        offsets = np.arange(theta.size) * stride_lat if stride_lat is not None else None
        return cls(theta, weight, nphi=nphi, phi0=phi0, offsets=offsets, stride=stride_lon)

    @classmethod
    def map_info_fejer1(cls, nrings, nphi, phi0=0, stride_lon=None, stride_lat=None):
        theta = np.linspace(0, np.pi, num=nrings + 1, endpoint=True)[:-1] + np.pi / nrings / 2
        weight = get_gridweights('F1', nrings) / nphi
        offsets = np.arange(theta.size) * stride_lat if stride_lat is not None else None
        return cls(theta, weight, nphi=nphi, phi0=phi0, offsets=offsets, stride=stride_lon)


def get_gridweights(kind, ntheta):
    '''Ring weights summing to 4 pi: 'CC' (Clenshaw-Curtis, ntheta rings at
    theta_k = k pi / (ntheta - 1), both poles included) or 'F1' (Fejer's
    first rule, theta_k = (k + 1/2) pi / ntheta): the rule's weights on
    x = cos(theta) times 2 pi.'''
    k = np.arange(ntheta)
    if kind == 'CC':
        n = ntheta - 1
        if n < 1:
            raise ValueError('Clenshaw-Curtis needs at least 2 rings')
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
    raise ValueError(f'geometry {kind} not supported')


def get_minfo(mtype, nrings, nphi, theta_min=None, theta_max=None):
    '''Rectangular grid of nrings rings of nphi pixels: 'GL', 'CC' or 'F1'
    (map_utils.py:819-883), optionally cut to theta_min <= theta <=
    theta_max.'''
    if mtype == 'GL':
        minfo = MapInfo.map_info_gauss_legendre(nrings, nphi)
    elif mtype == 'CC':
        minfo = MapInfo.map_info_clenshaw_curtis(nrings, nphi)
    elif mtype == 'F1':
        minfo = MapInfo.map_info_fejer1(nrings, nphi)
    else:
        raise ValueError(f"unknown grid {mtype!r}: use 'GL', 'CC' or 'F1'")
    return _theta_cut(minfo, nphi, theta_min, theta_max)


def _theta_cut(minfo, nphi, theta_min, theta_max):
    '''The rings of a rectangular grid with theta_min <= theta <= theta_max
    (an unset or zero bound leaves that side open), stored from offset 0.'''
    if not (theta_min or theta_max):
        return minfo
    lo = theta_min or 0
    hi = theta_max or np.pi
    keep = np.flatnonzero((minfo.theta >= lo) & (minfo.theta <= hi))
    return MapInfo(minfo.theta[keep], minfo.weight[keep], nphi=nphi, phi0=0,
                   offsets=minfo.offsets[keep] - minfo.offsets[keep[0]],
                   stride=minfo.stride[keep])


def fmul_pix(imap, minfo, fmat2d=None, fmat1d=None, ells=None, modlmap=None, inplace=False):
    '''Map (npol, npix) on a rectangular grid to 2-D Fourier coefficients,
    times a per-coefficient matrix, and back (map_utils.py:1751-1793).'''
    from optweight_amd import dft
    if imap.ndim < 2:
        imap = imap.reshape(1, -1)
    imap2d = view_2d(imap, minfo)
    fmap = dft.allocate_fmap(tuple(imap2d.shape), imap2d.dtype, device=imap.device)
    dft.rfft(imap2d, fmap)
    dft.fmul(fmap, fmat2d=fmat2d, fmat1d=fmat1d, ells=ells, modlmap=modlmap, out=fmap)
    omap2d = imap2d if inplace else torch.zeros_like(imap2d)
    dft.irfft(fmap, omap2d)
    return view_1d(omap2d, minfo)


def get_gauss_minfo(lmax, theta_min=None, theta_max=None):
    '''GL grid that supports band limit lmax: floor(lmax/2)+1 rings of
    lmax+1 pixels.'''
    nphi = lmax + 1
    return _theta_cut(MapInfo.map_info_gauss_legendre(lmax // 2 + 1, nphi), nphi, theta_min,
                      theta_max)


def get_mtype(minfo):
    '''Quadrature rule of a geometry's rings: 'GL' when they are unequally
    spaced; for equal spacing 'CC' when the last ring lies a whole number
    of steps from the south pole, 'F1' when half a step more (the fraction
    rounded to one decimal).'''
    theta = np.asarray(minfo.theta)
    if theta.size < 3:
        raise ValueError(f'the rule of a geometry is found from 3 or more rings, got '
                         f'{theta.size}')
    step = np.diff(theta)
    if not np.allclose(step, step[0]):
        return 'GL'
    frac = float(np.round(np.remainder((np.pi - theta.max()) / step[0], 1), 1))
    rule = {0.0: 'CC', 0.5: 'F1'}.get(frac)
    if rule is None:
        raise ValueError(f'equally spaced rings ending {frac} of a step short of a whole '
                         f'number of steps from the pole: neither Clenshaw-Curtis nor Fejer-1')
    return rule


def get_fullsky_geometry(minfo):
    '''The full-sky grid (same rule, same nphi) of which a geometry's rings
    are a consecutive run.'''
    nphi = int(minfo.nphi[0])
    if np.any(minfo.nphi != nphi):
        raise ValueError('the rings differ in nphi: no rectangular full-sky grid holds them')
    rule = get_mtype(minfo)
    step = np.abs(np.diff(minfo.theta))
    if rule == 'GL':
        # GL rings lie about pi / (n + 1/2) apart
        full = MapInfo.map_info_gauss_legendre(int(np.round(np.min(np.pi / step - 0.5))), nphi)
    elif rule == 'CC':
        full = MapInfo.map_info_clenshaw_curtis(int(np.round(np.pi / step[0] + 1)), nphi)
    else:
        full = MapInfo.map_info_fejer1(int(np.round(np.pi / step[0])), nphi)
    mine = np.sort(minfo.theta)
    theirs = np.sort(full.theta)
    first = np.flatnonzero(np.isclose(theirs, mine[0]))
    run = theirs[first[0]:first[0] + mine.size] if first.size == 1 else None
    if run is None or run.size != mine.size or not np.allclose(run, mine):
        raise ValueError(f'the rings are not a consecutive run of the full-sky {rule} grid of '
                         f'{full.nrow} rings')
    return full


def get_equal_area_minfo(minfo, orig_band=np.pi / 8, ratio_pow=1.):
    '''The same rings with nphi scaled by (w_ring / w_equator)^ratio_pow
    (rounded, at least 2; weights rescaled to keep each ring's area)
    outside the band pi/2 - orig_band < theta < pi/2 + orig_band, whose rings
    keep theirs. w_equator is the weight of the first ring at or south of
    the equator; some ring must lie within two ring spacings of it.'''
    if orig_band < 0:
        raise ValueError(f'orig_band is a half width, >= 0; got {orig_band}')
    theta = np.asarray(minfo.theta)
    spacing = np.min(np.abs(np.diff(theta)))
    if not np.any(np.abs(theta - np.pi / 2) < 2 * spacing):
        raise ValueError('the geometry has no ring near the equator to take the pixel '
                         'area from')
    w_eq = minfo.weight[np.searchsorted(theta, np.pi / 2, side='left')]
    in_band = ((np.pi / 2 - orig_band) < theta) & (theta < (np.pi / 2 + orig_band))
    scale = np.where(in_band, 1.0, (minfo.weight / w_eq) ** ratio_pow)
    nphi = np.maximum(2, np.round(minfo.nphi * scale).astype(minfo.nphi.dtype))
    return MapInfo(theta, nphi=nphi, weight=minfo.weight * minfo.nphi / nphi)


class CarWCS:
    '''The attributes of an astropy WCS that a cylindrical CAR geometry
    needs (``.wcs.cdelt``, ``.wcs.crpix``, ``.wcs.crval``, degrees and FITS
    1-based reference pixel), for callers without astropy.'''

    class _Params:
        def __init__(self, cdelt, crpix, crval):
            self.cdelt = np.array(cdelt, dtype=np.float64)
            self.crpix = np.array(crpix, dtype=np.float64)
            self.crval = np.array(crval, dtype=np.float64)

    def __init__(self, cdelt, crpix, crval=(0., 0.)):
        self.wcs = CarWCS._Params(cdelt, crpix, crval)


def car_pix2sky(wcs, y, x):
    '''(dec, ra) in radians of pixels (y, x) of a cylindrical CAR map
    (crval[1] = 0): ra = crval0 + (x + 1 - crpix0) cdelt0 and
    dec = crval1 + (y + 1 - crpix1) cdelt1 in degrees.'''
    w = wcs.wcs
    if abs(float(w.crval[1])) > 1e-12:
        raise NotImplementedError('only cylindrical CAR (crval[1] = 0) is supported')
    ra = w.crval[0] + (np.asarray(x, dtype=np.float64) + 1 - w.crpix[0]) * w.cdelt[0]
    dec = w.crval[1] + (np.asarray(y, dtype=np.float64) + 1 - w.crpix[1]) * w.cdelt[1]
    return np.radians(dec), np.radians(ra)


def match_enmap_minfo(shape, wcs, mtype='CC'):
    '''MapInfo of a CAR map of shape (..., ny, nx) without moving its
    pixels: ring r is map row r, its pixel j sits at offsets[r] + j stride[r]
    of the flattened map (stride -1 when RA decreases with x) at longitude
    phi0 + 2 pi j / nx. Raises ValueError if the map does not span the full
    RA range or its rows are not rings of the CC / fejer1 grid.'''
    if mtype not in ('CC', 'fejer1'):
        raise NotImplementedError(f"CAR maps come as 'CC' or 'fejer1', not {mtype!r}")
    yo = 1 if mtype == 'CC' else 0
    cdelt = wcs.wcs.cdelt
    nx = int(np.round(np.abs(360 / cdelt[0])))
    if nx != shape[-1]:
        raise ValueError(f'the map has {shape[-1]} columns; a full circle at this cdelt '
                         f'takes {nx}')
    ny = int(np.round(abs(180 / cdelt[1]) + yo))
    phi0 = float(car_pix2sky(wcs, 0, 0)[1])
    if mtype == 'CC':
        full = MapInfo.map_info_clenshaw_curtis(ny, nphi=nx)
    else:
        full = MapInfo.map_info_fejer1(ny, nphi=nx)
    ntheta = shape[-2]
    theta = np.pi / 2 - car_pix2sky(wcs, np.arange(ntheta), np.zeros(ntheta))[0]
    atol = 1e-6
    i0 = np.flatnonzero(np.abs(theta[0] - full.theta) < atol)
    i1 = np.flatnonzero(np.abs(theta[-1] - full.theta) < atol)
    if i0.size != 1 or i1.size != 1:
        raise ValueError(f'the first and last rows of the map are not rings of the {mtype} '
                         f'grid of {ny} rings (tolerance {atol} rad)')
    keep = np.s_[min(i0[0], i1[0]):max(i0[0], i1[0]) + 1]
    theta_k = full.theta[keep]
    weight_k = full.weight[keep]
    offsets = np.arange(theta_k.size) * nx
    if cdelt[0] < 0:
        offsets = offsets + nx - 1
        phi0 -= np.radians(cdelt[0])
    stride = np.full(theta_k.size, 1 if cdelt[0] > 0 else -1)
    if cdelt[1] > 0:
        theta_k = np.ascontiguousarray(theta_k[::-1])
        weight_k = np.ascontiguousarray(weight_k[::-1])
    return MapInfo(theta_k, weight_k, nphi=nx, phi0=phi0, offsets=offsets, stride=stride)


def unit_stride(minfo):
    '''(minfo_c, index): the same rings, in the same order, stored one after
    another with stride 1, and the int64 index with map_c = map[..., index]
    (map_c[..., index] = map_c scatters back). The transforms run on
    unit-stride geometries.'''
    nphi = minfo.nphi
    off_c = np.concatenate([[0], np.cumsum(nphi[:-1])]).astype(np.int64)
    index = np.concatenate([minfo.offsets[r] + minfo.stride[r] * np.arange(nphi[r])
                            for r in range(minfo.nrow)]).astype(np.int64)
    return MapInfo(minfo.theta, minfo.weight, nphi=nphi, phi0=minfo.phi0, offsets=off_c), index


def threshold_icov(icov_pix, q_low=0, q_high=1):
    '''Copy of icov_pix (npol, npix) or (npol, npol, npix) whose diagonal is
    clipped per component to its [q_low, q_high] quantiles over the pixels
    where it is positive (q_low / q_high scalars or one per component); a
    matrix is rescaled by the square root of each pixel's clipping factor on
    both sides, so it stays symmetric.'''
    icov_pix = np.asarray(icov_pix)
    if icov_pix.ndim not in (2, 3):
        raise ValueError(f'icov_pix of {icov_pix.ndim} dimensions: expected (npol, npix) or '
                         f'(npol, npol, npix)')
    npol = icov_pix.shape[0]
    diags = np.einsum('iip->ip', icov_pix) if icov_pix.ndim == 3 else icov_pix
    bounds = np.broadcast_to(np.stack(np.broadcast_arrays(q_low, q_high), -1), (npol, 2))
    factor = np.ones(diags.shape)
    for comp, (diag, qs) in enumerate(zip(diags, bounds)):
        positive = diag > 0
        if not positive.any():
            continue
        lo, hi = np.quantile(diag[positive], qs[0]), np.quantile(diag[positive], qs[1])
        clip = np.clip(diag, lo, hi)
        moved = positive & (clip != diag)
        factor[comp, moved] = clip[moved] / diag[moved]
    if icov_pix.ndim == 2:
        out = icov_pix * factor
    else:
        root = np.sqrt(factor)
        out = root[:, None] * icov_pix * root[None, :]
    return out.astype(icov_pix.dtype, copy=False)


def minfo2lmax(minfo):
    return int(0.5 * (np.max(minfo.nphi) - 1))


def ring_support(mask, minfo):
    '''Sorted indices of the rings of minfo on which ``mask`` (..., npix)
    is nonzero for some pixel and leading index (int64 array).'''
    m = np.asarray(mask).reshape(-1, minfo.npix) != 0
    anyp = m.any(axis=0)
    off = np.asarray(minfo.offsets, dtype=np.int64)
    nphi = np.asarray(minfo.nphi, dtype=np.int64)
    hit = np.array([anyp[o:o + n].any() for o, n in zip(off, nphi)], dtype=bool)
    return np.nonzero(hit)[0].astype(np.int64)


def compact_rings(minfo, rings):
    '''The geometry of the rings ``rings`` (sorted indices) of minfo alone,
    stored one after another in ring order, and for each of its pixels the
    index of the same pixel in a map of minfo: (MapInfo, int64 array).'''
    rings = np.asarray(rings, dtype=np.int64)
    nphi = np.asarray(minfo.nphi, dtype=np.int64)[rings]
    off = np.asarray(minfo.offsets, dtype=np.int64)[rings]
    sub = MapInfo(np.asarray(minfo.theta)[rings], np.asarray(minfo.weight)[rings], nphi=nphi,
                  phi0=np.asarray(minfo.phi0)[rings])
    pix = (np.concatenate([o + np.arange(n) for o, n in zip(off, nphi)]) if rings.size
           else np.zeros(0, dtype=np.int64))
    return sub, pix.astype(np.int64)


def view_2d(imap, minfo):
    nphi = minfo.nphi[0]
    if not np.all(minfo.nphi == nphi):
        raise ValueError('view_2d needs constant nphi')
    return imap.reshape(imap.shape[:-1] + (minfo.nrow, nphi))


def view_1d(imap, minfo):
    return imap.reshape(imap.shape[:-2] + (minfo.npix,))


def _gauss_coords(minfo):
    if not np.all(minfo.nphi == minfo.nphi[0]):
        raise ValueError('Input map_info nphi is not the same for all rings')
    nphi = minfo.nphi[0]
    phi0 = minfo.phi0[0]
    phis = np.linspace(phi0, phi0 + 2 * np.pi * minfo.stride[0], nphi, endpoint=False)
    return minfo.theta, np.mod(phis, 2 * np.pi)


def gauss2map(imap, minfo_in, minfo_out, order=3):
    '''Spline interpolation (scipy RectBivariateSpline) of a constant-nphi
    GL map onto any ring geometry.'''
    thetas, phis = _gauss_coords(minfo_in)
    tidx = np.argsort(thetas)
    pidx = np.argsort(phis)
    omap = np.zeros(imap.shape[:-1] + (minfo_out.npix,), dtype=imap.dtype)
    im2 = view_2d(imap, minfo_in)
    for index in np.ndindex(imap.shape[:-1]):
        m2 = im2[index][tidx, :][:, pidx]
        rbs = RectBivariateSpline(thetas[tidx], phis[pidx], m2, kx=order, ky=order)
        for t in range(minfo_out.nrow):
            n = minfo_out.nphi[t]
            p0 = minfo_out.phi0[t]
            ph = np.linspace(p0, 2 * np.pi + p0, n, endpoint=False)
            sl = slice(minfo_out.offsets[t], minfo_out.offsets[t] + n)
            omap[index][sl] = rbs(minfo_out.theta[t], ph)
    return omap


def gauss2gauss(imap, minfo_in, minfo_out, order=3, area_pow=0):
    '''Spline interpolation between constant-nphi Gauss-Legendre maps
    (map_utils.py:672-717): values scaled by (pixel area)^area_pow via the
    ring weights on both sides.'''
    imap = np.asarray(imap)
    thetas_in, phis_in = _gauss_coords(minfo_in)
    thetas_out, phis_out = _gauss_coords(minfo_out)
    im2 = view_2d(imap, minfo_in)
    if area_pow != 0:
        im2 = im2 * (np.asarray(minfo_in.weight) ** -float(area_pow))[:, np.newaxis]
    om2 = np.zeros(imap.shape[:-1] + (minfo_out.nrow, int(minfo_out.nphi[0])), dtype=imap.dtype)
    for idx in np.ndindex(im2.shape[:-2]):
        rbs = RectBivariateSpline(thetas_in, phis_in, im2[idx], kx=order, ky=order)
        om2[idx] = rbs(thetas_out, phis_out)
    if area_pow != 0:
        om2 *= (np.asarray(minfo_out.weight) ** float(area_pow))[:, np.newaxis]
    return view_1d(om2, minfo_out)


def get_ivar_ell(icov_wav, w_ell, mask=None, minfo_mask=None):
    '''Inverse-variance spectrum of a wavelet-diagonal inverse covariance
    (map_utils.py:1220-1266): per scale the isotropic inverse variance of
    block (j, j) (masked by the mask interpolated onto that scale's grid),
    weighted by w_ell[j]^2. Diagonal only, (npol, npol, nell).'''
    from optweight_amd import wavtrans
    npol = wavtrans.preshape2npol(icov_wav.preshape)
    w_ell = np.asarray(w_ell)
    itaus = np.zeros((npol, npol, w_ell.shape[0]))
    for jidx in range(w_ell.shape[0]):
        minfo_j = icov_wav.minfos[(jidx, jidx)]
        m = icov_wav.maps[(jidx, jidx)]
        m = m.detach().cpu().numpy() if hasattr(m, 'detach') else np.asarray(m)
        mask_j = None
        if mask is not None:
            mask_j = gauss2gauss(np.asarray(mask), minfo_mask, minfo_j, order=1, area_pow=0)
        itaus[:, :, jidx] = get_isotropic_ivar(m, minfo_j, mask=mask_j)
    return np.einsum('ijk, kl -> ijl', itaus, w_ell ** 2, optimize=True)


def _ringweight_np(imap, minfo, weight):
    for t in range(minfo.nrow):
        sl = slice(minfo.offsets[t], minfo.offsets[t] + minfo.nphi[t] * minfo.stride[t],
                   minfo.stride[t])
        imap[..., sl] *= imap.dtype.type(weight[t])
    return imap


def inv_qweight_map(imap, minfo, inplace=False, qweight=False):
    '''W^-1 m (or W m with qweight=True) for host arrays.'''
    out = imap if inplace else np.array(imap, copy=True)
    w = minfo.weight if qweight else 1 / minfo.weight
    return _ringweight_np(out, minfo, w)


def get_isotropic_ivar(icov_pix, minfo, mask=None):
    '''Isotropic inverse variance tau_ii = sum (icov_ii/w)^2 / sum (icov_ii/w)
    (Seljebotn et al. 2017, eq. A1); diagonal only.'''
    icov_pix = np.asarray(icov_pix)
    wcov = inv_qweight_map(icov_pix, minfo, inplace=False)
    if mask is not None:
        if wcov.ndim == 2:
            wcov = wcov * mask
        else:
            mask = mat_utils.atleast_nd(mask, 2)
            sm = np.sqrt(mask)
            wcov = wcov * np.einsum('ap, bp -> abp', sm, sm)
    num = np.sum(wcov ** 2, axis=-1)
    den = np.sum(wcov, axis=-1)
    if icov_pix.ndim == 2:
        return np.diag(num / den)
    return np.diag(np.diag(num) / np.diag(den))


def rand_map_pix(cov_pix, seed):
    '''Gaussian draw sqrt(cov) * N(0, 1) (host, numpy default_rng).'''
    cov_pix = np.asarray(cov_pix)
    npol, npix = cov_pix.shape[-2:]
    rng = np.random.default_rng(seed)
    uv = rng.standard_normal(npol * npix).astype(cov_pix.dtype, copy=False)
    uv = uv.reshape((npol, npix))
    if cov_pix.ndim == 2:
        return uv * np.sqrt(cov_pix)
    sq = mat_utils.matpow(cov_pix, 0.5)
    return np.einsum('ijk, jk -> ik', sq, uv, optimize=True)


# ------------------------------------------------------------ HDF5 files
# Datasets and attributes of the reference's files (map_utils.py:1902-2101):
# a map info is the datasets theta, nphi, phi0, offsets, stride, weight; a
# map file holds dataset 'map' and group 'minfo', and for maps stored as
# upper-triangular elements the attributes 'symm_axis' and 'shape_orig'.
# The append_* / *_from_hdf functions take any h5py-like group; opening
# files needs h5py.

def _h5py():
    try:
        import h5py
    except ImportError as exc:
        raise ImportError('reading and writing map files needs h5py') from exc
    return h5py


def _host(a):
    return a.detach().cpu().numpy() if isinstance(a, torch.Tensor) else np.asarray(a)


def _with_ext(fname):
    import os
    return fname if os.path.splitext(fname)[1] else fname + '.hdf5'


def append_minfo_to_hdf(hfile, minfo):
    for name in ('theta', 'nphi', 'phi0', 'offsets', 'stride', 'weight'):
        hfile.create_dataset(name, data=np.asarray(getattr(minfo, name)))


def minfo_from_hdf(hfile):
    return MapInfo(theta=hfile['theta'][()], nphi=hfile['nphi'][()], phi0=hfile['phi0'][()],
                   offsets=hfile['offsets'][()], stride=hfile['stride'][()],
                   weight=hfile['weight'][()])


def write_minfo(fname, minfo):
    with _h5py().File(_with_ext(fname), 'w') as hfile:
        append_minfo_to_hdf(hfile, minfo)


def read_minfo(fname):
    with _h5py().File(fname, 'r') as hfile:
        return minfo_from_hdf(hfile)


def _is_seq_of_seq(obj):
    try:
        iter(obj)
    except TypeError:
        return False
    for item in obj:
        try:
            iter(item)
        except TypeError:
            return False
    return True


def append_map_to_hdf(hfile, imap, minfo, symm_axes=None):
    """Dataset 'map' (only the upper-triangular elements when the map is
    symmetric in symm_axes, after merging axis groups if symm_axes is a
    sequence of sequences) and group 'minfo'."""
    imap = _host(imap)
    if symm_axes is not None:
        shape_orig = imap.shape
        if _is_seq_of_seq(symm_axes):
            imap, symm_axes = mat_utils.flattened_view(imap, [list(a) for a in symm_axes],
                                                       return_flat_axes=True)
        imap = mat_utils.symm2triu(imap, axes=symm_axes)
        hfile.attrs['symm_axis'] = symm_axes[0]
        hfile.attrs['shape_orig'] = shape_orig
    hfile.create_dataset('map', data=imap)
    append_minfo_to_hdf(hfile.create_group('minfo'), minfo)


def map_from_hdf(hfile):
    omap = hfile['map'][()]
    symm_axis = hfile.attrs.get('symm_axis', None)
    if symm_axis is not None:
        omap = mat_utils.triu2symm(omap, int(symm_axis))
    shape_orig = hfile.attrs.get('shape_orig', None)
    if shape_orig is not None:
        omap = omap.reshape(tuple(int(n) for n in shape_orig))
    return omap, minfo_from_hdf(hfile['minfo'])


def write_map(fname, imap, minfo, symm_axes=None):
    with _h5py().File(_with_ext(fname), 'w') as hfile:
        append_map_to_hdf(hfile, imap, minfo, symm_axes=symm_axes)


def read_map(fname):
    with _h5py().File(fname, 'r') as hfile:
        return map_from_hdf(hfile)
