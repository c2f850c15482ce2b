'''
Ring geometry and map helpers (host-side setup), mirroring optweight's
map_utils: MapInfo (map_utils.py:19-297), map_info_gauss_legendre (150-185),
get_mtype (299-347), get_fullsky_geometry (349-413), view_2d/view_1d
(415-475), gauss2map (719-785), get_minfo/get_gauss_minfo (819-941),
get_equal_area_minfo (943-992), rand_map_pix (1084-1122), inv_qweight_map
(1124-1170), get_isotropic_ivar (1172-1218), minfo2lmax (1795-1811).

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


def get_gauss_minfo(lmax, theta_min=None, theta_max=None):
    '''GL grid that supports band limit lmax: floor(lmax/2)+1 rings of
    lmax+1 pixels.'''
    nrings = int(np.floor(lmax / 2)) + 1
    nphi = lmax + 1
    minfo = MapInfo.map_info_gauss_legendre(nrings, nphi)
    if theta_min or theta_max:
        theta_min = theta_min if theta_min else 0
        theta_max = theta_max if theta_max else np.pi
        mask = (minfo.theta >= theta_min) & (minfo.theta <= theta_max)
        offsets = minfo.offsets[mask]
        offsets = offsets - offsets[0]
        minfo = MapInfo(minfo.theta[mask], nphi=nphi, phi0=0, offsets=offsets,
                        stride=minfo.stride[mask], weight=minfo.weight[mask])
    return minfo


def get_mtype(minfo):
    if minfo.theta.size < 3:
        raise ValueError(f'Need atleast 3 rings, got {minfo.theta.size=}')
    diff = np.diff(minfo.theta)
    if not np.allclose(diff, diff[0]):
        return 'GL'
    dtheta = diff[0]
    remain = np.round(np.remainder((np.pi - minfo.theta.max()) / dtheta, 1), 1)
    if remain == 0.0:
        return 'CC'
    if remain == 0.5:
        return 'F1'
    raise ValueError(f'Cannot determine mtype, {remain=}')


def get_fullsky_geometry(minfo):
    '''Full-sky GL geometry sharing the rings of a (cut-sky) GL geometry.'''
    nphi = minfo.nphi[0]
    if not all(minfo.nphi == nphi):
        raise ValueError('Geometries with varying nphi not supported')
    mtype = get_mtype(minfo)
    if mtype != 'GL':
        raise ValueError(f'{mtype=} not supported')
    diff = np.diff(minfo.theta)
    nrings = int(np.round(np.min(np.pi / np.abs(diff) - 0.5)))
    out = MapInfo.map_info_gauss_legendre(nrings, nphi)
    ts = np.sort(minfo.theta)
    tf = np.sort(out.theta)
    idx = np.argwhere(np.isclose(tf, ts[0]))
    if idx.size != 1:
        raise ValueError('theta_sort[0] of input not in fullsky output')
    idx = idx[0][0]
    if not np.allclose(tf[idx:idx + minfo.nrow], ts):
        raise ValueError('theta input does not overlap with fullsky output')
    return out


def get_equal_area_minfo(minfo, orig_band=np.pi / 8, ratio_pow=1.):
    '''Reduce nphi per ring to nphi * (w_ring / w_equator)^ratio_pow outside
    the band |theta - pi/2| < orig_band (at least 2 pixels).'''
    assert orig_band >= 0, f'orig_band must be >= 0, got {orig_band}'
    delta_theta = np.min(np.abs(np.diff(minfo.theta)))
    if not np.any(np.abs(minfo.theta - (np.pi / 2)) < 2 * delta_theta):
        raise ValueError('Input minfo has to include equator.')
    idx_ref = np.searchsorted(minfo.theta, np.pi / 2, side='left')
    area_ref = minfo.weight[idx_ref]
    nphi_r = np.zeros_like(minfo.nphi)
    w_r = np.zeros_like(minfo.weight)
    for t in range(minfo.nrow):
        area = minfo.weight[t]
        if (np.pi / 2 - orig_band) < minfo.theta[t] < (np.pi / 2 + orig_band):
            ratio = 1
        else:
            ratio = (area / area_ref) ** ratio_pow
        nphi_r[t] = max(2, int(np.round(minfo.nphi[t] * ratio)))
        w_r[t] = area * minfo.nphi[t] / nphi_r[t]
    return MapInfo(minfo.theta, nphi=nphi_r, weight=w_r)


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
