'''
2-D real Fourier transforms of flat-sky (CAR) maps and the Fourier-plane
helpers of the Fourier-wavelet noise model, mirroring optweight's dft
(dft.py:13-610).

Transforms are orthonormal (scaled by 1/sqrt(ny nx)), on device tensors
(..., ny, nx) <-> (..., ny, nx // 2 + 1), through the library FFT passes of
torch.fft (rocFFT) with the real-field conventions of a 2-D real FFT made
exact by owmi_fourier_fix: after the forward transform the self-conjugate
coefficients are exactly real, and the inverse uses only the real parts of
the columns lx = 0 and lx = nx / 2 (even nx) after the inverse pass along
y. Per-coefficient matrices and the real-field inner product run as owmi
HIP kernels. Wavenumber axes (host numpy): ly = 2 pi fftfreq(ny, dy),
lx = 2 pi rfftfreq(nx, dx) with (dy, dx) the WCS pixel steps in radians,
with no correction for sky curvature (dft.py:125-154).
'''
import numpy as np
import torch
from scipy.interpolate import interp1d

from optweight_amd import _lib

_CPLX = {torch.float32: torch.complex64, torch.float64: torch.complex128}
_REAL = {torch.complex64: torch.float32, torch.complex128: torch.float64}


def _fix(fmap, nlead, ny, nlx, nx_even, mode):
    if not fmap.is_contiguous():
        raise ValueError('the real-field fix-up needs a contiguous array')
    _lib.call('owmi_fourier_fix', _lib.dtype_code(fmap), _lib.ptr(fmap), nlead, ny, nlx,
              int(bool(nx_even)), mode, _lib.stream_ptr(fmap.device))


def rfft(imap, fmap):
    '''Real-to-complex orthonormal 2-D FFT of imap (..., ny, nx) into fmap
    (..., ny, nx // 2 + 1) (dft.py:13-39).'''
    ny, nx = imap.shape[-2:]
    if tuple(fmap.shape) != tuple(imap.shape[:-2]) + (ny, nx // 2 + 1):
        raise ValueError(f'Inconsistent shapes: imap : {tuple(imap.shape)}, '
                         f'fmap : {tuple(fmap.shape)}')
    if fmap.dtype != _CPLX.get(imap.dtype):
        raise TypeError(f'imap.dtype : {imap.dtype} and fmap.dtype : {fmap.dtype} do not match')
    _lib.require_cuda(imap, fmap)
    torch.fft.rfft2(imap, norm='ortho', out=fmap)
    nlead = fmap.numel() // max(ny * (nx // 2 + 1), 1)
    _fix(fmap, nlead, ny, nx // 2 + 1, nx % 2 == 0, 0)
    return fmap


def irfft(fmap, omap):
    '''Complex-to-real orthonormal 2-D FFT of fmap (..., ny, nx // 2 + 1)
    into omap (..., ny, nx); fmap is not modified (dft.py:41-67).'''
    ny, nx = omap.shape[-2:]
    if tuple(fmap.shape) != tuple(omap.shape[:-2]) + (ny, nx // 2 + 1):
        raise ValueError(f'Inconsistent shapes: fmap : {tuple(fmap.shape)}, '
                         f'omap : {tuple(omap.shape)}')
    if fmap.dtype != _CPLX.get(omap.dtype):
        raise TypeError(f'omap.dtype : {omap.dtype} and fmap.dtype : {fmap.dtype} do not match')
    _lib.require_cuda(fmap, omap)
    # the library may return the y pass with permuted strides
    tmp = torch.fft.ifft(fmap, dim=-2, norm='ortho').contiguous()
    nlx = nx // 2 + 1
    _fix(tmp, tmp.numel() // max(ny * nlx, 1), ny, nlx, nx % 2 == 0, 1)
    torch.fft.irfft(tmp, n=nx, dim=-1, norm='ortho', out=omap)
    return omap


def allocate_fmap(shape, dtype, fill_value=0, device=None):
    '''(..., ny, nx // 2 + 1) complex array for the rfft of maps of shape
    (..., ny, nx) and real dtype (dft.py:69-92).'''
    dtype = _torch_dtype(dtype)
    cdt = dtype if dtype.is_complex else _CPLX[dtype]
    ny, nx = shape[-2:]
    return torch.full(tuple(shape[:-2]) + (ny, nx // 2 + 1), fill_value, dtype=cdt,
                      device=device)


def allocate_map(fshape, dtype, fill_value=0, device=None):
    '''(..., ly, 2 (lx - 1) + 1) real array for the irfft of coefficients of
    shape (..., ly, lx): always odd in x (dft.py:94-123).'''
    dtype = _torch_dtype(dtype)
    rdt = _REAL[dtype] if dtype.is_complex else dtype
    ly, lx = fshape[-2:]
    return torch.full(tuple(fshape[:-2]) + (ly, 2 * (lx - 1) + 1), fill_value, dtype=rdt,
                      device=device)


def _torch_dtype(dtype):
    if isinstance(dtype, torch.dtype):
        return dtype
    return {np.dtype(np.float32): torch.float32, np.dtype(np.float64): torch.float64,
            np.dtype(np.complex64): torch.complex64,
            np.dtype(np.complex128): torch.complex128}[np.dtype(dtype)]


def laxes_real(shape, wcs):
    '''Wavenumber axes ly (ny,), lx (nx // 2 + 1,) of a CAR geometry.'''
    step = np.radians(np.asarray(wcs.wcs.cdelt, dtype=np.float64)[::-1])
    ly = np.fft.fftfreq(shape[-2], step[0]) * 2 * np.pi
    lx = np.fft.rfftfreq(shape[-1], step[1]) * 2 * np.pi
    return ly, lx


def laxes2lmap(ly, lx, dtype=np.float64):
    lmap = np.empty((2, ly.size, lx.size), dtype=dtype)
    lmap[0] = ly[:, np.newaxis]
    lmap[1] = lx[np.newaxis, :]
    return lmap


def lmap_real(shape, wcs, dtype=np.float64):
    return laxes2lmap(*laxes_real(shape, wcs), dtype=dtype)


def laxes2modlmap(ly, lx, dtype=np.float64):
    '''|l| = sqrt(ly^2 + lx^2) on the half plane (ny, nlx).'''
    lmap = laxes2lmap(ly, lx, dtype=dtype)
    return np.sum(lmap ** 2, axis=0) ** 0.5


def modlmap_real(shape, wcs, dtype=np.float64):
    return laxes2modlmap(*laxes_real(shape, wcs), dtype=dtype)


def lwcs_real(shape, wcs):
    '''CAR-style WCS of the half plane (dft.py:247-269): steps 2 pi / (n d)
    with the x step positive, reference pixel (0, ny // 2 + 1).'''
    from optweight_amd.map_utils import CarWCS
    lres = 2 * np.pi / (np.radians(np.asarray(wcs.wcs.cdelt, dtype=np.float64)[::-1]) *
                        np.asarray(shape[-2:], dtype=np.float64))
    lres[-1] = abs(lres[-1])
    ny = shape[-2]
    return CarWCS(cdelt=lres[::-1], crpix=[0, ny // 2 + 1], crval=[0, 0])


def cl2flat(c_ell, ells, modlmap):
    '''Linear interpolation of c_ell (..., nell) at |l| = modlmap, extending
    the first and last values to the range of modlmap (dft.py:367-414).'''
    c_ell = np.asarray(c_ell)
    ells = np.asarray(ells, dtype=np.float64)
    lmin_out, lmax_out = modlmap.min(), modlmap.max()
    ell_start, ell_end = [], []
    if ells[0] > lmin_out:
        ell_start = np.asarray([lmin_out])
        c_ell = np.concatenate((c_ell[..., :1], c_ell), axis=-1)
    if ells[-1] < lmax_out:
        ell_end = np.asarray([lmax_out])
        c_ell = np.concatenate((c_ell, c_ell[..., -1:]), axis=-1)
    ells = np.concatenate((ell_start, ells, ell_end))
    cs = interp1d(ells, c_ell, kind='linear', assume_sorted=True, bounds_error=True)
    return cs(modlmap).astype(c_ell.dtype, copy=False)


def fmul(fmap, fmat2d=None, fmat1d=None, ells=None, modlmap=None, out=None):
    '''f'[..., i, ly, lx] = m[i, j, ly, lx] f[..., j, ly, lx] with m given per
    coefficient (fmat2d) or as a function of |l| (fmat1d at ells, sampled
    on modlmap) (dft.py:297-330).'''
    if fmat2d is not None and fmat1d is not None:
        raise ValueError('Cannot have both fmat2d and fmat1d')
    if fmat1d is not None:
        fmat2d = cl2flat(fmat1d, ells, modlmap)
    return fmul_2d(fmap, fmat2d, out=out)


def fmul_2d(fmap, fmat2d, out=None):
    '''out = fmat2d fmap per coefficient: fmat2d (npol, npol, nly, nlx) is a
    matrix, (npol, nly, nlx) or (nly, nlx) a diagonal (dft.py:332-365).
    out may be fmap.'''
    _lib.require_cuda(fmap)
    if fmap.ndim < 3:
        fmap = fmap.reshape((1,) * (3 - fmap.ndim) + tuple(fmap.shape))
    rdt = _REAL[fmap.dtype]
    if not isinstance(fmat2d, torch.Tensor):
        fmat2d = torch.as_tensor(np.asarray(fmat2d), device=fmap.device)
    mat = fmat2d.to(device=fmap.device, dtype=rdt).contiguous()
    if out is None:
        out = torch.empty_like(fmap)
    elif tuple(out.shape) != tuple(fmap.shape):
        out = out.reshape(fmap.shape)
    _lib.require_cuda(out)
    npix = fmap.shape[-2] * fmap.shape[-1]
    code = _lib.dtype_code(fmap)
    st = _lib.stream_ptr(fmap.device)
    if mat.ndim in (2, 3):
        if tuple(mat.shape[-2:]) != tuple(fmap.shape[-2:]):
            raise ValueError(f'fmat2d {tuple(mat.shape)} does not match fmap {tuple(fmap.shape)}')
        if mat.ndim == 2:
            nrows = fmap.numel() // npix
            _lib.call('owmi_fmul', code, _lib.ptr(mat), 0, _lib.ptr(fmap), _lib.ptr(out), 1,
                      npix, nrows, 0, st)
        else:
            npol = mat.shape[0]
            if fmap.shape[-3] != npol:
                raise ValueError(f'fmat2d {tuple(mat.shape)} does not match fmap '
                                 f'{tuple(fmap.shape)}')
            _lib.call('owmi_fmul', code, _lib.ptr(mat), 0, _lib.ptr(fmap), _lib.ptr(out), npol,
                      npix, fmap.numel() // (npol * npix), npix, st)
        return out
    npol = mat.shape[0]
    if tuple(mat.shape) != (npol, npol) + tuple(fmap.shape[-2:]) or fmap.shape[-3] != npol:
        raise ValueError(f'fmat2d {tuple(mat.shape)} does not match fmap {tuple(fmap.shape)}')
    ntrans = fmap.numel() // (npol * npix)
    if npol <= 4:
        _lib.call('owmi_fmul', code, _lib.ptr(mat), 1, _lib.ptr(fmap), _lib.ptr(out), npol, npix,
                  ntrans, 0, st)
    else:
        tmp = torch.empty_like(fmap)
        _lib.call('owmi_fmul_generic', code, _lib.ptr(mat), _lib.ptr(fmap), _lib.ptr(out),
                  _lib.ptr(tmp), npol, npix, ntrans, st)
    return out


def contract_fxg_dev(fmap, gmap, nx_even=None, out=None):
    '''Real-field inner product of half-plane coefficients as a 0-d float64
    device tensor: 2 sum Re(f g*) minus the column lx = 0 and, when nx is
    even, the column lx = nx / 2. nx_even None decides the parity as
    dft.py:449-496 does: even when f is real at (ly = 0, lx = last) and,
    for even ny, at (ly = ny / 2, lx = last).'''
    if tuple(fmap.shape) != tuple(gmap.shape):
        raise ValueError(f'Shape fmap ({tuple(fmap.shape)}) != shape gmap ({tuple(gmap.shape)})')
    _lib.require_cuda(fmap, gmap)
    if fmap.dtype != gmap.dtype:
        raise ValueError('dtype mismatch')
    ny, nlx = fmap.shape[-2:]
    nlead = fmap.numel() // max(ny * nlx, 1)
    if out is None:
        out = torch.empty((), dtype=torch.float64, device=fmap.device)
    parity = -1 if nx_even is None else int(bool(nx_even))
    _lib.call('owmi_fourier_dot', _lib.dtype_code(fmap), _lib.ptr(fmap), _lib.ptr(gmap), nlead,
              ny, nlx, parity, _lib.ptr(out), _lib.stream_ptr(fmap.device))
    return out


def contract_fxg(fmap, gmap):
    '''Python float of contract_fxg_dev (synchronises).'''
    return float(contract_fxg_dev(fmap, gmap).item())


def slice_fmap(fmap, slices_y, slice_x, laxes=None):
    '''Rows slices_y[0] then slices_y[1] and columns slice_x of fmap as a
    new contiguous array (dft.py:498-533).'''
    parts = [fmap[..., slices_y[0], slice_x], fmap[..., slices_y[1], slice_x]]
    if isinstance(fmap, torch.Tensor):
        out = torch.cat(parts, dim=-2).contiguous()
    else:
        out = np.concatenate(parts, axis=-2)
    if laxes is not None:
        ly, lx = laxes
        out = out, (np.concatenate((ly[slices_y[0]], ly[slices_y[1]])), lx[slice_x])
    return out


def add_to_fmap(fmap_large, fmap_small, slices_y, slice_x):
    '''In-place fmap_large[rows, cols] += fmap_small (dft.py:535-560).'''
    fmap_large[..., slices_y[0], slice_x] += fmap_small[..., slices_y[0], :]
    fmap_large[..., slices_y[1], slice_x] += fmap_small[..., slices_y[1], :]
    return fmap_large


def get_optimal_fftlen(len_min, even=True):
    '''Smallest 2^a 3^b 5^c 7^d >= len_min (even if requested), with the
    search bounds of dft.py:562-610 (exponents below ceil(log_p len_min),
    starting from 2^ceil(log2 len_min)).'''
    if len_min == 0:
        return len_min
    if len_min == 1 and even:
        return 2
    max_a = int(np.ceil(np.log(len_min) / np.log(2)))
    max_b = int(np.ceil(np.log(len_min) / np.log(3)))
    max_c = int(np.ceil(np.log(len_min) / np.log(5)))
    max_d = int(np.ceil(np.log(len_min) / np.log(7)))
    best = 2 ** max_a
    for a in range(max_a):
        for b in range(max_b):
            for c in range(max_c):
                for d in range(max_d):
                    n = 2 ** a * 3 ** b * 5 ** c * 7 ** d
                    if (even and n % 2) or n < len_min:
                        continue
                    best = min(best, n)
    return best
