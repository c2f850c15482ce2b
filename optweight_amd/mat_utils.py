'''
Small symmetric matrices per pixel / per multipole, mirroring optweight's
mat_utils: full_matrix (mat_utils.py:125-155), matpow / _eigpow (220-365),
wavmatpow (367-411),
atleast_nd (445-473), matvec (475-532), and mat_c_utils.eigpow
(mat_c_utils.c:9-157: V diag(lam^p) V^T, eigenvalues below lim * max set
to zero, all zero when max < lim0).

Powers are set-up work: they run on the host in float64 (numpy.linalg.eigh,
as the reference computes them in float64 through LAPACK). matvec on device
tensors runs through the HIP pixel kernel.
'''
import numpy as np
import torch

from optweight_amd import _lib


def atleast_nd(mat, ndim, append=False):
    mat = np.asanyarray(mat)
    axes = np.arange(ndim - mat.ndim)
    if append:
        axes -= axes.size
    return np.expand_dims(mat, tuple(axes))


def full_matrix(mat):
    if isinstance(mat, torch.Tensor):
        if mat.ndim == 2:
            npol = mat.shape[0]
            return torch.eye(npol, dtype=mat.dtype, device=mat.device)[:, :, None] * mat
        return mat
    mat = np.asarray(mat)
    if mat.ndim not in (2, 3):
        raise ValueError(f'Matrix dimension : {mat.ndim} not supported. Requires dim = 2 or 3.')
    if mat.ndim == 2:
        npol = mat.shape[0]
        mat = np.eye(npol, dtype=mat.dtype)[:, :, np.newaxis] * mat
    return mat


def eigpow(imat, power, lim=None, lim0=None):
    '''(ncomp, ncomp, nsamp) -> V diag(lam^power) V^T per sample.'''
    imat = np.asarray(imat)
    ishape = imat.shape
    if imat.ndim not in (2, 3):
        raise ValueError(f'Input ndim : {imat.ndim} != 2 or 3.')
    if imat.ndim == 2:
        imat = imat[:, :, np.newaxis]
    if imat.shape[0] != imat.shape[1]:
        raise ValueError(f'Input shape : {ishape} does not match (ncomp, ncomp, nsamp) shape')
    if imat.dtype not in (np.float32, np.float64):
        raise ValueError(f'Input dtype : {imat.dtype} not supported')
    if lim0 is None:
        lim0 = np.finfo(imat.dtype).tiny ** 0.5
    if lim is None:
        lim = 1e-6
    m = np.moveaxis(imat.astype(np.float64), -1, 0)
    w, v = np.linalg.eigh(m)
    wmax = w[:, -1:]
    keep = (w >= lim * wmax) & (wmax >= lim0)
    wp = np.zeros_like(w)
    wp[keep] = w[keep] ** power
    out = np.einsum('nij,nj,nkj->nik', v, wp, v)
    return np.moveaxis(out, 0, -1).astype(imat.dtype).reshape(ishape)


def matpow(mat, power, return_diag=False, skip_unit_pow=True, lim=1e-9, lim0=None,
           chunksize=100_000):
    '''Positive semidefinite matrix power for (npol, npol, N) or (npol, N)
    input (diagonal input: elementwise, non-positive entries set to 0).'''
    mat = np.asarray(mat)
    if mat.ndim == 1:
        mat = mat[np.newaxis, :]
    shape_in = mat.shape
    if not 2 <= mat.ndim <= 3:
        raise ValueError(f'Matrix is not (npol, npol, N) or (npol, N), got : {mat.shape}')
    if power != 1 or not skip_unit_pow:
        if mat.ndim == 2:
            mat = mat.copy()
            mask = mat <= 0
            mat[mask] = 0
            np.power(mat, power, out=mat, where=~mask)
        else:
            dtype_in = mat.dtype
            out = np.empty(mat.shape, dtype=dtype_in)
            for start in range(0, mat.shape[-1], chunksize):
                sl = slice(start, start + chunksize)
                out[:, :, sl] = eigpow(mat[:, :, sl].astype(np.float64), power, lim=lim,
                                       lim0=lim0).astype(dtype_in)
            mat = out
    mat = mat.reshape(shape_in)
    if not return_diag:
        mat = full_matrix(mat)
    return np.array(mat, copy=True)


def wavmatpow(m_wav, power, **matpow_kwargs):
    '''Power of a block-diagonal wavelet matrix (mat_utils.py:367-411):
    each diagonal block through matpow (host float64).'''
    from optweight_amd import wavtrans
    if not np.array_equal(m_wav.indices[:, 0], m_wav.indices[:, 1]):
        raise NotImplementedError('Can only raise diagonal block matrix to a power')
    out = wavtrans.Wav(2, dtype=m_wav.dtype)
    for jidx in range(m_wav.shape[0]):
        key = (jidx, jidx)
        mat = m_wav.maps[key]
        if isinstance(mat, torch.Tensor):
            mat = mat.detach().cpu().numpy()
        out.add(key, matpow(mat, power, **matpow_kwargs).astype(m_wav.dtype, copy=False),
                m_wav.minfos[key])
    return out


def matvec(mat, vec, inplace=False):
    '''Pixel-diagonal matrix-vector product on device tensors: elementwise
    when shapes match, else (n, n, N) x (n, N).'''
    if not isinstance(vec, torch.Tensor):
        # host arrays: the reference semantics in numpy
        mat = np.asarray(mat)
        if mat.shape == vec.shape:
            return np.multiply(mat, vec, out=vec if inplace else None)
        out = np.einsum('ijk, jk -> ik', mat, vec, optimize=True)
        if inplace:
            vec[...] = out
            return vec
        return out
    _lib.require_cuda(vec)
    if not isinstance(mat, torch.Tensor):
        mat = torch.as_tensor(np.asarray(mat), device=vec.device, dtype=vec.dtype)
    if mat.dtype != vec.dtype:
        mat = mat.to(vec.dtype)
    mat = mat.contiguous()
    _lib.require_cuda(mat)
    st = _lib.stream_ptr(vec.device)
    code = _lib.dtype_code(vec)
    out = vec if inplace else torch.empty_like(vec)
    if tuple(mat.shape) == tuple(vec.shape):
        _lib.call('owmi_pix_matvec', code, _lib.ptr(mat), 0, _lib.ptr(vec), _lib.ptr(out), 1,
                  vec.numel(), 1, st)
        return out
    if vec.ndim >= 2 and tuple(mat.shape) == tuple(vec.shape[-2:]):
        ntrans = vec.numel() // mat.numel()
        _lib.call('owmi_pix_matvec', code, _lib.ptr(mat), 0, _lib.ptr(vec), _lib.ptr(out), 1,
                  mat.numel(), ntrans, st)
        return out
    npol, npix = vec.shape[-2], vec.shape[-1]
    if tuple(mat.shape) != (npol, npol, npix):
        raise ValueError(f'matrix shape : {tuple(mat.shape)} does not match up with vector '
                         f'shape : {tuple(vec.shape)}')
    ntrans = vec.numel() // (npol * npix)
    # each thread reads all components of its pixel before writing: in-place safe
    _lib.call('owmi_pix_matvec', code, _lib.ptr(mat), 1, _lib.ptr(vec), _lib.ptr(out), npol,
              npix, ntrans, st)
    return out


def symm2triu(mat, axes, return_axis=False):
    """Copy of the upper-triangular elements (row-major) of a matrix that is
    symmetric in two adjacent axes, which become one axis of N (N + 1) / 2
    (mat_utils.py:7-65)."""
    mat = np.asarray(mat)
    if len(axes) != 2:
        raise ValueError(f'Need two axes, got {len(axes)} ({axes})')
    axes = np.sort([a + mat.ndim if a < 0 else a for a in axes])
    if mat.shape[axes[0]] != mat.shape[axes[1]]:
        raise ValueError(f'Matrix (shape : {mat.shape}) is not N x N in specified axes : {axes}.')
    if axes[1] - axes[0] != 1:
        raise ValueError(f'Axes are not adjacent: {axes}')
    n = mat.shape[axes[0]]
    mask = np.triu(np.ones((n, n), dtype=bool))
    out = np.ascontiguousarray(mat[(slice(None),) * axes[0] + (mask,)])
    return (out, int(axes[0])) if return_axis else out


def triu2symm(mat, axis):
    """Inverse of symm2triu: the upper-triangular elements in axis expanded to
    a symmetric N x N pair of axes (mat_utils.py:67-117)."""
    mat = np.asarray(mat)
    if axis < 0:
        axis += mat.ndim
    n = (-1 + np.sqrt(1 + 8 * mat.shape[axis])) / 2
    if np.floor(n) != n:
        raise ValueError(f'Cannot intepret axis : {axis} of shape {mat.shape} '
                         f'as upper triangular elements')
    n = int(n)
    out = np.zeros(mat.shape[:axis] + (n, n) + mat.shape[axis + 1:], dtype=mat.dtype)
    lead = (slice(None),) * axis
    for k, (i, j) in enumerate(zip(*np.triu_indices(n))):
        out[lead + (i, j)] = mat[lead + (k,)]
        out[lead + (j, i)] = mat[lead + (k,)]
    return out


def flattened_view(mat, axes, return_flat_axes=False):
    """View of mat with each group of axes merged into one (the first axis
    of the group keeps its place); also the merged axes' indices in the
    output (mat_utils.py:157-218)."""
    shape_in = np.asarray(mat.shape)
    axes = [[a + mat.ndim if a < 0 else a for a in sub] for sub in axes]
    flat = [a for sub in axes for a in sub]
    if np.unique(flat).size != len(flat):
        raise ValueError(f'Axes cannot contain duplicates, got : {axes}')
    shape_out = shape_in.copy()
    merged = np.zeros_like(shape_out)
    for sub in axes:
        sub = np.asarray(sub)
        shape_out[sub[0]] = np.prod(shape_in[sub])
        shape_out[sub[1:]] = 0
        merged[sub[0]] = 1
    keep = shape_out != 0
    view = mat.reshape(tuple(int(n) for n in shape_out[keep]))
    if return_flat_axes:
        return view, [int(i) for i in np.nonzero(merged[keep])[0]]
    return view
