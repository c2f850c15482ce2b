'''
Wavelet block vectors and matrices of maps, and the axisymmetric wavelet
transforms between alm and per-scale maps, mirroring optweight's wavtrans:
Wav (wavtrans.py:10-369), wav2alm (406-465), alm2wav (467-515),
preshape2npol (517-546), and the 2-D Fourier-wavelet transforms f2wav /
wav2f (687-745).

A Wav holds one map per index (an int for a block vector, an (i, j) tuple
for a block matrix) with its own MapInfo; maps are numpy arrays or device
tensors (the transforms need device tensors). Wavelet scale j of an alm is
w^j_lm = w_ell[j, l] a_lm for l <= lmax_j, mapped onto the Gauss-Legendre
grid of band limit lmax_j (lmax_j + 1 rings, 2 lmax_j + 1 pixels per ring).
'''
import numpy as np
import torch

from optweight_amd import map_utils, alm_utils, sht


def _np_dtype(dtype):
    if isinstance(dtype, torch.dtype):
        return {torch.float32: np.dtype(np.float32), torch.float64: np.dtype(np.float64),
                torch.complex64: np.dtype(np.complex64),
                torch.complex128: np.dtype(np.complex128)}[dtype]
    return np.dtype(dtype)


class Wav:
    '''Block vector (ndim 1) or matrix (ndim 2) of wavelet maps.'''

    def __init__(self, ndim, indices=None, minfos=None, preshape=None, dtype=np.float64):
        if ndim not in (1, 2):
            raise NotImplementedError(f'{ndim}-D matrices not supported')
        if indices is None:
            indices = np.zeros((0, ndim), dtype=int)
        else:
            indices = np.asarray(indices)
            if indices.ndim == 1:
                indices = indices[:, np.newaxis]
        self.ndim = ndim
        self.preshape = () if preshape is None else tuple(preshape)
        self.dtype = _np_dtype(dtype)
        self.maps = {}
        self.minfos = {}
        self.indices = np.zeros((0, ndim), dtype=int)
        for idx in range(indices.shape[0]):
            minfo = minfos[idx]
            self.add(indices[idx], np.zeros(self.preshape + (minfo.npix,), dtype=self.dtype),
                     minfo)

    @property
    def shape(self):
        '''Shape of the corresponding dense block vector / matrix.'''
        if self.indices.shape[0] == 0:
            return (0,) * self.indices.shape[1]
        return tuple(int(1 + np.max(self.indices[:, i])) for i in range(self.indices.shape[1]))

    def add(self, index, m_arr, minfo):
        '''Place m_arr (preshape + (npix,)) at index.'''
        index = np.asarray(index)
        if index.ndim == 0:
            index = index[np.newaxis]
        if index.size != self.ndim:
            raise ValueError(f'Wrong dimension for index, got : {index.size} '
                             f'expected : {self.ndim}')
        if not issubclass(index.dtype.type, np.integer):
            raise ValueError(f'Index should be integer, instead got : {index.dtype}')
        if np.any(index < 0):
            raise ValueError(f'Index cannot be negative, got {index}')
        key = int(index[0]) if index.size == 1 else tuple(int(i) for i in index)
        if _np_dtype(m_arr.dtype) != self.dtype:
            raise ValueError(f'Adding array of wrong dtype : {m_arr.dtype}, '
                             f'expected {self.dtype}')
        preshape = tuple(m_arr.shape[:-1])
        if self.maps:
            if preshape != self.preshape:
                raise ValueError(f'Wrong leading dimensions of map array: expected : '
                                 f'{self.preshape}, got {preshape}')
        else:
            self.preshape = preshape
        if m_arr.shape[-1] != minfo.npix:
            raise ValueError(f'map has {m_arr.shape[-1]} pixels, minfo {minfo.npix}')
        self.maps[key] = m_arr
        self.minfos[key] = minfo
        self.indices = np.unique(np.append(self.indices, index[np.newaxis, :], axis=0), axis=0)

    def get_indices_diag(self):
        '''Keys of the diagonal (every key for a vector).'''
        if self.ndim == 1:
            return [int(i) for i in self.indices[:, 0]]
        return [(int(i), int(j)) for i, j in self.indices if i == j]

    def get_minfos_diag(self, copy=True):
        '''MapInfo objects of the diagonal, in key order (wavtrans.py:253-280).'''
        import copy as _copy
        out = [self.minfos[k] for k in sorted(self.get_indices_diag())]
        return [_copy.deepcopy(m) for m in out] if copy else out

    def diag(self):
        '''New block vector holding the diagonal (or a copy of a vector).'''
        out = Wav(1, preshape=self.preshape, dtype=self.dtype)
        for key in self.get_indices_diag():
            m = self.maps[key]
            m = m.clone() if isinstance(m, torch.Tensor) else m.copy()
            out.add(key if self.ndim == 1 else key[0], m, self.minfos[key])
        return out

    def astype(self, dtype):
        '''Change the map dtype (copies when needed).'''
        for key in self.maps:
            m = self.maps[key]
            if isinstance(m, torch.Tensor):
                tdt = {np.dtype(np.float32): torch.float32,
                       np.dtype(np.float64): torch.float64}[np.dtype(dtype)]
                self.maps[key] = m.to(tdt)
            else:
                self.maps[key] = m.astype(dtype, copy=False)
        self.dtype = np.dtype(dtype)

    def to(self, device):
        '''Copy with every map as a tensor on device.'''
        out = Wav(self.ndim, preshape=self.preshape, dtype=self.dtype)
        for key in self.maps:
            m = self.maps[key]
            t = m.to(device) if isinstance(m, torch.Tensor) else torch.as_tensor(
                np.ascontiguousarray(m), device=device)
            out.add(np.asarray(key), t, self.minfos[key])
        return out


def preshape2npol(preshape):
    '''npol of a block vector (preshape (npol,)) or matrix ((npol, npol)).'''
    if len(preshape) > 2 or len(set(preshape)) > 1:
        raise ValueError(f'Could not determine npol from preshape : {preshape}')
    return 1 if len(preshape) == 0 else preshape[0]


def wav_lmax(w_ell, idx):
    '''Band limit of kernel idx: the last multipole where it is positive.'''
    lmax = w_ell.shape[-1] - 1
    return int(lmax - np.argmax(np.asarray(w_ell)[idx, ::-1] > 0))


_GL_MINFO = {}


def _gl_minfo(lmax):
    '''The Gauss-Legendre geometry of band limit lmax, one object per lmax:
    the transform plans are cached on the geometry object, so repeated
    wavelet transforms reuse them.'''
    if lmax not in _GL_MINFO:
        _GL_MINFO[lmax] = map_utils.MapInfo.map_info_gauss_legendre(lmax + 1, 2 * lmax + 1)
    return _GL_MINFO[lmax]


def alm2wav(alm, ainfo, spin, w_ell, wav=None, adjoint=False, lmaxs=None):
    '''Wavelet maps of device alm (..., npol, nelem): scale j is Y (or W Y
    if adjoint) of w_ell[j] a_lm on its Gauss-Legendre grid.'''
    wlms, winfos = alm_utils.alm2wlm_axisym(alm, ainfo, w_ell, lmaxs=lmaxs)
    indices = np.arange(len(wlms))
    rdt = torch.float64 if alm.dtype == torch.complex128 else torch.float32
    if wav is None:
        minfos = [_gl_minfo(w.lmax) for w in winfos]
        wav = Wav(1, preshape=tuple(alm.shape[:-1]),
                  dtype=np.float64 if rdt == torch.float64 else np.float32)
        for idx in indices:
            wav.add(idx, torch.zeros(tuple(alm.shape[:-1]) + (minfos[idx].npix,), dtype=rdt,
                                     device=alm.device), minfos[idx])
    for idx in indices:
        sht.alm2map(wlms[idx], wav.maps[int(idx)], winfos[idx], wav.minfos[int(idx)], spin,
                    adjoint=adjoint)
    return wav


def wav2alm(wav, alm, ainfo, spin, w_ell, adjoint=False):
    '''alm = sum_j w_ell[j] Y^T W (or Y^T if adjoint) of wavelet map j;
    alm is overwritten. Returns (alm, ainfo).'''
    if len(wav.shape) != 1:
        raise ValueError('Input wavelet container is not a vector')
    alm.zero_()
    for idx in range(w_ell.shape[0]):
        lmax_w = wav_lmax(w_ell, idx)
        imap = wav.maps[idx]
        if imap.ndim == 1:
            imap = imap[None, :]
        winfo = alm_utils.AlmInfo(lmax_w)
        cdt = torch.complex128 if imap.dtype == torch.float64 else torch.complex64
        wlm = torch.zeros(tuple(alm.shape[:-1]) + (winfo.nelem,), dtype=cdt, device=alm.device)
        sht.map2alm(imap, wlm, wav.minfos[idx], winfo, spin, adjoint=adjoint)
        alm_utils.wlm2alm_axisym([wlm], [winfo], w_ell[idx:idx + 1, :], alm=alm, ainfo=ainfo)
    return alm, ainfo


def _fourier_plan(fmap, fkernelset):
    from optweight_amd import dft
    rdt = dft._REAL[fmap.dtype]
    plan = fkernelset.device_plan(fmap.device, rdt)
    ny, nlx = fmap.shape[-2:]
    if (ny, nlx) != tuple(fkernelset.shape_full):
        raise ValueError(f'fmap {tuple(fmap.shape)} does not match the kernels '
                         f'{tuple(fkernelset.shape_full)}')
    nlead = fmap.numel() // max(ny * nlx, 1)
    return plan, nlead


def f2wav(fmap, wav, fkernelset):
    """Wavelet maps of 2-D Fourier coefficients fmap (..., ny, nx // 2 + 1):
    map j = irfft(slice_j(fmap) k_j) on the (ny_j, 2 (nx_j - 1) + 1) grid of
    wav.maps[j] (wavtrans.py:687-712). One gather kernel cuts and weights
    every slice; the inverse FFT of each runs per scale."""
    from optweight_amd import dft, _lib
    _lib.require_cuda(fmap)
    fmap = fmap.contiguous()
    plan, nlead = _fourier_plan(fmap, fkernelset)
    lead = tuple(fmap.shape[:-2])
    ny, nlx = fmap.shape[-2:]
    small = torch.empty(max(nlead * plan['total'], 1), dtype=fmap.dtype, device=fmap.device)
    _lib.call('owmi_fwav_gather', _lib.dtype_code(fmap), _lib.ptr(fmap), nlead, ny, nlx,
              len(plan['shapes']), _lib.ptr(plan['desc']), _lib.ptr(plan['rows']),
              _lib.ptr(plan['kern']), plan['nblocks'], _lib.ptr(small),
              _lib.stream_ptr(fmap.device))
    for widx, off, ny_j, nx_j in plan['shapes']:
        blk = small[nlead * off:nlead * (off + ny_j * nx_j)].view(lead + (ny_j, nx_j))
        map_2d = map_utils.view_2d(wav.maps[widx], wav.minfos[widx])
        dft.irfft(blk, map_2d)
    return wav


def wav2f(wav, fmap, fkernelset):
    """2-D Fourier coefficients of a wavelet vector, the adjoint of f2wav:
    fmap = sum_j pad_j(rfft(wav.maps[j]) k_j), overwriting fmap
    (wavtrans.py:714-745). The forward FFT of each scale runs per scale;
    one scatter kernel weights and adds them in scale order."""
    from optweight_amd import dft, _lib
    _lib.require_cuda(fmap)
    if not fmap.is_contiguous():
        raise ValueError('fmap must be contiguous')
    plan, nlead = _fourier_plan(fmap, fkernelset)
    lead = tuple(fmap.shape[:-2])
    ny, nlx = fmap.shape[-2:]
    small = torch.empty(max(nlead * plan['total'], 1), dtype=fmap.dtype, device=fmap.device)
    for widx, off, ny_j, nx_j in plan['shapes']:
        blk = small[nlead * off:nlead * (off + ny_j * nx_j)].view(lead + (ny_j, nx_j))
        map_2d = map_utils.view_2d(wav.maps[widx], wav.minfos[widx])
        dft.rfft(map_2d, blk)
    _lib.call('owmi_fwav_scatter', _lib.dtype_code(fmap), _lib.ptr(small), nlead, ny, nlx,
              len(plan['shapes']), _lib.ptr(plan['desc']), _lib.ptr(plan['invrows']),
              _lib.ptr(plan['kern']), _lib.ptr(fmap), _lib.stream_ptr(fmap.device))
    return fmap


def _map_key(idx):
    idx = np.asarray(idx)
    if idx.size == 1:
        return f'm_{int(idx.ravel()[0])}', int(idx.ravel()[0])
    return f'm_{int(idx[0])}_{int(idx[1])}', (int(idx[0]), int(idx[1]))


def append_wav_to_hdf(hfile, wav, extra=None, **kwargs):
    """Attributes ndim, indices, preshape and dtype (numpy type character),
    optional extra datasets, and one group m_j or m_j_k per map written by
    map_utils.append_map_to_hdf (wavtrans.py:548-590)."""
    hfile.attrs['ndim'] = wav.ndim
    hfile.attrs['indices'] = wav.indices
    hfile.attrs['preshape'] = wav.preshape
    hfile.attrs['dtype'] = np.dtype(wav.dtype).char
    if extra is not None:
        for key in extra:
            hfile.create_dataset(key, data=extra[key])
    for idx in wav.indices:
        dname, key = _map_key(idx)
        map_utils.append_map_to_hdf(hfile.create_group(dname), wav.maps[key], wav.minfos[key],
                                    **kwargs)


def wav_from_hdf(hfile, extra=None):
    extra_dict = {key: hfile[key][()] for key in extra} if extra is not None else None
    wav = Wav(int(hfile.attrs['ndim']), preshape=tuple(int(n) for n in hfile.attrs['preshape']),
              dtype=np.dtype(hfile.attrs['dtype']))
    for idx in hfile.attrs['indices']:
        dname, key = _map_key(idx)
        m_arr, minfo = map_utils.map_from_hdf(hfile[dname])
        wav.add(np.asarray(key), m_arr, minfo)
    return (wav, extra_dict) if extra is not None else wav


def write_wav(fname, wav, extra=None, **kwargs):
    with map_utils._h5py().File(map_utils._with_ext(fname), 'w') as hfile:
        append_wav_to_hdf(hfile, wav, extra=extra, **kwargs)


def read_wav(fname, extra=None):
    with map_utils._h5py().File(fname, 'r') as hfile:
        return wav_from_hdf(hfile, extra=extra)
