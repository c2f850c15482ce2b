'''
2-D Fourier wavelet kernels, mirroring optweight's fkernel (fkernel.py:
9-709): FKernel (one kernel cut to the rectangle of the half plane where
it is nonzero), FKernelSet (the kernels of one transform, iterated in
insertion order), the digitised (on/off) versions of smooth kernels and
the scale-discrete kernels sampled on |l| of a CAR geometry.

A kernel holds the values k[r, c] of rows slices_y[0] then slices_y[1]
and columns slice_x of the full (ny, nx // 2 + 1) half plane. For the
device transforms (wavtrans.f2wav / wav2f) a set packs, per device and
real dtype, a descriptor table (block offset, ny_j, nx_j, first column,
row-table offset, kernel offset and first gather block per scale), the
full-plane row of every kernel row, the inverse row map (ny per scale, -1
outside the slice) and the kernel values.

File output (write / from_hdf) needs h5py.
'''
import os

import numpy as np
import torch
from scipy.interpolate import interp1d

from optweight_amd import dft, wlm_utils

DESC_LEN = 8
GATHER_BLOCK = 256


class FKernel:
    '''One 2-D Fourier kernel (nly, nlx) with wavenumbers ly, lx and its
    place in the full half plane (fkernel.py:205-382).'''

    def __init__(self, fkernel, ly, lx, slices_y=None, slice_x=None, shape_full=None):
        if fkernel.shape[-2:] != (ly.size, lx.size):
            raise ValueError(f'kernel of shape {fkernel.shape[-2:]} for {ly.size} ly and '
                             f'{lx.size} lx wavenumbers')
        self.fkernel = fkernel
        self.ly = ly
        self.lx = lx
        self.slices_y = slices_y
        self.slice_x = slice_x
        self.shape_full = shape_full
        self.lmax = int(np.ceil(np.sqrt(np.max(np.abs(ly)) ** 2 + np.max(np.abs(lx)) ** 2)))

    def modlmap(self, dtype=np.float64):
        return dft.laxes2modlmap(self.ly, self.lx, dtype=dtype)

    @property
    def dtype(self):
        return self.fkernel.dtype

    def to_full_array(self):
        if None in (self.slices_y, self.slice_x, self.shape_full):
            raise ValueError('this kernel has no place in a full plane (slices_y, slice_x '
                             'and shape_full unset)')
        full = np.zeros(self.shape_full, dtype=self.dtype)
        dft.add_to_fmap(full, self.fkernel, self.slices_y, self.slice_x)
        return full

    def rows(self):
        '''Full-plane row of each kernel row (int64, nly).'''
        ny = self.shape_full[0]
        idx = np.arange(ny)
        return np.concatenate((idx[self.slices_y[0]], idx[self.slices_y[1]]))

    def cols(self):
        nlx = self.shape_full[1]
        return np.arange(nlx)[self.slice_x]

    def append_to_hdf(self, hfile):
        if self.shape_full is not None:
            hfile.attrs['shape_full'] = self.shape_full
        for name, sl in (('slice_x', self.slice_x), ('slice_ypos', self.slices_y[0]),
                         ('slice_yneg', self.slices_y[1])):
            for part in ('start', 'stop', 'step'):
                val = getattr(sl, part)
                if val is not None:
                    hfile.attrs[f'{name}_{part}'] = val
        hfile.create_dataset('fkernel', data=self.fkernel)
        hfile.create_dataset('ly', data=self.ly)
        hfile.create_dataset('lx', data=self.lx)

    @classmethod
    def from_hdf(cls, hfile):
        def get_slice(name):
            parts = [hfile.attrs.get(f'{name}_{p}') for p in ('start', 'stop', 'step')]
            if parts.count(None) < 3:
                return slice(*[None if p is None else int(p) for p in parts])
            return None
        slice_x = get_slice('slice_x')
        ypos, yneg = get_slice('slice_ypos'), get_slice('slice_yneg')
        slices_y = None if ypos is None and yneg is None else (ypos, yneg)
        shape_full = hfile.attrs.get('shape_full')
        if shape_full is not None:
            shape_full = tuple(int(s) for s in shape_full)
        return cls(hfile['fkernel'][()], hfile['ly'][()], hfile['lx'][()], slices_y=slices_y,
                   slice_x=slice_x, shape_full=shape_full)


class FKernelSet:
    '''Kernels of one Fourier-wavelet transform keyed by integer scale
    index (fkernel.py:9-203).'''

    def __init__(self):
        self.fkernel_dict = {}
        self.dtype = np.float64
        self.shape_full = None
        self._plans = {}

    def __getitem__(self, key):
        return self.fkernel_dict[int(key)]

    def __setitem__(self, key, fkernel):
        if not isinstance(fkernel, FKernel):
            raise TypeError(f'a kernel set holds FKernel objects, not {type(fkernel).__name__}')
        if int(key) != key:
            raise TypeError(f'kernel index {key!r} is not an integer')
        if not self.fkernel_dict:
            self.dtype = fkernel.dtype
            self.shape_full = fkernel.shape_full
        if fkernel.dtype != self.dtype:
            raise TypeError(f'kernel of dtype {fkernel.dtype} added to a set of {self.dtype}')
        if fkernel.shape_full != self.shape_full:
            raise TypeError(f'kernel on a {fkernel.shape_full} plane added to a set on '
                            f'{self.shape_full}')
        self.fkernel_dict[int(key)] = fkernel
        self._plans = {}

    def __len__(self):
        return len(self.fkernel_dict)

    def __iter__(self):
        return iter(self.fkernel_dict.items())

    def astype(self, dtype, copy=True):
        for key in self.fkernel_dict:
            fk = self.fkernel_dict[key]
            fk.fkernel = fk.fkernel.astype(dtype, copy=copy)
            self.dtype = fk.fkernel.dtype
        self._plans = {}
        return self

    def to_full_array(self):
        if not self.fkernel_dict:
            raise ValueError('the kernel set is empty')
        full = np.zeros((len(self),) + tuple(self.shape_full), dtype=self.dtype)
        for key, fk in self:
            dft.add_to_fmap(full[key], fk.fkernel, fk.slices_y, fk.slice_x)
        return full

    def get_wav_vec(self, preshape, dtype=None, device=None):
        '''Wavelet vector with, per kernel, a zero map (preshape, ny_j,
        2 (nx_j - 1) + 1) flattened on a Clenshaw-Curtis geometry of that
        size (fkernel.py:117-152); device tensors when device is given.'''
        from optweight_amd import wavtrans, map_utils
        dtype = np.dtype(self.dtype if dtype is None else dtype)
        if dtype == np.bool_ or np.issubdtype(dtype, np.integer):
            dtype = np.dtype(np.float64)
        wav = wavtrans.Wav(1, preshape=tuple(preshape), dtype=dtype)
        for widx, fk in self:
            ny_j, nx_j = fk.fkernel.shape
            n_j = 2 * (nx_j - 1) + 1
            minfo = map_utils.get_minfo('CC', ny_j, n_j)
            shape = tuple(preshape) + (minfo.npix,)
            if device is None:
                m = np.zeros(shape, dtype=dtype)
            else:
                m = torch.zeros(shape, dtype=dft._torch_dtype(dtype), device=device)
            wav.add(widx, m, minfo)
        return wav

    def device_plan(self, device, dtype):
        '''Packed device tables of the set for real dtype (cached).'''
        key = (str(device), dtype)
        if key in self._plans:
            return self._plans[key]
        ny, nlx = self.shape_full
        desc, rows, kern = [], [], []
        invrows = np.full((len(self), ny), -1, dtype=np.int32)
        shapes, off, roff, koff, nblk = [], 0, 0, 0, 0
        for j, (widx, fk) in enumerate(self):
            k = np.asarray(fk.fkernel, dtype=np.float64)
            ny_j, nx_j = k.shape
            r = fk.rows()
            c = fk.cols()
            if r.size != ny_j or c.size != nx_j:
                raise ValueError(f'kernel {widx}: slices do not match the kernel shape')
            if nx_j and not np.array_equal(c, np.arange(c[0], c[0] + nx_j)):
                raise NotImplementedError('kernel column slices must have unit step')
            if np.unique(r).size != r.size:
                raise ValueError(f'kernel {widx}: repeated rows')
            invrows[j, r] = np.arange(ny_j, dtype=np.int32)
            desc.append([off, ny_j, nx_j, int(c[0]) if nx_j else 0, roff, koff, nblk, 0])
            nblk += -(-ny_j * nx_j // GATHER_BLOCK)
            shapes.append((widx, off, ny_j, nx_j))
            rows.append(r.astype(np.int32))
            kern.append(k.ravel())
            off += ny_j * nx_j
            roff += ny_j
            koff += ny_j * nx_j
        tdt = dtype
        plan = {
            'desc': torch.as_tensor(np.asarray(desc, dtype=np.int64).reshape(-1, DESC_LEN),
                                    device=device),
            'rows': torch.as_tensor(np.concatenate(rows) if rows else np.zeros(1, np.int32),
                                    device=device),
            'invrows': torch.as_tensor(invrows, device=device),
            'kern': torch.as_tensor(np.concatenate(kern) if kern else np.zeros(1),
                                    device=device).to(tdt),
            'shapes': shapes,
            'total': off,
            'max_count': max([s[2] * s[3] for s in shapes], default=0),
            'nblocks': nblk,
        }
        self._plans[key] = plan
        return plan

    def write(self, fname):
        h5py = _h5py()
        if not os.path.splitext(fname)[1]:
            fname = fname + '.hdf5'
        with h5py.File(fname, 'w') as hfile:
            for widx, fk in self:
                fk.append_to_hdf(hfile.create_group(f'k_{widx}', track_order=True))

    @classmethod
    def from_hdf(cls, fname):
        h5py = _h5py()
        fks = cls()
        with h5py.File(fname, 'r') as hfile:
            for key, val in hfile.items():
                if not isinstance(val, h5py.Group):
                    raise ValueError(f'{fname}: entry {key!r} is not a kernel group')
                fks[int(key.split('_')[-1])] = FKernel.from_hdf(val)
        return fks


def _h5py():
    try:
        import h5py
    except ImportError as exc:
        raise ImportError('reading and writing kernel files needs h5py') from exc
    return h5py


def digitize_1d(arr):
    '''On/off approximation of a smooth kernel with values in [0, 1]
    (fkernel.py:384-410): sample i is on when the rounded running sum of
    the kernel steps there, so the count of samples that are on tracks the
    kernel's integral.'''
    arr = np.asarray(arr)
    if arr.min() < 0 or arr.max() > 1:
        raise ValueError(f'a kernel to digitise must lie in [0, 1]; got values in '
                         f'[{arr.min():.3g}, {arr.max():.3g}]')
    return np.diff(np.round(np.cumsum(arr)), prepend=0) != 0


def _check_disjoint_non_neighbours(w_ell, eps=1e-8):
    '''Kernels two apart must not overlap: within the even kernels and within
    the odd ones (groups of two or more), the product of the group stays
    below eps at every l.'''
    for parity in (0, 1):
        group = w_ell[parity::2]
        if group.shape[0] > 1 and not np.all(np.prod(group, axis=0) < eps):
            raise ValueError(f'kernels {parity}, {parity + 2}, ... overlap: only neighbouring '
                             f'kernels may share multipoles')


def digitize_kernels(w_ell, upsamp_fact=10):
    '''Digital (bool) kernels on ells_fine = linspace(0, lmax + 1,
    upsamp_fact (lmax + 1)) (fkernel.py:412-467): the kernels are linearly
    interpolated there; the even ones are digitised, the odd ones are on
    wherever no even one is and their smooth value is at least 1e-5.
    Returns (d_fine (nwav, nfine) bool, ells_fine).'''
    w_ell = np.asarray(w_ell)
    _check_disjoint_non_neighbours(w_ell)
    nell = w_ell.shape[-1]
    ells_fine = np.linspace(0, nell, upsamp_fact * nell)
    w_fine = interp1d(np.arange(nell), w_ell, kind='linear', fill_value='extrapolate',
                      axis=-1)(ells_fine)
    even = np.array([digitize_1d(w) for w in w_fine[0::2]], dtype=bool)
    d_fine = np.empty(w_fine.shape, dtype=bool)
    d_fine[0::2] = even
    d_fine[1::2] = ~even.any(axis=0) & ~(w_fine[1::2] < 1e-5)
    return d_fine, ells_fine


def _prefix_len(flags):
    '''Length of the shortest prefix of ``flags`` holding all its True
    entries (0 if there is none).'''
    on = np.flatnonzero(flags)
    return int(on[-1]) + 1 if on.size else 0


def find_kernel_slice(fkernel_arr, minval=1e-7, optimize_len=False):
    '''Support of a kernel on the half plane (ny, nlx) (fkernel.py:503-598):
    ((rows from ly = 0 up, rows of negative ly), columns), holding every
    value >= minval. Columns start at lx = 0; the rows reach equally far
    from ly = 0 on both sides (row 0 always kept). With optimize_len the
    row count grows to an FFT-friendly length (positive side half of the
    extra, negative side the rest) when that length is below ny.'''
    fkernel_arr = np.asarray(fkernel_arr)
    if fkernel_arr.ndim != 2:
        raise ValueError(f'a kernel is a 2-D array, got {fkernel_arr.ndim} dimensions')
    support = fkernel_arr >= minval
    ncol = _prefix_len(support.any(axis=0))
    if ncol == 0:
        raise ValueError(f'the kernel has no value >= {minval}')
    rows = support.any(axis=1)
    ny = rows.size
    mid = -(-ny // 2)
    above, below = rows[1:mid], rows[mid:][::-1]
    reach = max(_prefix_len(above), _prefix_len(below))
    npos = 1 + min(reach, above.size)
    nneg = min(reach, below.size)
    if optimize_len:
        nopt = dft.get_optimal_fftlen(npos + nneg, even=False)
        if nopt < ny:
            extra = nopt - npos - nneg
            npos += extra // 2
            nneg += extra - extra // 2
    slice_yneg = slice(-nneg, None) if nneg else slice(0, 0)
    return (slice(0, npos), slice_yneg), slice(0, ncol)


def w_ell2fkernels(w_ell, ells, ly, lx, interp_kind, optimize_len=False):
    '''Kernels w_ell (nwav, nell) at ells sampled on |l| of the half plane
    and cut to their support (fkernel.py:600-651); beyond ells the last
    kernel is 1 and the others 0.'''
    w_ell = np.asarray(w_ell)
    nwav = w_ell.shape[0]
    modlmap = dft.laxes2modlmap(ly, lx)
    fks = FKernelSet()
    for idx in range(nwav):
        fill = (0., 1.) if idx == nwav - 1 else (0., 0.)
        cs = interp1d(ells, w_ell[idx], kind=interp_kind, bounds_error=False, fill_value=fill)
        full = cs(modlmap)
        slices_y, slice_x = find_kernel_slice(full, optimize_len=optimize_len)
        cut, laxes = dft.slice_fmap(full, slices_y, slice_x, laxes=(ly, lx))
        fks[idx] = FKernel(cut, *laxes, slices_y=slices_y, slice_x=slice_x,
                           shape_full=(ly.size, lx.size))
    return fks


def get_sd_kernels_fourier(ly, lx, lamb, j0=None, lmin=None, jmax=None, lmax_j=None,
                           digital=True, oversample=10):
    '''Scale-discrete kernels on the half plane (fkernel.py:653-709);
    digital kernels (bool) come from kernels upsampled to about
    `oversample` samples per Fourier spacing.'''
    lmax = int(np.ceil(np.sqrt(np.max(np.abs(ly)) ** 2 + np.max(np.abs(lx)) ** 2)))
    w_ell, _ = wlm_utils.get_sd_kernels(lamb, lmax, j0=j0, lmin=lmin, jmax=jmax, lmax_j=lmax_j)
    if digital:
        delta_ell = np.mean([np.abs(lx[0] - lx[1]), np.abs(ly[0] - ly[1])])
        upsamp_fact = max(1, int(np.round(oversample / delta_ell)))
        w_ell, ells = digitize_kernels(w_ell, upsamp_fact=upsamp_fact)
        dtype = bool
    else:
        ells = np.arange(w_ell.shape[-1])
        dtype = w_ell.dtype
    return w_ell2fkernels(w_ell, ells, ly, lx, 'nearest', optimize_len=True).astype(
        dtype, copy=False)
