'''
Map, map-info and wavelet-vector files (map_utils.write_map / read_map /
write_minfo / read_minfo, wavtrans.write_wav / read_wav; reference
map_utils.py:1902-2101, wavtrans.py:548-686) on the host: the dataset and
attribute layout and the round trips, through an in-memory stand-in for an
HDF5 group (h5py is not installed here; the file-level functions need it
and are checked when it is), and the upper-triangular packing of symmetric
maps against the literal answers of test_mat_utils.py:351-440.
'''
import numpy as np
import pytest

from optweight_amd import map_utils, mat_utils, wavtrans


class _Dataset:
    def __init__(self, data):
        self._data = np.array(data, copy=True)

    def __getitem__(self, key):
        assert key == ()
        return self._data.copy()


class MemGroup:
    '''The subset of h5py.Group the readers and writers use.'''

    def __init__(self):
        self.attrs = {}
        self._items = {}

    def create_dataset(self, name, data):
        assert name not in self._items
        self._items[name] = _Dataset(data)
        return self._items[name]

    def create_group(self, name):
        assert name not in self._items
        self._items[name] = MemGroup()
        return self._items[name]

    def __getitem__(self, name):
        return self._items[name]

    def keys(self):
        return sorted(self._items)


def test_triu_packing_known_answers():
    mat = np.arange(9, dtype=np.float32).reshape(3, 3)
    triu = mat_utils.symm2triu(mat, [0, 1])
    np.testing.assert_array_equal(triu, [0, 1, 2, 4, 5, 8])
    assert triu.dtype == np.float32 and not np.shares_memory(triu, mat)
    np.testing.assert_array_equal(mat_utils.triu2symm(np.asarray([0, 1, 2, 4, 5, 8.]), 0),
                                  [[0, 1, 2], [1, 4, 5], [2, 5, 8]])
    m = np.ones((4, 2, 2, 5), dtype=np.float32) * np.asarray([[0, 1], [1, 3]])[None, :, :, None]
    for axes in ([1, 2], [-2, -3], [-3, -2], [-3, 2]):
        np.testing.assert_array_equal(
            mat_utils.symm2triu(m, axes),
            np.ones((4, 3, 5)) * np.asarray([0, 1, 3])[None, :, None])
    with pytest.raises(ValueError):
        # not adjacent
        mat_utils.symm2triu(np.ones((3, 2, 2, 4)) * np.arange(4.).reshape(2, 2)[None, :, :, None],
                            [0, 2])
    with pytest.raises(ValueError):
        # not N x N
        mat_utils.symm2triu(np.ones((3, 2, 1, 4)), [0, 2])
    with pytest.raises(ValueError):
        mat_utils.triu2symm(np.ones(5), 0)
    v, flat = mat_utils.flattened_view(np.zeros((3, 2, 5, 10)), [[0, 1], [3]],
                                       return_flat_axes=True)
    assert v.shape == (6, 5, 10) and flat == [0, 2]


def test_map_and_minfo_round_trip():
    minfo = map_utils.MapInfo.map_info_gauss_legendre(6, 11)
    rng = np.random.default_rng(0)
    imap = rng.normal(size=(3, minfo.npix))
    g = MemGroup()
    map_utils.append_map_to_hdf(g, imap, minfo)
    assert g.keys() == ['map', 'minfo']
    assert g['minfo'].keys() == ['nphi', 'offsets', 'phi0', 'stride', 'theta', 'weight']
    omap, mi = map_utils.map_from_hdf(g)
    np.testing.assert_array_equal(omap, imap)
    for name in ('theta', 'weight', 'nphi', 'phi0', 'offsets', 'stride'):
        np.testing.assert_array_equal(getattr(mi, name), getattr(minfo, name))


def test_symmetric_map_is_stored_as_upper_triangle():
    minfo = map_utils.MapInfo.map_info_gauss_legendre(4, 7)
    rng = np.random.default_rng(1)
    a = rng.normal(size=(3, 3, minfo.npix))
    cov = a + a.transpose(1, 0, 2)
    g = MemGroup()
    map_utils.append_map_to_hdf(g, cov, minfo, symm_axes=[0, 1])
    assert g['map'][()].shape == (6, minfo.npix)
    assert g.attrs['symm_axis'] == 0 and tuple(g.attrs['shape_orig']) == cov.shape
    np.testing.assert_array_equal(map_utils.map_from_hdf(g)[0], cov)
    b = rng.normal(size=(2, 3, 2, 3, minfo.npix))
    b = b + b.transpose(2, 3, 0, 1, 4)
    g2 = MemGroup()
    map_utils.append_map_to_hdf(g2, b, minfo, symm_axes=[[0, 1], [2, 3]])
    assert g2['map'][()].shape == (21, minfo.npix)
    np.testing.assert_array_equal(map_utils.map_from_hdf(g2)[0], b)


def test_wav_round_trip():
    rng = np.random.default_rng(2)
    wav = wavtrans.Wav(2, dtype=np.float32)
    for j in range(3):
        mi = map_utils.MapInfo.map_info_gauss_legendre(j + 2, 2 * j + 3)
        m = rng.normal(size=(2, 2, mi.npix)).astype(np.float32)
        wav.add((j, j), m + m.transpose(1, 0, 2), mi)
    g = MemGroup()
    wavtrans.append_wav_to_hdf(g, wav, extra={'w_ell': np.ones((3, 5))}, symm_axes=[0, 1])
    assert g.keys() == ['m_0_0', 'm_1_1', 'm_2_2', 'w_ell']
    assert g.attrs['dtype'] == 'f' and g.attrs['ndim'] == 2
    out, extra = wavtrans.wav_from_hdf(g, extra=['w_ell'])
    np.testing.assert_array_equal(extra['w_ell'], np.ones((3, 5)))
    assert out.ndim == 2 and out.dtype == np.float32 and out.preshape == (2, 2)
    for key in wav.maps:
        np.testing.assert_array_equal(out.maps[key], wav.maps[key])
        np.testing.assert_array_equal(out.minfos[key].theta, wav.minfos[key].theta)
    vec = wavtrans.Wav(1, dtype=np.float64)
    mi = map_utils.MapInfo.map_info_gauss_legendre(3, 5)
    vec.add(0, rng.normal(size=(1, mi.npix)), mi)
    g = MemGroup()
    wavtrans.append_wav_to_hdf(g, vec)
    assert g.keys() == ['m_0']
    np.testing.assert_array_equal(wavtrans.wav_from_hdf(g).maps[0], vec.maps[0])


def test_file_functions(tmp_path):
    try:
        import h5py  # noqa: F401
    except ImportError:
        with pytest.raises(ImportError):
            map_utils.write_minfo(str(tmp_path / 'minfo'), map_utils.get_minfo('CC', 3, 4))
        pytest.skip('h5py not installed: file round trips not run')
    minfo = map_utils.get_minfo('CC', 5, 8)
    imap = np.random.default_rng(3).normal(size=(2, minfo.npix))
    map_utils.write_map(str(tmp_path / 'm'), imap, minfo)
    omap, mi = map_utils.read_map(str(tmp_path / 'm.hdf5'))
    np.testing.assert_array_equal(omap, imap)
    np.testing.assert_array_equal(mi.theta, minfo.theta)
