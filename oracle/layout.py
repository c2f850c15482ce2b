'''
alm storage layout (oracle). Triangular, mmax = lmax, m-major, stride 1:
index(l, m) = m (2 lmax + 1 - m) / 2 + l   (alm_c_utils.c:3-5)
nelem = (lmax^2 + 3 lmax) / 2 + 1          (alm_c_utils.c:7-10)
'''
import numpy as np


def nelem(lmax):
    return (lmax * lmax + 3 * lmax) // 2 + 1


def mstart(m, lmax):
    return m * (2 * lmax + 1 - m) // 2


def getlmax(n):
    '''Inverse of nelem (healpy Alm.getlmax semantics, alm_utils.py:36).'''
    lmax = int((-3 + np.sqrt(1 + 8 * n)) / 2)
    if nelem(lmax) != n:
        return -1
    return lmax


def lm_arrays(lmax):
    '''Return (ell, m) index arrays of length nelem in storage order.'''
    ells = []
    ms = []
    for m in range(lmax + 1):
        ells.append(np.arange(m, lmax + 1))
        ms.append(np.full(lmax + 1 - m, m))
    return np.concatenate(ells), np.concatenate(ms)
