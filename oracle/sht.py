'''
Spherical harmonic transforms (oracle). Restates optweight's sht.alm2map
(sht.py:95-178) and sht.map2alm (sht.py:10-93), i.e. ducc0's synthesis and
adjoint_synthesis on a ring geometry, plus default_spin (sht.py:180-215).

Definitions (real fields, m-major triangular alm, see oracle.layout):
  spin 0:  f(theta, phi) = Re[ sum_l a_l0 lam_l0 ] + 2 Re sum_{m>0} sum_l a_lm lam_lm e^{i m phi}
  spin 2:  Q_m = -sum_l (aE F1 + i aB F2),  U_m = -sum_l (aB F1 - i aE F2)
           F1 = (2lam + -2lam)/2, F2 = (2lam - -2lam)/2   (HEALPix convention)
  adjoint: a = Y^T f, exact adjoint under <a,b> = sum_{m=0} Re(a b*) + 2 sum_{m>0} Re(a b*)
           spin 0: a_lm = sum_r lam_lm(theta_r) G_m(r)
           spin 2: aE = -sum_r (F1 q_m + i F2 u_m), aB = -sum_r (F1 u_m - i F2 q_m)
  where G_m(r) = e^{-i m phi0_r} sum_j f_rj e^{-2 pi i m j / nphi_r}.
Ring sums with nphi_r < 2 mmax + 1 alias frequency m onto bin m mod nphi_r,
as ducc0 does for reduced (equal-area) rings. Mirror rings (theta_i +
theta_j = pi within 1e-12) are evaluated at x_S = -x_N with x_N = cos(theta_N)
rounded to float64 (oracle.geometry.ring_x), the libsharp/ducc0 practice of
sharing one Legendre evaluation between the two rings of a symmetric pair.

The Legendre recurrence runs in longdouble (oracle.legendre.lam_m); ring
Fourier sums use numpy.fft.
'''
import numpy as np

from oracle import layout, legendre, geometry


def default_spin(shape):
    if len(shape) > 2:
        raise ValueError('cannot determine default spin')
    if len(shape) == 1 or shape[0] == 1:
        return 0
    if shape[0] == 2:
        return 2
    if shape[0] == 3:
        return [0, 2]
    raise ValueError('cannot determine default spin')


def spin_groups(spin, npol):
    '''(s, i1, i2) triples, the semantics of pixell's enmap.spin_helper:
    spin=[0,2] with npol=3 gives comp 0 at spin 0 and comps 1:3 at spin 2.'''
    spins = np.atleast_1d(spin).tolist()
    out = []
    i = 0
    k = 0
    while i < npol:
        s = spins[min(k, len(spins) - 1)]
        n = 1 if s == 0 else 2
        out.append((s, i, i + n))
        i += n
        k += 1
    if i != npol:
        raise ValueError('spin incompatible with npol')
    return out


def _ring_synth(F, geom, mmax):
    '''F: (ncomp, nrow, mmax+1) complex ring coefficients -> (ncomp, npix).'''
    ncomp = F.shape[0]
    out = np.zeros((ncomp, geom.npix))
    ms = np.arange(mmax + 1)
    for r in range(geom.nrow):
        n = int(geom.nphi[r])
        C = F[:, r, :] * np.exp(1j * ms * geom.phi0[r])
        Z = np.zeros((ncomp, n), dtype=np.complex128)
        Z[:, 0] += C[:, 0].real
        if mmax >= 1:
            np.add.at(Z, (slice(None), ms[1:] % n), C[:, 1:])
            np.add.at(Z, (slice(None), (-ms[1:]) % n), np.conj(C[:, 1:]))
        sl = slice(geom.offsets[r], geom.offsets[r] + n)
        out[:, sl] = (np.fft.ifft(Z, axis=-1) * n).real
    return out


def _ring_adjoint(imap, geom, mmax):
    '''(ncomp, npix) -> G (ncomp, nrow, mmax+1).'''
    ncomp = imap.shape[0]
    G = np.zeros((ncomp, geom.nrow, mmax + 1), dtype=np.complex128)
    ms = np.arange(mmax + 1)
    for r in range(geom.nrow):
        n = int(geom.nphi[r])
        sl = slice(geom.offsets[r], geom.offsets[r] + n)
        X = np.fft.fft(imap[:, sl].astype(np.float64), axis=-1)
        G[:, r, :] = X[:, ms % n] * np.exp(-1j * ms * geom.phi0[r])
    return G


def _synth_group(alm, geom, lmax, s):
    '''alm (n, nelem) complex for one spin group -> map (n, npix).'''
    theta, x = geometry.ring_x(geom)
    F = np.zeros((alm.shape[0], geom.nrow, lmax + 1), dtype=np.complex128)
    if s == 0:
        seeds = legendre.seeds_spin0(lmax, theta)
    for m in range(lmax + 1):
        i0 = layout.mstart(m, lmax) + m
        a = alm[:, i0:i0 + lmax + 1 - m].astype(np.complex128)
        if s == 0:
            lam = legendre.lam_m(m, lmax, theta, 0, seed=seeds[m], x=x).astype(np.float64)
            F[:, :, m] = a @ lam
        else:
            lp = legendre.lam_m(m, lmax, theta, s, x=x).astype(np.float64)
            lm = legendre.lam_m(m, lmax, theta, -s, x=x).astype(np.float64)
            f1 = 0.5 * (lp + lm)
            f2 = 0.5 * (lp - lm)
            aE, aB = a[0], a[1]
            F[0, :, m] = -(aE @ f1 + 1j * (aB @ f2))
            F[1, :, m] = -(aB @ f1 - 1j * (aE @ f2))
    return _ring_synth(F, geom, lmax)


def _adjoint_group(imap, geom, lmax, s):
    theta, x = geometry.ring_x(geom)
    G = _ring_adjoint(imap, geom, lmax)
    out = np.zeros((imap.shape[0], layout.nelem(lmax)), dtype=np.complex128)
    if s == 0:
        seeds = legendre.seeds_spin0(lmax, theta)
    for m in range(lmax + 1):
        i0 = layout.mstart(m, lmax) + m
        n = lmax + 1 - m
        if s == 0:
            lam = legendre.lam_m(m, lmax, theta, 0, seed=seeds[m], x=x).astype(np.float64)
            out[:, i0:i0 + n] = G[:, :, m] @ lam.T
        else:
            lp = legendre.lam_m(m, lmax, theta, s, x=x).astype(np.float64)
            lm = legendre.lam_m(m, lmax, theta, -s, x=x).astype(np.float64)
            f1 = 0.5 * (lp + lm)
            f2 = 0.5 * (lp - lm)
            q, u = G[0, :, m], G[1, :, m]
            out[0, i0:i0 + n] = -(f1 @ q + 1j * (f2 @ u))
            out[1, i0:i0 + n] = -(f1 @ u - 1j * (f2 @ q))
    return out


def _check(alm_n, map_n, lmax, geom, spin):
    if np.asarray(spin).max() > lmax:
        raise ValueError('Spin exceeds lmax')
    if map_n != geom.npix:
        raise ValueError('Wrong map size')
    if alm_n != layout.nelem(lmax):
        raise ValueError('Wrong alm size')


def ringweight(imap, geom, weight=None):
    '''In-place per-ring multiply (map_c_utils.c:7-37).'''
    w = geom.weight if weight is None else weight
    for r in range(geom.nrow):
        sl = slice(geom.offsets[r], geom.offsets[r] + geom.nphi[r])
        imap[..., sl] *= imap.dtype.type(w[r])
    return imap


def alm2map(alm, omap, lmax, geom, spin, adjoint=False):
    '''Y (adjoint=False) or W Y (adjoint=True); writes omap in place.'''
    _check(alm.shape[-1], omap.shape[-1], lmax, geom, spin)
    alm3 = alm.reshape((-1,) + alm.shape[-2:]) if alm.ndim >= 2 else alm.reshape(1, 1, -1)
    om3 = omap.reshape((-1,) + omap.shape[-2:]) if omap.ndim >= 2 else omap.reshape(1, 1, -1)
    if alm3.shape[:-1] != om3.shape[:-1]:
        raise ValueError('Mismatch leading dimension of map and alm')
    npol = om3.shape[-2]
    for t in range(om3.shape[0]):
        for s, i1, i2 in spin_groups(spin, npol):
            m = _synth_group(alm3[t, i1:i2], geom, lmax, s)
            if adjoint:
                ringweight(m, geom)
            om3[t, i1:i2] = m
    return omap


def map2alm(imap, alm, geom, lmax, spin, adjoint=False):
    '''Y^T W (adjoint=False) or Y^T (adjoint=True); writes alm in place.'''
    _check(alm.shape[-1], imap.shape[-1], lmax, geom, spin)
    im3 = imap.reshape((-1,) + imap.shape[-2:]) if imap.ndim >= 2 else imap.reshape(1, 1, -1)
    alm3 = alm.reshape((-1,) + alm.shape[-2:]) if alm.ndim >= 2 else alm.reshape(1, 1, -1)
    if alm3.shape[:-1] != im3.shape[:-1]:
        raise ValueError('Mismatch leading dimension of map and alm')
    npol = im3.shape[-2]
    for t in range(im3.shape[0]):
        for s, i1, i2 in spin_groups(spin, npol):
            m = np.array(im3[t, i1:i2], dtype=np.float64)
            if not adjoint:
                ringweight(m, geom)
            alm3[t, i1:i2] = _adjoint_group(m, geom, lmax, s)
    return alm
