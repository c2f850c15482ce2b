'''
Normalised spin-weighted associated Legendre functions (oracle).

    sY_lm(theta, phi) = slam_lm(theta) exp(i m phi),
    slam_lm(theta)    = (-1)^s sqrt((2l+1)/(4 pi)) d^l_{m,-s}(theta),

with Wigner's d in the convention where d^1_{1,0}(b) = -sin(b)/sqrt(2), so
that s=0 gives the Condon-Shortley orthonormal Y_lm used by ducc0/HEALPix
(the functions behind ducc0.sht.experimental.synthesis, called at
sht.py:78-93,159-174 of optweight).

Two independent evaluations:
  * ``lam_exact``: Wigner's explicit finite sum, in mpmath (40 digits).
  * ``lam_m``: three-term recurrence in l at fixed (m, s), in numpy
    longdouble (x87 80-bit: 64-bit mantissa, exponent range 1e+-4932), so
    sin^m(theta) never underflows for lmax <= 8191 and no rescaling trick is
    needed here (the HIP kernels use one; this oracle deliberately does not).

Recurrence (d^l_{m m'} with m' = -s, r_l = sqrt((l^2-m^2)(l^2-m'^2))):
    slam_{l+1} = (alpha_l x + beta_l) slam_l - gamma_l slam_{l-1}
    alpha_l = sqrt((2l+3)(2l+1)) (l+1) / r_{l+1}
    beta_l  = -sqrt((2l+3)(2l+1)) m m' / (l r_{l+1})
    gamma_l = sqrt((2l+3)/(2l-1)) (l+1) r_l / (l r_{l+1})
'''
import numpy as np
import mpmath

LD = np.longdouble


def wigner_d_exact(j, mp, m, beta, dps=40):
    '''d^j_{mp,m}(beta) from Wigner's explicit sum, mpmath precision.'''
    with mpmath.workdps(dps):
        b = mpmath.mpf(beta)
        c = mpmath.cos(b / 2)
        s = mpmath.sin(b / 2)
        pref = mpmath.sqrt(mpmath.factorial(j + mp) * mpmath.factorial(j - mp)
                           * mpmath.factorial(j + m) * mpmath.factorial(j - m))
        tot = mpmath.mpf(0)
        for k in range(0, 2 * j + 1):
            a1 = j + m - k
            a2 = mp - m + k
            a3 = j - mp - k
            if a1 < 0 or a2 < 0 or a3 < 0:
                continue
            term = (-1) ** (mp - m + k) * pref / (
                mpmath.factorial(a1) * mpmath.factorial(k)
                * mpmath.factorial(a2) * mpmath.factorial(a3))
            term *= c ** (2 * j + m - mp - 2 * k) * s ** (mp - m + 2 * k)
            tot += term
        return tot


def lam_exact(l, m, s, theta, dps=40):
    '''slam_lm(theta) from the exact sum (0 if l < max(|m|,|s|)).'''
    l, m, s = int(l), int(m), int(s)
    if l < max(abs(m), abs(s)):
        return mpmath.mpf(0)
    with mpmath.workdps(dps):
        return ((-1) ** s * mpmath.sqrt(mpmath.mpf(2 * l + 1) / (4 * mpmath.pi))
                * wigner_d_exact(l, m, -s, theta, dps=dps))


def _wigner_d_ld(j, mp, m, theta):
    '''Wigner's sum in longdouble, for the l=2 seeds of spin 2 at m < 2.'''
    from math import factorial
    c = np.cos(theta / LD(2))
    s = np.sin(theta / LD(2))
    pref = np.sqrt(LD(factorial(j + mp) * factorial(j - mp)
                      * factorial(j + m) * factorial(j - m)))
    tot = np.zeros_like(c)
    for k in range(0, 2 * j + 1):
        a1, a2, a3 = j + m - k, mp - m + k, j - mp - k
        if a1 < 0 or a2 < 0 or a3 < 0:
            continue
        den = LD(factorial(a1) * factorial(k) * factorial(a2) * factorial(a3))
        tot += ((-1) ** (mp - m + k) * pref / den
                * c ** (2 * j + m - mp - 2 * k) * s ** (mp - m + 2 * k))
    return tot


def seeds_spin0(mmax, theta):
    '''slam_mm for m = 0..mmax, shape (mmax+1, ntheta), longdouble.
    slam_00 = 1/sqrt(4pi), slam_mm = -sqrt((2m+1)/(2m)) sin(theta) slam_{m-1,m-1}.'''
    th = np.asarray(theta, dtype=LD)
    st = np.sin(th)
    out = np.zeros((mmax + 1, th.size), dtype=LD)
    out[0] = LD(1) / np.sqrt(LD(4) * LD(np.pi))
    for m in range(1, mmax + 1):
        out[m] = -np.sqrt(LD(2 * m + 1) / LD(2 * m)) * st * out[m - 1]
    return out


def seed_spin(m, s, theta, pref=None):
    '''slam_lm at l = lstart = max(m, |s|) for s != 0 (longdouble, theta a
    longdouble array). pref may pass sqrt(binom(2m, m+s)) for m >= |s|.'''
    th = np.asarray(theta, dtype=LD)
    if m >= abs(s):
        # d^m_{m,-s} = (-1)^{m+s} sqrt((2m)!/((m-s)!(m+s)!)) c^{m-s} sn^{m+s}
        c = np.cos(th / LD(2))
        sn = np.sin(th / LD(2))
        if pref is None:
            # sqrt((2m)! / ((m-s)! (m+s)!)) = sqrt(binom(2m, m+s)), as a
            # product to keep longdouble precision
            pref = LD(1)
            k = m + s
            for i in range(1, k + 1):
                pref *= LD(2 * m - k + i) / LD(i)
            pref = np.sqrt(pref)
        return ((-1) ** m * np.sqrt(LD(2 * m + 1) / (LD(4) * LD(np.pi)))
                * pref * c ** (m - s) * sn ** (m + s))
    lstart = abs(s)
    return ((-1) ** s * np.sqrt(LD(2 * lstart + 1) / (LD(4) * LD(np.pi)))
            * _wigner_d_ld(lstart, m, -s, th))


def seed_tables(lmax, theta, s):
    '''Seeds slam at l = max(m, |s|) for m = 0..lmax on the given colatitudes
    as float64 mantissas and int32 exponents (value = mant * 2^exp), shape
    (nfunc, lmax + 1, ntheta): one function for s = 0, two (+|s|, -|s|) for
    spin |s| > 0. The seeds are the longdouble values ``lam_m`` starts from.'''
    th = np.asarray(theta, dtype=LD)
    if s == 0:
        vals = [seeds_spin0(lmax, th)]
    else:
        vals = []
        for sg in (abs(s), -abs(s)):
            v = np.zeros((lmax + 1, th.size), dtype=LD)
            # sqrt(binom(2m, m + sg)) for m >= |s|, by its ratio in m
            pref2 = LD(1)    # binom(2|s|, |s| + sg) = 1 for both signs
            for m in range(lmax + 1):
                if m < abs(s):
                    v[m] = seed_spin(m, sg, th)
                    continue
                if m > abs(s):
                    pref2 *= LD((2 * m) * (2 * m - 1)) / LD((m + sg) * (m - sg))
                v[m] = seed_spin(m, sg, th, pref=np.sqrt(pref2))
            vals.append(v)
    mant, expo = np.frexp(np.stack(vals))
    return mant.astype(np.float64), expo.astype(np.int32)


def lam_m(m, lmax, theta, s, seed=None, x=None):
    '''slam_lm(theta) for l = m..lmax at fixed m, spin s (0, +2 or -2).

    Returns longdouble array (lmax + 1 - m, ntheta); rows with l < |s| are 0.
    ``seed`` may pass slam_mm for s = 0 (from ``seeds_spin0``). ``x`` may
    pass the cos(theta) values the recurrence uses (default cos(theta)); the
    seeds are always evaluated from theta.
    '''
    th = np.asarray(theta, dtype=LD)
    x = np.cos(th) if x is None else np.asarray(x, dtype=LD)
    nl = lmax + 1 - m
    out = np.zeros((nl, th.size), dtype=LD)
    if nl <= 0:
        return out
    mp = -s
    lstart = max(m, abs(s))
    if lstart > lmax:
        return out
    if s == 0:
        if seed is None:
            seed = seeds_spin0(m, theta)[m]
        first = seed
    else:
        first = seed_spin(m, s, th)
    out[lstart - m] = first
    lam_prev2 = np.zeros_like(x)
    lam_prev = first
    m2 = LD(m * m)
    mp2 = LD(mp * mp)
    for l in range(lstart, lmax):
        L = LD(l)
        r_l = np.sqrt((L * L - m2) * (L * L - mp2))
        r_l1 = np.sqrt(((L + 1) ** 2 - m2) * ((L + 1) ** 2 - mp2))
        alpha = np.sqrt((2 * L + 3) * (2 * L + 1)) * (L + 1) / r_l1
        if m * mp != 0:
            beta = -np.sqrt((2 * L + 3) * (2 * L + 1)) * LD(m * mp) / (L * r_l1)
        else:
            beta = LD(0)
        if l == lstart or l == 0:
            gamma = LD(0)
        else:
            gamma = (np.sqrt((2 * L + 3) / (2 * L - 1)) * (L + 1) * r_l
                     / (L * r_l1))
        lam = (alpha * x + beta) * lam_prev - gamma * lam_prev2
        out[l + 1 - m] = lam
        lam_prev2 = lam_prev
        lam_prev = lam
    return out


def f1f2_healpix(l, m, theta):
    '''HEALPix/Zaldarriaga-Seljak closed form of the spin-2 functions,
    F1 = (2lam + -2lam)/2 and F2 = (2lam - -2lam)/2, written with the
    Condon-Shortley P_l^m (an independent check of the Wigner-d route):
      F1 = N [ -((l - m^2)/s^2 + l(l-1)/2) P_l^m + (l+m) c/s^2 P_{l-1}^m ]
      F2 = N m/s^2 [ -(l-1) c P_l^m + (l+m) P_{l-1}^m ]
      N  = 2 sqrt((2l+1)/(4pi) (l-2)!/(l+2)! (l-m)!/(l+m)!)
    Valid for l >= 2, 0 < theta < pi.'''
    with mpmath.workdps(40):
        th = mpmath.mpf(theta)
        c = mpmath.cos(th)
        s2 = mpmath.sin(th) ** 2
        P = lambda ll: mpmath.legenp(ll, m, c) if ll >= m else mpmath.mpf(0)
        N = 2 * mpmath.sqrt(mpmath.mpf(2 * l + 1) / (4 * mpmath.pi)
                            * mpmath.factorial(l - 2) / mpmath.factorial(l + 2)
                            * mpmath.factorial(l - m) / mpmath.factorial(l + m))
        F1 = N * (-((l - m * m) / s2 + mpmath.mpf(l * (l - 1)) / 2) * P(l)
                  + (l + m) * c / s2 * P(l - 1))
        F2 = N * m / s2 * (-(l - 1) * c * P(l) + (l + m) * P(l - 1))
        return F1, F2
