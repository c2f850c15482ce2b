/*
 * Oracle (test infrastructure, CPU): the Legendre stage of the spherical
 * harmonic transform pair in float64 with OpenMP threads over m. It
 * restates the same definitions as oracle/sht.py (optweight sht.py:10-178,
 * ducc0's synthesis / adjoint_synthesis) and is pinned against that
 * long-double numpy restatement by tests/test_oracle_fast.py. It exists so
 * that the oracle can run full-size problems (lmax 1023) in seconds: the
 * bench's CPU baseline and the full-size golden fixtures.
 *
 * Functions: the normalised spin-weighted associated Legendre functions
 * s_lambda_lm(theta) = (-1)^s sqrt((2l+1)/4pi) d^l_{m,-s}(theta), generated
 * for l >= lstart = max(m, |s|) by the three-term recurrence
 *   lam_{l+1} = (alpha_l x + beta_l) lam_l - gamma_l lam_{l-1}
 *   alpha_l = sqrt((2l+3)(2l+1)) (l+1) / r_{l+1},
 *   beta_l  = -sqrt((2l+3)(2l+1)) m mp / (l r_{l+1}),
 *   gamma_l = sqrt((2l+3)/(2l-1)) (l+1) r_l / (l r_{l+1}),
 *   r_l = sqrt((l^2 - m^2)(l^2 - mp^2)),  mp = -s
 * (the recurrence of oracle/legendre.py lam_m). The seed at lstart comes
 * from the caller as mantissa * 2^exponent (it underflows float64 near the
 * poles at large m); the recurrence runs on a scaled value (times 2^-E)
 * until the function exceeds 2^-900 and plainly from there on; terms below
 * 2^-900 of the function's O(1) maximum are dropped.
 *
 * Mirror rings (theta_S = pi - theta_N) share one evaluation:
 *   spin 0:  lam(pi - theta) = (-1)^(l+m) lam(theta)
 *   spin 2:  F1(pi - theta) = (-1)^(l+m) F1(theta),
 *            F2(pi - theta) = -(-1)^(l+m) F2(theta),
 *            F1 = (2lam + -2lam) / 2, F2 = (2lam - -2lam) / 2.
 * Spin 2 (HEALPix convention): Q_m = -sum_l (aE F1 + i aB F2),
 * U_m = -sum_l (aB F1 - i aE F2); adjoint aE = -sum_r (F1 q + i F2 u),
 * aB = -sum_r (F1 u - i F2 q).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define SCALE_UP 300
#define START_EXP (-900)

typedef struct {
  double a, b, g;
} coef_t;

static long long mstart_of(int m, int lmax) { return (long long)m * (2LL * lmax + 1 - m) / 2; }

/* Recurrence coefficients producing lam_{l+1} for lstart <= l < lmax. */
static void make_coefs(int m, int mp, int lstart, int lmax, coef_t *c) {
  const double m2 = (double)m * m, mp2 = (double)mp * mp;
  for (int l = lstart; l < lmax; ++l) {
    const double L = l;
    const double rl = sqrt((L * L - m2) * (L * L - mp2));
    const double rl1 = sqrt(((L + 1) * (L + 1) - m2) * ((L + 1) * (L + 1) - mp2));
    const double sq = sqrt((2 * L + 3) * (2 * L + 1));
    c[l].a = sq * (L + 1) / rl1;
    c[l].b = (m * mp != 0) ? -sq * (double)(m * mp) / (L * rl1) : 0.0;
    c[l].g = (l == lstart || l == 0) ? 0.0 : sqrt((2 * L + 3) / (2 * L - 1)) * (L + 1) * rl / (L * rl1);
  }
}

/*
 * Advance one function from its seed (mant * 2^e at lstart) to the first l
 * where |lam_l| >= 2^START_EXP; returns that l (lmax + 1 if none) and the
 * float64 values lam_{l-1}, lam_l there.
 */
static int start_fn(const coef_t *cf, int lstart, int lmax, double x, double mant, int e,
                    double *pv, double *cv) {
  double p = 0.0, c = mant;
  int E = e, l = lstart;
  for (;;) {
    if (c != 0.0 && E + ilogb(c) >= START_EXP) {
      *pv = ldexp(p, E);
      *cv = ldexp(c, E);
      return l;
    }
    if (l >= lmax) return lmax + 1;
    const double n = (cf[l].a * x + cf[l].b) * c - cf[l].g * p;
    p = c;
    c = n;
    ++l;
    const double ac = fabs(c);
    if (ac > 0x1p300) {
      c = ldexp(c, -SCALE_UP);
      p = ldexp(p, -SCALE_UP);
      E += SCALE_UP;
    } else if (ac < 0x1p-300 && ac > 0.0) {
      c = ldexp(c, SCALE_UP);
      p = ldexp(p, SCALE_UP);
      E -= SCALE_UP;
    }
  }
}

/* Plain recurrence step from l to l + 1. */
static inline void step(const coef_t *cf, int l, double x, double *p, double *c) {
  const double n = (cf[l].a * x + cf[l].b) * (*c) - cf[l].g * (*p);
  *p = *c;
  *c = n;
}

/*
 * Ring pairs are processed NB at a time: their recurrences are independent
 * chains, interleaved for instruction-level parallelism. A pair joins the
 * block's sums at its own start multipole (before it, its values are held
 * at zero).
 */
#define NB 8

typedef struct {
  int l0;          /* first accumulated multipole (lmax + 1: none) */
  double p, c;     /* lam_{l0 - 1}, lam_{l0} */
} start_t;

/* Start of one function at every pair of a block. */
static int block_starts(const coef_t *cf, int lstart, int lmax, const double *x,
                        const double *sm, const int *se, int nb, start_t *st) {
  int lmin = lmax + 1;
  for (int b = 0; b < nb; ++b) {
    st[b].l0 = start_fn(cf, lstart, lmax, x[b], sm[b], se[b], &st[b].p, &st[b].c);
    if (st[b].l0 < lmin) lmin = st[b].l0;
  }
  return lmin;
}

/* Spin 2: both functions of a pair from a common multipole (the later of
 * their two starts; the earlier one is stepped on to it). */
static void align_pair(const coef_t *c0, const coef_t *c1, int lmax, double x, start_t *a,
                       start_t *b) {
  while (a->l0 < b->l0 && a->l0 <= lmax) {
    if (a->l0 < lmax) step(c0, a->l0, x, &a->p, &a->c);
    ++a->l0;
  }
  while (b->l0 < a->l0 && b->l0 <= lmax) {
    if (b->l0 < lmax) step(c1, b->l0, x, &b->p, &b->c);
    ++b->l0;
  }
}

/*
 * Synthesis of one spin group: F[comp][ring][m] (complex) from alm[comp]
 * (complex, m-major triangular at lmax, components nelem apart).
 * Pairs: north[k] (ring index), south[k] (mirror ring or -1), x[k] =
 * cos(theta_north). Seeds: seed_m / seed_e [f][m][k], f = 0 for spin 0,
 * f = 0 (s = +2, mp = -2) and f = 1 (s = -2, mp = +2) for spin 2.
 * F has nring rows of (lmax + 1) complex per component; every entry of
 * the given rings is written.
 */
void ol_legendre_synth(int lmax, int spin, int npair, const int *north, const int *south,
                       const double *x, const double *seed_m, const int *seed_e,
                       const double *alm, int nring, double *F, int nthreads) {
  const long long nelem = (long long)(lmax + 1) * (lmax + 2) / 2;
  const long long fstride = (long long)nring * (lmax + 1);   /* complex per component */
  const long long sstride = (long long)(lmax + 1) * npair;    /* per seed function */
#pragma omp parallel num_threads(nthreads)
  {
    coef_t *cf0 = malloc(sizeof(coef_t) * (lmax + 1));
    coef_t *cf1 = malloc(sizeof(coef_t) * (lmax + 1));
#pragma omp for schedule(dynamic, 1)
    for (int m = 0; m <= lmax; ++m) {
      const long long i0 = mstart_of(m, lmax);   /* alm index of (l, m) is i0 + l */
      const int ls = spin == 0 ? m : (m > 2 ? m : 2);
      make_coefs(m, spin == 0 ? 0 : -2, ls, lmax, cf0);
      if (spin != 0) make_coefs(m, 2, ls, lmax, cf1);
      for (int k0 = 0; k0 < npair; k0 += NB) {
        const int nb = npair - k0 < NB ? npair - k0 : NB;
        const double *xb = x + k0;
        const long long si = (long long)m * npair + k0;
        start_t s0[NB], s1[NB];
        int lmin = block_starts(cf0, ls, lmax, xb, seed_m + si, seed_e + si, nb, s0);
        if (spin != 0) {
          block_starts(cf1, ls, lmax, xb, seed_m + sstride + si, seed_e + sstride + si, nb, s1);
          lmin = lmax + 1;
          for (int b = 0; b < nb; ++b) {
            align_pair(cf0, cf1, lmax, xb[b], &s0[b], &s1[b]);
            if (s0[b].l0 < lmin) lmin = s0[b].l0;
          }
        }
        double p0[NB] = {0}, c0[NB] = {0}, p1[NB] = {0}, c1[NB] = {0};
        /* sums [term][parity][re, im][pair]: spin 0 uses term 0 (a lam);
           spin 2 terms aE F1, aB F2, aB F1, aE F2 */
        double acc[4][2][2][NB];
        memset(acc, 0, sizeof(acc));
        for (int l = lmin; l <= lmax; ++l) {
          for (int b = 0; b < nb; ++b)
            if (l == s0[b].l0) {
              p0[b] = s0[b].p;
              c0[b] = s0[b].c;
              if (spin != 0) {
                p1[b] = s1[b].p;
                c1[b] = s1[b].c;
              }
            }
          const int par = (l - m) & 1;
          const double *ae = alm + 2 * (i0 + l);
          if (spin == 0) {
            for (int b = 0; b < NB; ++b) {
              acc[0][par][0][b] += ae[0] * c0[b];
              acc[0][par][1][b] += ae[1] * c0[b];
            }
          } else {
            const double *ab = alm + 2 * (nelem + i0 + l);
            for (int b = 0; b < NB; ++b) {
              const double f1 = 0.5 * (c0[b] + c1[b]), f2 = 0.5 * (c0[b] - c1[b]);
              acc[0][par][0][b] += ae[0] * f1;
              acc[0][par][1][b] += ae[1] * f1;
              acc[1][par][0][b] += ab[0] * f2;
              acc[1][par][1][b] += ab[1] * f2;
              acc[2][par][0][b] += ab[0] * f1;
              acc[2][par][1][b] += ab[1] * f1;
              acc[3][par][0][b] += ae[0] * f2;
              acc[3][par][1][b] += ae[1] * f2;
            }
          }
          if (l < lmax) {
            const double A0 = cf0[l].a, B0 = cf0[l].b, G0 = cf0[l].g;
            for (int b = 0; b < NB; ++b) {
              const double n = (A0 * xb[b < nb ? b : 0] + B0) * c0[b] - G0 * p0[b];
              p0[b] = c0[b];
              c0[b] = n;
            }
            if (spin != 0) {
              const double A1 = cf1[l].a, B1 = cf1[l].b, G1 = cf1[l].g;
              for (int b = 0; b < NB; ++b) {
                const double n = (A1 * xb[b < nb ? b : 0] + B1) * c1[b] - G1 * p1[b];
                p1[b] = c1[b];
                c1[b] = n;
              }
            }
          }
        }
        for (int b = 0; b < nb; ++b) {
          const int k = k0 + b;
          if (spin == 0) {
            double *fn = F + 2 * ((long long)north[k] * (lmax + 1) + m);
            fn[0] = acc[0][0][0][b] + acc[0][1][0][b];
            fn[1] = acc[0][0][1][b] + acc[0][1][1][b];
            if (south[k] >= 0) {
              double *fs = F + 2 * ((long long)south[k] * (lmax + 1) + m);
              fs[0] = acc[0][0][0][b] - acc[0][1][0][b];
              fs[1] = acc[0][0][1][b] - acc[0][1][1][b];
            }
            continue;
          }
          /* north: F1, F2 as they are; south: F1 -> sigma F1, F2 -> -sigma F2 */
          for (int side = 0; side < 2; ++side) {
            const int ring = side == 0 ? north[k] : south[k];
            if (ring < 0) continue;
            const double sg = side == 0 ? 1.0 : -1.0;
            double e1[2], b2[2], b1[2], e2[2];
            for (int q = 0; q < 2; ++q) {
              e1[q] = acc[0][0][q][b] + sg * acc[0][1][q][b];
              b2[q] = sg * acc[1][0][q][b] + acc[1][1][q][b];
              b1[q] = acc[2][0][q][b] + sg * acc[2][1][q][b];
              e2[q] = sg * acc[3][0][q][b] + acc[3][1][q][b];
            }
            /* Q = -(E1 + i B2), U = -(B1 - i E2) */
            double *fq = F + 2 * ((long long)ring * (lmax + 1) + m);
            double *fu = F + 2 * (fstride + (long long)ring * (lmax + 1) + m);
            fq[0] = -(e1[0] - b2[1]);
            fq[1] = -(e1[1] + b2[0]);
            fu[0] = -(b1[0] + e2[1]);
            fu[1] = -(b1[1] - e2[0]);
          }
        }
      }
    }
    free(cf0);
    free(cf1);
  }
}

/*
 * Adjoint of the Legendre stage: alm[comp] (complex, m-major triangular)
 * from G[comp][ring][m] (complex); every alm entry with l >= m is written
 * (zero below max(m, |s|)).
 */
void ol_legendre_adjoint(int lmax, int spin, int npair, const int *north, const int *south,
                         const double *x, const double *seed_m, const int *seed_e,
                         const double *G, int nring, double *alm, int nthreads) {
  const long long nelem = (long long)(lmax + 1) * (lmax + 2) / 2;
  const long long gstride = (long long)nring * (lmax + 1);
  const long long sstride = (long long)(lmax + 1) * npair;
#pragma omp parallel num_threads(nthreads)
  {
    coef_t *cf0 = malloc(sizeof(coef_t) * (lmax + 1));
    coef_t *cf1 = malloc(sizeof(coef_t) * (lmax + 1));
#pragma omp for schedule(dynamic, 1)
    for (int m = 0; m <= lmax; ++m) {
      const long long i0 = mstart_of(m, lmax);
      const int ncomp = spin == 0 ? 1 : 2;
      for (int c = 0; c < ncomp; ++c)
        memset(alm + 2 * (c * nelem + i0 + m), 0, sizeof(double) * 2 * (lmax + 1 - m));
      const int ls = spin == 0 ? m : (m > 2 ? m : 2);
      make_coefs(m, spin == 0 ? 0 : -2, ls, lmax, cf0);
      if (spin != 0) make_coefs(m, 2, ls, lmax, cf1);
      double *oe = alm + 2 * i0, *ob = alm + 2 * (nelem + i0);
      for (int k0 = 0; k0 < npair; k0 += NB) {
        const int nb = npair - k0 < NB ? npair - k0 : NB;
        const double *xb = x + k0;
        const long long si = (long long)m * npair + k0;
        start_t s0[NB], s1[NB];
        int lmin = block_starts(cf0, ls, lmax, xb, seed_m + si, seed_e + si, nb, s0);
        if (spin != 0) {
          block_starts(cf1, ls, lmax, xb, seed_m + sstride + si, seed_e + sstride + si, nb, s1);
          lmin = lmax + 1;
          for (int b = 0; b < nb; ++b) {
            align_pair(cf0, cf1, lmax, xb[b], &s0[b], &s1[b]);
            if (s0[b].l0 < lmin) lmin = s0[b].l0;
          }
        }
        /* ring sums per pair: [sum N + S, difference N - S][re, im][pair]
           for the first and (spin 2) second component; unpaired rings have
           S = 0 */
        double g0[2][2][NB] = {{{0}}}, g1[2][2][NB] = {{{0}}};
        for (int b = 0; b < nb; ++b) {
          const int k = k0 + b;
          for (int c = 0; c < ncomp; ++c) {
            double(*g)[2][NB] = c == 0 ? g0 : g1;
            const double *gn = G + 2 * (c * gstride + (long long)north[k] * (lmax + 1) + m);
            double gs[2] = {0.0, 0.0};
            if (south[k] >= 0) {
              const double *p = G + 2 * (c * gstride + (long long)south[k] * (lmax + 1) + m);
              gs[0] = p[0];
              gs[1] = p[1];
            }
            for (int q = 0; q < 2; ++q) {
              g[0][q][b] = gn[q] + gs[q];
              g[1][q][b] = gn[q] - gs[q];
            }
          }
        }
        double p0[NB] = {0}, c0[NB] = {0}, p1[NB] = {0}, c1[NB] = {0};
        for (int l = lmin; l <= lmax; ++l) {
          for (int b = 0; b < nb; ++b)
            if (l == s0[b].l0) {
              p0[b] = s0[b].p;
              c0[b] = s0[b].c;
              if (spin != 0) {
                p1[b] = s1[b].p;
                c1[b] = s1[b].c;
              }
            }
          const int odd = (l - m) & 1;
          if (spin == 0) {
            /* a_l = sum_pairs lam (G_N + sigma G_S) */
            double re = 0.0, im = 0.0;
            for (int b = 0; b < NB; ++b) {
              re += c0[b] * g0[odd][0][b];
              im += c0[b] * g0[odd][1][b];
            }
            oe[2 * l] += re;
            oe[2 * l + 1] += im;
          } else {
            /* F1 pairs with (N + sigma S), F2 with (N - sigma S);
               aE = -(F1 q + i F2 u), aB = -(F1 u - i F2 q) */
            double er = 0.0, ei = 0.0, br = 0.0, bi = 0.0;
            for (int b = 0; b < NB; ++b) {
              const double f1 = 0.5 * (c0[b] + c1[b]), f2 = 0.5 * (c0[b] - c1[b]);
              er += f1 * g0[odd][0][b] - f2 * g1[1 - odd][1][b];
              ei += f1 * g0[odd][1][b] + f2 * g1[1 - odd][0][b];
              br += f1 * g1[odd][0][b] + f2 * g0[1 - odd][1][b];
              bi += f1 * g1[odd][1][b] - f2 * g0[1 - odd][0][b];
            }
            oe[2 * l] -= er;
            oe[2 * l + 1] -= ei;
            ob[2 * l] -= br;
            ob[2 * l + 1] -= bi;
          }
          if (l < lmax) {
            const double A0 = cf0[l].a, B0 = cf0[l].b, G0 = cf0[l].g;
            for (int b = 0; b < NB; ++b) {
              const double n = (A0 * xb[b < nb ? b : 0] + B0) * c0[b] - G0 * p0[b];
              p0[b] = c0[b];
              c0[b] = n;
            }
            if (spin != 0) {
              const double A1 = cf1[l].a, B1 = cf1[l].b, G1 = cf1[l].g;
              for (int b = 0; b < NB; ++b) {
                const double n = (A1 * xb[b < nb ? b : 0] + B1) * c1[b] - G1 * p1[b];
                p1[b] = c1[b];
                c1[b] = n;
              }
            }
          }
        }
      }
    }
    free(cf0);
    free(cf1);
  }
}
