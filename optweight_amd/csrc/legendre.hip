// Legendre transforms between alm (triangular, m-major) and ring
// coefficients G[row][m] for spin 0 and spin 2 (HEALPix convention).
//
// Functions: slam_lm(x) = c_l mu_l with the normalised three-term recurrence
//     mu_l = (A'_l x + B'_l) mu_{l-1} - mu_{l-2}
// (2 VALU ops per step for spin 0), started for each (m, ring pair,
// function) at its first significant multipole l_f with plan-time seeds.
// The normalisation c_l is folded into the staged alm (synthesis) or the
// final sums (adjoint). Below l_f the functions are < 1e-30 and skipped, so
// no exponent tracking is needed at run time.
//
// One thread owns one (synthesis) or K (adjoint) mirror ring pairs
// (theta, pi - theta). The south ring follows from
// slam_lm(-x) = (-1)^(l+m) slam_lm(x) (spin 0) and
// 2lam_lm(pi - theta) = (-1)^(l+m) (-2)lam_lm(theta) (spin 2): every
// recurrence step serves both rings through even/odd (l - m) sums.
// Per-m recurrence coefficients (and, for synthesis, the alm) are staged in
// LDS in chunks and read as broadcasts.
//
// Spin-2 synthesis per m (P = -(aE + i aB)/2, Mm = -(aE - i aB)/2):
//     Q_m = sum_l (P 2lam + Mm (-2)lam),  U_m = -i sum_l (P 2lam - Mm (-2)lam)
// Spin-2 adjoint per m (Zp = q + i u, Zm = q - i u from ring coefficients):
//     S+ = sum_r 2lam Zp,  S- = sum_r (-2)lam Zm,
//     aE = -(S+ + S-)/2,  aB = (i/2)(S+ - S-).
// The adjoint sums over ring pairs in registers (K pairs per thread), then
// across the block through an LDS transposition after every chunk of CH
// multipoles.
#include "sht_internal.hpp"

#include <climits>

namespace owmi {

namespace {

struct LegArgs {
  const int *pair_n;
  const int *pair_s;
  long long npairs;
  long long nrings;
  int lmax;
  int mmax;
};

template <typename T>
__device__ __forceinline__ cplx_t<T> cfma(T a, cplx_t<T> b, cplx_t<T> c) {
  return {fma(a, b.x, c.x), fma(a, b.y, c.y)};
}

// ------------------------------------------------------------ seed finder

__device__ void find_first(double x, double v, int k, int lstart, int lmax, const double *U,
                           int stride, int sign, long long ms, int *lf_out, double *s_out) {
  const int S = ScaleStep<double>::bits;
  const double big = ldexp(1.0, S), small = ldexp(1.0, -S);
  double l1 = v, l2 = 0;
  if (k <= -1000) {
    *lf_out = lmax + 1;
    s_out[0] = s_out[1] = 0;
    return;
  }
  for (int l = lstart; l <= lmax; ++l) {
    if (l > lstart) {
      const double *u = U + stride * (ms + l);
      const double A = u[0];
      const double B = stride == 3 ? sign * u[1] : 0.0;
      const double C = u[stride - 1];
      double lam = (A * x + B) * l1 - C * l2;
      l2 = l1;
      l1 = lam;
      if (k < 0 && fabs(l1) > big) {
        l1 *= small;
        l2 *= small;
        ++k;
      }
    }
    double tv = k >= -7 ? ldexp(l1, S * k) : 0.0;
    if (fabs(tv) >= kNegligible) {
      *lf_out = l;
      s_out[0] = l > lstart ? ldexp(l2, S * k) : 0.0;
      s_out[1] = tv;
      return;
    }
  }
  *lf_out = lmax + 1;
  s_out[0] = s_out[1] = 0;
}

__global__ void __launch_bounds__(256)
seed_finder_kernel(LegArgs g, const double *xs, const double *v0, const int *k0,
                   const double *vp, const int *kp, const double *vm, const int *km,
                   const double *U0, const double *U2, int *lf0, double *s0, int *lfp,
                   double *sp, int *lfm, double *sm) {
  const long long ip = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const int m = blockIdx.y;
  if (ip >= g.npairs) return;
  const long long si = (long long)m * g.npairs + ip;
  const long long ms = mstart_of(m, g.lmax);
  const double x = xs[ip];
  find_first(x, v0[si], k0[si], m, g.lmax, U0, 2, 1, ms, lf0 + si, s0 + 2 * si);
  const int ls2 = m > 2 ? m : 2;
  find_first(x, vp[si], kp[si], ls2, g.lmax, U2, 3, 1, ms, lfp + si, sp + 2 * si);
  find_first(x, vm[si], km[si], ls2, g.lmax, U2, 3, -1, ms, lfm + si, sm + 2 * si);
}

// Block-wide minimum of v (all threads must call).
__device__ __forceinline__ int block_min(int v, int *slot) {
  if (threadIdx.x == 0) *slot = INT_MAX;
  __syncthreads();
  atomicMin(slot, v);
  __syncthreads();
  int r = *slot;
  __syncthreads();
  return r;
}

// ------------------------------------------------------------- synthesis

constexpr int kSynCH = 64;

template <typename T>
__global__ void __launch_bounds__(256)
synth0_kernel(LegArgs g, const T *__restrict__ xs, const int *__restrict__ lfs,
              const T *__restrict__ seeds, const T *__restrict__ coef,
              const cplx_t<T> *__restrict__ alm, long long alm_ts, cplx_t<T> *G) {
  __shared__ T sA[kSynCH];
  __shared__ cplx_t<T> sa[kSynCH];
  __shared__ int smin;
  const long long ip = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const int m = blockIdx.y;
  const long long t = blockIdx.z;
  const bool active = ip < g.npairs;
  const int lmax = g.lmax;
  const long long ms = mstart_of(m, lmax);
  const long long si = (long long)m * g.npairs + ip;
  const int lf = active ? lfs[si] : INT_MAX;
  const int lmin = block_min(lf, &smin);
  cplx_t<T> E = {0, 0}, O = {0, 0};
  if (lmin <= lmax) {
    const T x = active ? xs[ip] : (T)0;
    const T sA0 = active ? seeds[2 * si] : (T)0, sB0 = active ? seeds[2 * si + 1] : (T)0;
    const cplx_t<T> *a = alm + t * alm_ts + ms;
    const T *cf = coef + 2 * ms;
    T l1 = 0, l2 = 0;
    for (int l0 = lmin; l0 <= lmax; l0 += kSynCH) {
      const int nl = min(kSynCH, lmax - l0 + 1);
      __syncthreads();
      for (int j = threadIdx.x; j < nl; j += blockDim.x) {
        const int l = l0 + j;
        sA[j] = cf[2 * l];
        const T c = cf[2 * l + 1];
        const cplx_t<T> av = a[l];
        sa[j] = {av.x * c, av.y * c};
      }
      __syncthreads();
      const bool inj = active && lf >= l0 && lf < l0 + nl;
      if (__syncthreads_or(inj)) {
        for (int i = 0; i < nl; ++i) {
          const int l = l0 + i;
          if (l == lf) {
            l2 = sA0;
            l1 = sB0;
          } else {
            const T lam = fma(sA[i] * x, l1, -l2);
            l2 = l1;
            l1 = lam;
          }
          if ((l - m) & 1)
            O = cfma(l1, sa[i], O);
          else
            E = cfma(l1, sa[i], E);
        }
      } else {
        for (int i = 0; i < nl; ++i) {
          const T lam = fma(sA[i] * x, l1, -l2);
          l2 = l1;
          l1 = lam;
          if ((l0 + i - m) & 1)
            O = cfma(l1, sa[i], O);
          else
            E = cfma(l1, sa[i], E);
        }
      }
    }
  }
  if (!active) return;
  const int rn = g.pair_n[ip], rs = g.pair_s[ip];
  const long long stride = g.mmax + 1;
  G[(t * g.nrings + rn) * stride + m] = {E.x + O.x, E.y + O.y};
  if (rs >= 0) G[(t * g.nrings + rs) * stride + m] = {E.x - O.x, E.y - O.y};
}

template <typename T>
__global__ void __launch_bounds__(256)
synth2_kernel(LegArgs g, const T *__restrict__ xs, const int *__restrict__ lfps,
              const T *__restrict__ sps, const int *__restrict__ lfms, const T *__restrict__ sms,
              const T *__restrict__ coef, const cplx_t<T> *__restrict__ alm, long long alm_ts,
              cplx_t<T> *G) {
  __shared__ T sA[kSynCH];
  __shared__ T sB[kSynCH];
  __shared__ cplx_t<T> sP[kSynCH];
  __shared__ cplx_t<T> sM[kSynCH];
  __shared__ int smin;
  const long long ip = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const int m = blockIdx.y;
  const long long t = blockIdx.z;
  const bool active = ip < g.npairs;
  const int lmax = g.lmax;
  const long long nelem = nelem_of(lmax);
  const long long ms = mstart_of(m, lmax);
  const long long si = (long long)m * g.npairs + ip;
  const int lfp = active ? lfps[si] : INT_MAX;
  const int lfm = active ? lfms[si] : INT_MAX;
  const int lmin = block_min(min(lfp, lfm), &smin);
  // accumulators by parity of (l - m): P*2lam, P*(-2)lam, Mm*2lam, Mm*(-2)lam
  cplx_t<T> PPe = {0, 0}, PMe = {0, 0}, MPe = {0, 0}, MMe = {0, 0};
  cplx_t<T> PPo = {0, 0}, PMo = {0, 0}, MPo = {0, 0}, MMo = {0, 0};
  if (lmin <= lmax) {
    const T x = active ? xs[ip] : (T)0;
    const T pa = active ? sps[2 * si] : (T)0, pb = active ? sps[2 * si + 1] : (T)0;
    const T qa = active ? sms[2 * si] : (T)0, qb = active ? sms[2 * si + 1] : (T)0;
    const cplx_t<T> *aE = alm + t * alm_ts + ms;
    const cplx_t<T> *aB = aE + nelem;
    const T *cf = coef + 3 * ms;
    T p1 = 0, p2 = 0, q1 = 0, q2 = 0;
    for (int l0 = lmin; l0 <= lmax; l0 += kSynCH) {
      const int nl = min(kSynCH, lmax - l0 + 1);
      __syncthreads();
      for (int j = threadIdx.x; j < nl; j += blockDim.x) {
        const int l = l0 + j;
        sA[j] = cf[3 * l];
        sB[j] = cf[3 * l + 1];
        const T c = (T)-0.5 * cf[3 * l + 2];
        const cplx_t<T> e = aE[l], b = aB[l];
        // P = -(e + i b)/2, Mm = -(e - i b)/2, scaled by c_l
        sP[j] = {c * (e.x - b.y), c * (e.y + b.x)};
        sM[j] = {c * (e.x + b.y), c * (e.y - b.x)};
      }
      __syncthreads();
      const bool inj = active && ((lfp >= l0 && lfp < l0 + nl) || (lfm >= l0 && lfm < l0 + nl));
      const bool slow = __syncthreads_or(inj);
      for (int i = 0; i < nl; ++i) {
        const int l = l0 + i;
        const T ax = sA[i] * x;
        const T np_ = fma(ax + sB[i], p1, -p2);
        const T nq_ = fma(ax - sB[i], q1, -q2);
        if (slow) {
          if (l == lfp) {
            p2 = pa;
            p1 = pb;
          } else {
            p2 = p1;
            p1 = np_;
          }
          if (l == lfm) {
            q2 = qa;
            q1 = qb;
          } else {
            q2 = q1;
            q1 = nq_;
          }
        } else {
          p2 = p1;
          p1 = np_;
          q2 = q1;
          q1 = nq_;
        }
        const cplx_t<T> P = sP[i], Mm = sM[i];
        if ((l - m) & 1) {
          PPo = cfma(p1, P, PPo);
          PMo = cfma(q1, P, PMo);
          MPo = cfma(p1, Mm, MPo);
          MMo = cfma(q1, Mm, MMo);
        } else {
          PPe = cfma(p1, P, PPe);
          PMe = cfma(q1, P, PMe);
          MPe = cfma(p1, Mm, MPe);
          MMe = cfma(q1, Mm, MMe);
        }
      }
    }
  }
  if (!active) return;
  const int rn = g.pair_n[ip], rs = g.pair_s[ip];
  const long long stride = g.mmax + 1;
  cplx_t<T> *Gq = G + (t * 2 * g.nrings) * stride;
  cplx_t<T> *Gu = Gq + g.nrings * stride;
  // North: A+ = sum P 2lam, A- = sum Mm (-2)lam; Q = A+ + A-, U = -i (A+ - A-).
  const cplx_t<T> ApN = {PPe.x + PPo.x, PPe.y + PPo.y};
  const cplx_t<T> AmN = {MMe.x + MMo.x, MMe.y + MMo.y};
  Gq[rn * stride + m] = {ApN.x + AmN.x, ApN.y + AmN.y};
  Gu[rn * stride + m] = {ApN.y - AmN.y, -(ApN.x - AmN.x)};
  if (rs >= 0) {
    // South: A+ = sum (-1)^(l+m) P (-2)lam, A- = sum (-1)^(l+m) Mm 2lam.
    const cplx_t<T> ApS = {PMe.x - PMo.x, PMe.y - PMo.y};
    const cplx_t<T> AmS = {MPe.x - MPo.x, MPe.y - MPo.y};
    Gq[rs * stride + m] = {ApS.x + AmS.x, ApS.y + AmS.y};
    Gu[rs * stride + m] = {ApS.y - AmS.y, -(ApS.x - AmS.x)};
  }
}

// --------------------------------------------------------------- adjoint

constexpr int kAdjThreads = 256;

// Sum of red[row][*] over the block: thread tid serves row tid / R (R =
// blockDim / CH threads per row) and returns the row total in lanes with
// tid % R == 0.
template <typename T, int CH>
__device__ __forceinline__ cplx_t<T> row_sum(cplx_t<T> (*red)[kAdjThreads], int nthr) {
  const int R = nthr / CH;
  const int row = threadIdx.x / R, seg = threadIdx.x % R;
  cplx_t<T> acc = {0, 0};
#pragma unroll
  for (int j = 0; j < CH; ++j) {
    const cplx_t<T> v = red[row][seg + j * R];
    acc.x += v.x;
    acc.y += v.y;
  }
  for (int off = R >> 1; off >= 1; off >>= 1) {
    acc.x += __shfl_xor(acc.x, off);
    acc.y += __shfl_xor(acc.y, off);
  }
  return acc;
}

template <typename T, int K, int CH>
__global__ void __launch_bounds__(kAdjThreads)
adj0_kernel(LegArgs g, const T *__restrict__ xs, const int *__restrict__ lfs,
            const T *__restrict__ seeds, const T *__restrict__ coef,
            const cplx_t<T> *__restrict__ G, cplx_t<T> *alm, long long alm_ts) {
  __shared__ T sA[CH];
  __shared__ T sC[CH];
  __shared__ cplx_t<T> red[CH][kAdjThreads];
  __shared__ int smin;
  const int m = blockIdx.x;
  const long long t = blockIdx.y;
  const int nthr = blockDim.x;
  const int lmax = g.lmax;
  const long long ms = mstart_of(m, lmax);
  const T *cf = coef + 2 * ms;
  cplx_t<T> *out = alm + t * alm_ts + ms;
  const long long stride = g.mmax + 1;
  const int R = nthr / CH;
  const int row = threadIdx.x / R;
  const bool writer = (threadIdx.x % R) == 0;

  for (long long base = 0; base < g.npairs; base += (long long)nthr * K) {
    T x[K], l1[K], l2[K], sa0[K], sb0[K];
    int lf[K];
    cplx_t<T> He[K], Ho[K];
    int mylf = INT_MAX;
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const long long ip = base + threadIdx.x + (long long)k * nthr;
      x[k] = 0;
      l1[k] = l2[k] = sa0[k] = sb0[k] = 0;
      lf[k] = INT_MAX;
      He[k] = Ho[k] = {0, 0};
      if (ip < g.npairs) {
        const long long si = (long long)m * g.npairs + ip;
        lf[k] = lfs[si];
        if (lf[k] <= lmax) {
          x[k] = xs[ip];
          sa0[k] = seeds[2 * si];
          sb0[k] = seeds[2 * si + 1];
          const int rn = g.pair_n[ip], rs = g.pair_s[ip];
          const cplx_t<T> gn = G[(t * g.nrings + rn) * stride + m];
          if (rs >= 0) {
            const cplx_t<T> gs = G[(t * g.nrings + rs) * stride + m];
            He[k] = {gn.x + gs.x, gn.y + gs.y};
            Ho[k] = {gn.x - gs.x, gn.y - gs.y};
          } else {
            He[k] = gn;
            Ho[k] = gn;
          }
        }
        mylf = min(mylf, lf[k]);
      }
    }
    const int lmin = block_min(mylf, &smin);
    if (base == 0)
      for (int l = m + threadIdx.x; l < min(lmin, lmax + 1); l += nthr) out[l] = {0, 0};

    for (int l0 = lmin; l0 <= lmax; l0 += CH) {
      const int nl = min(CH, lmax - l0 + 1);
      __syncthreads();
      if (threadIdx.x < nl) {
        sA[threadIdx.x] = cf[2 * (l0 + threadIdx.x)];
        sC[threadIdx.x] = cf[2 * (l0 + threadIdx.x) + 1];
      }
      __syncthreads();
      bool inj = false;
#pragma unroll
      for (int k = 0; k < K; ++k) inj |= (lf[k] >= l0 && lf[k] < l0 + nl);
      const bool slow = __syncthreads_or(inj);
      cplx_t<T> s[CH];
#pragma unroll
      for (int i = 0; i < CH; ++i) {
        s[i] = {0, 0};
        if (i < nl) {
          const int l = l0 + i;
          const bool odd = (l - m) & 1;
#pragma unroll
          for (int k = 0; k < K; ++k) {
            const T lam = fma(sA[i] * x[k], l1[k], -l2[k]);
            if (slow && l == lf[k]) {
              l2[k] = sa0[k];
              l1[k] = sb0[k];
            } else {
              l2[k] = l1[k];
              l1[k] = lam;
            }
            s[i] = cfma(l1[k], odd ? Ho[k] : He[k], s[i]);
          }
        }
      }
#pragma unroll
      for (int i = 0; i < CH; ++i) red[i][threadIdx.x] = s[i];
      __syncthreads();
      const cplx_t<T> tot = row_sum<T, CH>(red, nthr);
      if (writer && row < nl) {
        const int l = l0 + row;
        const T c = sC[row];
        if (base == 0) {
          out[l] = {c * tot.x, c * tot.y};
        } else {
          const cplx_t<T> o = out[l];
          out[l] = {o.x + c * tot.x, o.y + c * tot.y};
        }
      }
    }
  }
}

template <typename T, int K, int CH>
__global__ void __launch_bounds__(kAdjThreads)
adj2_kernel(LegArgs g, const T *__restrict__ xs, const int *__restrict__ lfps,
            const T *__restrict__ sps, const int *__restrict__ lfms, const T *__restrict__ sms,
            const T *__restrict__ coef, const cplx_t<T> *__restrict__ G, cplx_t<T> *alm,
            long long alm_ts) {
  __shared__ T sA[CH];
  __shared__ T sB[CH];
  __shared__ T sC[CH];
  __shared__ cplx_t<T> redP[CH][kAdjThreads];
  __shared__ cplx_t<T> redM[CH][kAdjThreads];
  __shared__ int smin;
  const int m = blockIdx.x;
  const long long t = blockIdx.y;
  const int nthr = blockDim.x;
  const int lmax = g.lmax;
  const long long nelem = nelem_of(lmax);
  const long long ms = mstart_of(m, lmax);
  const T *cf = coef + 3 * ms;
  cplx_t<T> *outE = alm + t * alm_ts + ms;
  cplx_t<T> *outB = outE + nelem;
  const long long stride = g.mmax + 1;
  const cplx_t<T> *Gq = G + (t * 2 * g.nrings) * stride;
  const cplx_t<T> *Gu = Gq + g.nrings * stride;
  const int R = nthr / CH;
  const int row = threadIdx.x / R;
  const bool writer = (threadIdx.x % R) == 0;

  for (long long base = 0; base < g.npairs; base += (long long)nthr * K) {
    T x[K], p1[K], p2[K], q1[K], q2[K], pa[K], pb[K], qa[K], qb[K];
    int lfp[K], lfm[K];
    cplx_t<T> ZpN[K], ZmN[K], ZpS[K], ZmS[K];
    int mylf = INT_MAX;
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const long long ip = base + threadIdx.x + (long long)k * nthr;
      x[k] = p1[k] = p2[k] = q1[k] = q2[k] = pa[k] = pb[k] = qa[k] = qb[k] = 0;
      lfp[k] = lfm[k] = INT_MAX;
      ZpN[k] = ZmN[k] = ZpS[k] = ZmS[k] = {0, 0};
      if (ip < g.npairs) {
        const long long si = (long long)m * g.npairs + ip;
        lfp[k] = lfps[si];
        lfm[k] = lfms[si];
        if (min(lfp[k], lfm[k]) <= lmax) {
          x[k] = xs[ip];
          pa[k] = sps[2 * si];
          pb[k] = sps[2 * si + 1];
          qa[k] = sms[2 * si];
          qb[k] = sms[2 * si + 1];
          const int rn = g.pair_n[ip], rs = g.pair_s[ip];
          const cplx_t<T> qn = Gq[rn * stride + m], un = Gu[rn * stride + m];
          ZpN[k] = {qn.x - un.y, qn.y + un.x};
          ZmN[k] = {qn.x + un.y, qn.y - un.x};
          if (rs >= 0) {
            const cplx_t<T> qs = Gq[rs * stride + m], us = Gu[rs * stride + m];
            ZpS[k] = {qs.x - us.y, qs.y + us.x};
            ZmS[k] = {qs.x + us.y, qs.y - us.x};
          }
        }
        mylf = min(mylf, min(lfp[k], lfm[k]));
      }
    }
    const int lmin = block_min(mylf, &smin);
    if (base == 0)
      for (int l = m + threadIdx.x; l < min(lmin, lmax + 1); l += nthr) {
        outE[l] = {0, 0};
        outB[l] = {0, 0};
      }

    for (int l0 = lmin; l0 <= lmax; l0 += CH) {
      const int nl = min(CH, lmax - l0 + 1);
      __syncthreads();
      if (threadIdx.x < nl) {
        const int l = l0 + threadIdx.x;
        sA[threadIdx.x] = cf[3 * l];
        sB[threadIdx.x] = cf[3 * l + 1];
        sC[threadIdx.x] = cf[3 * l + 2];
      }
      __syncthreads();
      bool inj = false;
#pragma unroll
      for (int k = 0; k < K; ++k)
        inj |= (lfp[k] >= l0 && lfp[k] < l0 + nl) || (lfm[k] >= l0 && lfm[k] < l0 + nl);
      const bool slow = __syncthreads_or(inj);
      cplx_t<T> sPl[CH], sMl[CH];
#pragma unroll
      for (int i = 0; i < CH; ++i) {
        sPl[i] = sMl[i] = {0, 0};
        if (i < nl) {
          const int l = l0 + i;
          const T sg = ((l - m) & 1) ? (T)-1 : (T)1;
#pragma unroll
          for (int k = 0; k < K; ++k) {
            const T ax = sA[i] * x[k];
            const T np_ = fma(ax + sB[i], p1[k], -p2[k]);
            const T nq_ = fma(ax - sB[i], q1[k], -q2[k]);
            if (slow && l == lfp[k]) {
              p2[k] = pa[k];
              p1[k] = pb[k];
            } else {
              p2[k] = p1[k];
              p1[k] = np_;
            }
            if (slow && l == lfm[k]) {
              q2[k] = qa[k];
              q1[k] = qb[k];
            } else {
              q2[k] = q1[k];
              q1[k] = nq_;
            }
            // even: S+ = 2lam ZpN + (-2)lam ZpS, S- = (-2)lam ZmN + 2lam ZmS
            // odd : S+ = 2lam ZpN - (-2)lam ZpS, S- = (-2)lam ZmN - 2lam ZmS
            sPl[i] = cfma(p1[k], ZpN[k], sPl[i]);
            sPl[i] = cfma(sg * q1[k], ZpS[k], sPl[i]);
            sMl[i] = cfma(q1[k], ZmN[k], sMl[i]);
            sMl[i] = cfma(sg * p1[k], ZmS[k], sMl[i]);
          }
        }
      }
#pragma unroll
      for (int i = 0; i < CH; ++i) {
        redP[i][threadIdx.x] = sPl[i];
        redM[i][threadIdx.x] = sMl[i];
      }
      __syncthreads();
      const cplx_t<T> sp = row_sum<T, CH>(redP, nthr);
      const cplx_t<T> sm = row_sum<T, CH>(redM, nthr);
      if (writer && row < nl) {
        const int l = l0 + row;
        const T c = (T)0.5 * sC[row];
        // aE = -(S+ + S-)/2, aB = (i/2)(S+ - S-)
        const cplx_t<T> e = {-c * (sp.x + sm.x), -c * (sp.y + sm.y)};
        const cplx_t<T> b = {-c * (sp.y - sm.y), c * (sp.x - sm.x)};
        if (base == 0) {
          outE[l] = e;
          outB[l] = b;
        } else {
          const cplx_t<T> oe = outE[l], ob = outB[l];
          outE[l] = {oe.x + e.x, oe.y + e.y};
          outB[l] = {ob.x + b.x, ob.y + b.y};
        }
      }
    }
  }
}

template <typename T>
PlanTables<T> &tab(owmi_plan_impl *p);
template <>
PlanTables<double> &tab<double>(owmi_plan_impl *p) { return p->t64; }
template <>
PlanTables<float> &tab<float>(owmi_plan_impl *p) { return p->t32; }

LegArgs leg_args(owmi_plan_impl *p) {
  LegArgs a;
  a.pair_n = p->pair_n;
  a.pair_s = p->pair_s;
  a.npairs = p->npairs;
  a.nrings = p->nrings;
  a.lmax = p->lmax;
  a.mmax = p->mmax;
  return a;
}

// Chunk lengths of the adjoint: LDS holds CH x 256 complex values per sum
// (32 KB).
template <typename T> struct AdjCH;
template <> struct AdjCH<double> { static constexpr int s0 = 8, s2 = 4; };
template <> struct AdjCH<float> { static constexpr int s0 = 16, s2 = 8; };

}  // namespace

void launch_seed_finder(owmi_plan_impl *p, const double *xs, const double *v0, const int *k0,
                        const double *vp, const int *kp, const double *vm, const int *km,
                        const double *U0, const double *U2, int *lf0, double *s0, int *lfp,
                        double *sp, int *lfm, double *sm) {
  dim3 grid(ceil_div(p->npairs, 256), (unsigned)(p->mmax + 1));
  hipLaunchKernelGGL(seed_finder_kernel, grid, dim3(256), 0, 0, leg_args(p), xs, v0, k0, vp, kp,
                     vm, km, U0, U2, lf0, s0, lfp, sp, lfm, sm);
  OWMI_HIP(hipGetLastError());
}

template <typename T>
void launch_synthesis(owmi_plan_impl *p, int spin, const cplx_t<T> *alm, long long alm_ts,
                      cplx_t<T> *G, long long ntrans, hipStream_t st) {
  OWMI_REQUIRE(ntrans < 65536, "too many transforms for one launch");
  PlanTables<T> &tb = tab<T>(p);
  int bs = (int)std::min<long long>(256, ((p->npairs + 63) / 64) * 64);
  dim3 grid(ceil_div(p->npairs, bs), (unsigned)(p->mmax + 1), (unsigned)ntrans);
  if (spin == 0)
    hipLaunchKernelGGL(synth0_kernel<T>, grid, dim3(bs), 0, st, leg_args(p), tb.x, tb.lf0, tb.s0,
                       tb.coef0, alm, alm_ts, G);
  else
    hipLaunchKernelGGL(synth2_kernel<T>, grid, dim3(bs), 0, st, leg_args(p), tb.x, tb.lfp, tb.sp,
                       tb.lfm, tb.sm, tb.coef2, alm, alm_ts, G);
}

template <typename T, int K>
static void launch_adj_k(owmi_plan_impl *p, int spin, int bs, const cplx_t<T> *G, cplx_t<T> *alm,
                         long long alm_ts, long long ntrans, hipStream_t st) {
  PlanTables<T> &tb = tab<T>(p);
  dim3 grid((unsigned)(p->mmax + 1), (unsigned)ntrans);
  if (spin == 0)
    hipLaunchKernelGGL((adj0_kernel<T, K, AdjCH<T>::s0>), grid, dim3(bs), 0, st, leg_args(p),
                       tb.x, tb.lf0, tb.s0, tb.coef0, G, alm, alm_ts);
  else
    hipLaunchKernelGGL((adj2_kernel<T, K, AdjCH<T>::s2>), grid, dim3(bs), 0, st, leg_args(p),
                       tb.x, tb.lfp, tb.sp, tb.lfm, tb.sm, tb.coef2, G, alm, alm_ts);
}

template <typename T>
void launch_adjoint(owmi_plan_impl *p, int spin, const cplx_t<T> *G, cplx_t<T> *alm,
                    long long alm_ts, long long ntrans, hipStream_t st) {
  OWMI_REQUIRE(ntrans < 65536, "too many transforms for one launch");
  // threads: a power of two in [64, 256] covering the pairs with K <= 2
  int bs = 64;
  while (bs < kAdjThreads && bs * 2 < p->npairs) bs *= 2;
  if (p->npairs > (long long)bs)
    launch_adj_k<T, 2>(p, spin, bs, G, alm, alm_ts, ntrans, st);
  else
    launch_adj_k<T, 1>(p, spin, bs, G, alm, alm_ts, ntrans, st);
}

template void launch_synthesis<double>(owmi_plan_impl *, int, const double2 *, long long,
                                       double2 *, long long, hipStream_t);
template void launch_synthesis<float>(owmi_plan_impl *, int, const float2 *, long long, float2 *,
                                      long long, hipStream_t);
template void launch_adjoint<double>(owmi_plan_impl *, int, const double2 *, double2 *, long long,
                                     long long, hipStream_t);
template void launch_adjoint<float>(owmi_plan_impl *, int, const float2 *, float2 *, long long,
                                    long long, hipStream_t);

}  // namespace owmi
