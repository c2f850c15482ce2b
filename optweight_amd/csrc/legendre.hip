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

// The recurrence (x, coefficients, seeds, mu) always runs in float64: near
// the poles 1 - x ~ 1e-6 at lmax 2047 is not resolved by float32 and the
// three-term recurrence is ill-conditioned there (float32 transforms were
// off by up to 3e-2 on the polar ring). The float32 kernels keep the alm,
// ring coefficients and accumulations in float32; float64 VALU runs at the
// float32 (non-packed) rate on this hardware.
using R = double;

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
//
// NB transforms per block share one recurrence per ring pair: the staged
// (scaled) alm of the NB transforms sit side by side in LDS, and each
// recurrence value feeds NB accumulations.

constexpr int kSynCH = 64;

template <typename T, int NB>
__global__ void __launch_bounds__(256)
synth0_kernel(LegArgs g, const R *__restrict__ xs, const int *__restrict__ lfs,
              const R *__restrict__ seeds, const R *__restrict__ coef,
              const cplx_t<T> *__restrict__ alm, long long alm_ts, long long ntrans,
              cplx_t<T> *G) {
  __shared__ R sA[kSynCH];
  __shared__ cplx_t<T> sa[kSynCH][NB];
  __shared__ int smin;
  const long long ip = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const int m = blockIdx.y;
  const long long t0 = (long long)blockIdx.z * NB;
  const int nt = (int)min((long long)NB, ntrans - t0);
  const bool active = ip < g.npairs;
  const int lmax = g.lmax;
  const long long ms = mstart_of(m, lmax);
  const long long si = (long long)m * g.npairs + ip;
  const int lf = active ? lfs[si] : INT_MAX;
  const int lmin = block_min(lf, &smin);
  cplx_t<T> E[NB], O[NB];
#pragma unroll
  for (int b = 0; b < NB; ++b) E[b] = O[b] = {0, 0};
  if (lmin <= lmax) {
    const R x = active ? xs[ip] : (R)0;
    const R sA0 = active ? seeds[2 * si] : (R)0, sB0 = active ? seeds[2 * si + 1] : (R)0;
    const cplx_t<T> *a = alm + t0 * alm_ts + ms;
    const R *cf = coef + 2 * ms;
    R l1 = 0, l2 = 0;
    for (int l0 = lmin; l0 <= lmax; l0 += kSynCH) {
      const int nl = min(kSynCH, lmax - l0 + 1);
      __syncthreads();
      for (int j = threadIdx.x; j < kSynCH * NB; j += blockDim.x) {
        const int jl = j % kSynCH, b = j / kSynCH;
        if (jl < nl) {
          const int l = l0 + jl;
          if (b == 0) sA[jl] = cf[2 * l];
          const T c = (T)cf[2 * l + 1];
          const cplx_t<T> av = b < nt ? a[b * alm_ts + l] : cplx_t<T>{0, 0};
          sa[jl][b] = {av.x * c, av.y * c};
        }
      }
      __syncthreads();
      const bool inj = active && lf >= l0 && lf < l0 + nl;
      const bool slow = __syncthreads_or(inj);
      for (int i = 0; i < nl; ++i) {
        const int l = l0 + i;
        const R lam = fma(sA[i] * x, l1, -l2);
        if (slow && l == lf) {
          l2 = sA0;
          l1 = sB0;
        } else {
          l2 = l1;
          l1 = lam;
        }
        const T lt = (T)l1;
        if ((l - m) & 1) {
#pragma unroll
          for (int b = 0; b < NB; ++b) O[b] = cfma(lt, sa[i][b], O[b]);
        } else {
#pragma unroll
          for (int b = 0; b < NB; ++b) E[b] = cfma(lt, sa[i][b], E[b]);
        }
      }
    }
  }
  if (!active) return;
  const int rn = g.pair_n[ip], rs = g.pair_s[ip];
  const long long stride = g.mmax + 1;
#pragma unroll
  for (int b = 0; b < NB; ++b) {
    if (b < nt) {
      const long long t = t0 + b;
      G[(t * g.nrings + rn) * stride + m] = {E[b].x + O[b].x, E[b].y + O[b].y};
      if (rs >= 0) G[(t * g.nrings + rs) * stride + m] = {E[b].x - O[b].x, E[b].y - O[b].y};
    }
  }
}

template <typename T, int NB>
__global__ void __launch_bounds__(256)
synth2_kernel(LegArgs g, const R *__restrict__ xs, const int *__restrict__ lfps,
              const R *__restrict__ sps, const int *__restrict__ lfms, const R *__restrict__ sms,
              const R *__restrict__ coef, const cplx_t<T> *__restrict__ alm, long long alm_ts,
              long long ntrans, cplx_t<T> *G) {
  __shared__ R sA[kSynCH];
  __shared__ R sB[kSynCH];
  __shared__ cplx_t<T> sP[kSynCH][NB];
  __shared__ cplx_t<T> sM[kSynCH][NB];
  __shared__ int smin;
  const long long ip = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const int m = blockIdx.y;
  const long long t0 = (long long)blockIdx.z * NB;
  const int nt = (int)min((long long)NB, ntrans - t0);
  const bool active = ip < g.npairs;
  const int lmax = g.lmax;
  const long long nelem = nelem_of(lmax);
  const long long ms = mstart_of(m, lmax);
  const long long si = (long long)m * g.npairs + ip;
  const int lfp = active ? lfps[si] : INT_MAX;
  const int lfm = active ? lfms[si] : INT_MAX;
  const int lmin = block_min(min(lfp, lfm), &smin);
  // accumulators by parity of (l - m): P*2lam, P*(-2)lam, Mm*2lam, Mm*(-2)lam
  cplx_t<T> PPe[NB], PMe[NB], MPe[NB], MMe[NB], PPo[NB], PMo[NB], MPo[NB], MMo[NB];
#pragma unroll
  for (int b = 0; b < NB; ++b)
    PPe[b] = PMe[b] = MPe[b] = MMe[b] = PPo[b] = PMo[b] = MPo[b] = MMo[b] = {0, 0};
  if (lmin <= lmax) {
    const R x = active ? xs[ip] : (R)0;
    const R pa = active ? sps[2 * si] : (R)0, pb = active ? sps[2 * si + 1] : (R)0;
    const R qa = active ? sms[2 * si] : (R)0, qb = active ? sms[2 * si + 1] : (R)0;
    const cplx_t<T> *aE = alm + t0 * alm_ts + ms;
    const R *cf = coef + 3 * ms;
    R p1 = 0, p2 = 0, q1 = 0, q2 = 0;
    for (int l0 = lmin; l0 <= lmax; l0 += kSynCH) {
      const int nl = min(kSynCH, lmax - l0 + 1);
      __syncthreads();
      for (int j = threadIdx.x; j < kSynCH * NB; j += blockDim.x) {
        const int jl = j % kSynCH, b = j / kSynCH;
        if (jl < nl) {
          const int l = l0 + jl;
          if (b == 0) {
            sA[jl] = cf[3 * l];
            sB[jl] = cf[3 * l + 1];
          }
          const T c = (T)(-0.5 * cf[3 * l + 2]);
          cplx_t<T> e = {0, 0}, bb = {0, 0};
          if (b < nt) {
            e = aE[b * alm_ts + l];
            bb = aE[b * alm_ts + nelem + l];
          }
          // P = -(e + i b)/2, Mm = -(e - i b)/2, scaled by c_l
          sP[jl][b] = {c * (e.x - bb.y), c * (e.y + bb.x)};
          sM[jl][b] = {c * (e.x + bb.y), c * (e.y - bb.x)};
        }
      }
      __syncthreads();
      const bool inj = active && ((lfp >= l0 && lfp < l0 + nl) || (lfm >= l0 && lfm < l0 + nl));
      const bool slow = __syncthreads_or(inj);
      for (int i = 0; i < nl; ++i) {
        const int l = l0 + i;
        const R ax = sA[i] * x;
        const R np_ = fma(ax + sB[i], p1, -p2);
        const R nq_ = fma(ax - sB[i], q1, -q2);
        if (slow && l == lfp) {
          p2 = pa;
          p1 = pb;
        } else {
          p2 = p1;
          p1 = np_;
        }
        if (slow && l == lfm) {
          q2 = qa;
          q1 = qb;
        } else {
          q2 = q1;
          q1 = nq_;
        }
        const T pt = (T)p1, qt = (T)q1;
        if ((l - m) & 1) {
#pragma unroll
          for (int b = 0; b < NB; ++b) {
            const cplx_t<T> P = sP[i][b], Mm = sM[i][b];
            PPo[b] = cfma(pt, P, PPo[b]);
            PMo[b] = cfma(qt, P, PMo[b]);
            MPo[b] = cfma(pt, Mm, MPo[b]);
            MMo[b] = cfma(qt, Mm, MMo[b]);
          }
        } else {
#pragma unroll
          for (int b = 0; b < NB; ++b) {
            const cplx_t<T> P = sP[i][b], Mm = sM[i][b];
            PPe[b] = cfma(pt, P, PPe[b]);
            PMe[b] = cfma(qt, P, PMe[b]);
            MPe[b] = cfma(pt, Mm, MPe[b]);
            MMe[b] = cfma(qt, Mm, MMe[b]);
          }
        }
      }
    }
  }
  if (!active) return;
  const int rn = g.pair_n[ip], rs = g.pair_s[ip];
  const long long stride = g.mmax + 1;
#pragma unroll
  for (int b = 0; b < NB; ++b) {
    if (b >= nt) continue;
    const long long t = t0 + b;
    cplx_t<T> *Gq = G + (t * 2 * g.nrings) * stride;
    cplx_t<T> *Gu = Gq + g.nrings * stride;
    // North: A+ = sum P 2lam, A- = sum Mm (-2)lam; Q = A+ + A-, U = -i (A+ - A-).
    const cplx_t<T> ApN = {PPe[b].x + PPo[b].x, PPe[b].y + PPo[b].y};
    const cplx_t<T> AmN = {MMe[b].x + MMo[b].x, MMe[b].y + MMo[b].y};
    Gq[rn * stride + m] = {ApN.x + AmN.x, ApN.y + AmN.y};
    Gu[rn * stride + m] = {ApN.y - AmN.y, -(ApN.x - AmN.x)};
    if (rs >= 0) {
      // South: A+ = sum (-1)^(l+m) P (-2)lam, A- = sum (-1)^(l+m) Mm 2lam.
      const cplx_t<T> ApS = {PMe[b].x - PMo[b].x, PMe[b].y - PMo[b].y};
      const cplx_t<T> AmS = {MPe[b].x - MPo[b].x, MPe[b].y - MPo[b].y};
      Gq[rs * stride + m] = {ApS.x + AmS.x, ApS.y + AmS.y};
      Gu[rs * stride + m] = {ApS.y - AmS.y, -(ApS.x - AmS.x)};
    }
  }
}

// --------------------------------------------------------------- adjoint
//
// One wavefront per work item: an m paired with lmax - m (uniform work per
// item) and a group of 64 K ring pairs (K per lane). The wave sums over its
// pairs in registers and across its lanes through a wave-private LDS region
// after every CH multipoles; no block-wide barrier is used. Ring-pair groups
// write separate partial sums that a second kernel adds in fixed order.

constexpr int kAdjWaves = 4;  // independent waves per block

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

__device__ __forceinline__ int wave_min(int v) {
  for (int off = 32; off >= 1; off >>= 1) v = min(v, __shfl_xor(v, off));
  return v;
}

// Row totals of red[i][lane] (i < CH): lane L serves row L / R (R = 64 / CH
// lanes per row); lanes with L % R == 0 return the row total.
template <typename T, int CH>
__device__ __forceinline__ cplx_t<T> wave_rows(const cplx_t<T> (*red)[64], int lane) {
  constexpr int R = 64 / CH;
  const int row = lane / R, seg = lane % R;
  cplx_t<T> acc = {0, 0};
#pragma unroll
  for (int j = 0; j < CH; ++j) {
    const cplx_t<T> v = red[row][seg + j * R];
    acc.x += v.x;
    acc.y += v.y;
  }
#pragma unroll
  for (int off = R >> 1; off >= 1; off >>= 1) {
    acc.x += __shfl_xor(acc.x, off);
    acc.y += __shfl_xor(acc.y, off);
  }
  return acc;
}

struct AdjItems {
  int nmpairs;        // ceil((mmax + 1) / 2)
  int ngroups;        // ring-pair groups of 64 K pairs
  long long grp_ts;   // elements between the partial buffers of two groups
};

template <typename T, int K, int CH, int NB>
__global__ void __launch_bounds__(64 * kAdjWaves)
adj0_kernel(LegArgs g, AdjItems it, const R *__restrict__ xs, const int *__restrict__ lfs,
            const R *__restrict__ seeds, const R *__restrict__ coef,
            const cplx_t<T> *__restrict__ G, long long ntrans, cplx_t<T> *out,
            long long out_ts) {
  constexpr int ROWS = CH * NB;  // reduction rows per chunk: (multipole, transform)
  __shared__ cplx_t<T> red_all[kAdjWaves][ROWS][64];
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
  const long long item = (long long)blockIdx.x * kAdjWaves + wave;
  if (item >= (long long)it.nmpairs * it.ngroups) return;
  cplx_t<T>(*red)[64] = red_all[wave];
  const int q = (int)(item / it.ngroups), grp = (int)(item % it.ngroups);
  const long long t0 = (long long)blockIdx.y * NB;
  const int nt = (int)min((long long)NB, ntrans - t0);
  const int lmax = g.lmax;
  const long long stride = g.mmax + 1;
  constexpr int RL = 64 / ROWS;  // lanes per reduction row
  const int row = lane / RL, ri = row / NB, rb = row % NB;
  const bool writer = (lane % RL) == 0;

  for (int pass = 0; pass < 2; ++pass) {
    const int m = pass == 0 ? q : g.mmax - q;
    if (pass == 1 && m == q) break;
    const long long ms = mstart_of(m, lmax);
    const R *cf = coef + 2 * ms;
    cplx_t<T> *om = out + grp * it.grp_ts + t0 * out_ts + ms;
    R x[K], l1[K], l2[K], sa0[K], sb0[K];
    int lf[K];
    cplx_t<T> He[K][NB], Ho[K][NB];
    int mylf = INT_MAX;
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const long long ip = (long long)grp * 64 * K + lane + 64 * k;
      x[k] = l1[k] = l2[k] = sa0[k] = sb0[k] = 0;
      lf[k] = INT_MAX;
#pragma unroll
      for (int b = 0; b < NB; ++b) He[k][b] = Ho[k][b] = {0, 0};
      if (ip < g.npairs) {
        const long long si = (long long)m * g.npairs + ip;
        lf[k] = lfs[si];
        if (lf[k] <= lmax) {
          x[k] = xs[ip];
          sa0[k] = seeds[2 * si];
          sb0[k] = seeds[2 * si + 1];
          const int rn = g.pair_n[ip], rs = g.pair_s[ip];
#pragma unroll
          for (int b = 0; b < NB; ++b) {
            if (b < nt) {
              const long long t = t0 + b;
              const cplx_t<T> gn = G[(t * g.nrings + rn) * stride + m];
              if (rs >= 0) {
                const cplx_t<T> gs = G[(t * g.nrings + rs) * stride + m];
                He[k][b] = {gn.x + gs.x, gn.y + gs.y};
                Ho[k][b] = {gn.x - gs.x, gn.y - gs.y};
              } else {
                He[k][b] = gn;
                Ho[k][b] = gn;
              }
            }
          }
        }
        mylf = min(mylf, lf[k]);
      }
    }
    const int lmin = __builtin_amdgcn_readfirstlane(wave_min(mylf));
    for (int b = 0; b < nt; ++b)
      for (int l = m + lane; l < min(lmin, lmax + 1); l += 64) om[b * out_ts + l] = {0, 0};

    for (int l0 = lmin; l0 <= lmax; l0 += CH) {
      const int nl = min(CH, lmax - l0 + 1);
      bool inj = false;
#pragma unroll
      for (int k = 0; k < K; ++k) inj |= (lf[k] >= l0 && lf[k] < l0 + nl);
      // wave-uniform coefficients of the chunk (clamped index: scalar loads)
      R Ac[CH];
#pragma unroll
      for (int i = 0; i < CH; ++i) Ac[i] = cf[2 * min(l0 + i, lmax)];
      // weights of the first and second multipole of the chunk (parity
      // alternates with l, so the selection is made once per chunk)
      const bool odd0 = (l0 - m) & 1;
      cplx_t<T> H0[K][NB], H1[K][NB];
#pragma unroll
      for (int k = 0; k < K; ++k)
#pragma unroll
        for (int b = 0; b < NB; ++b) {
          H0[k][b] = odd0 ? Ho[k][b] : He[k][b];
          H1[k][b] = odd0 ? He[k][b] : Ho[k][b];
        }
      cplx_t<T> s[CH][NB];
#pragma unroll
      for (int i = 0; i < CH; ++i)
#pragma unroll
        for (int b = 0; b < NB; ++b) s[i][b] = {0, 0};
      if (nl == CH && !__any(inj)) {
        // steady state: plain recurrence over a full chunk
#pragma unroll
        for (int i = 0; i < CH; ++i) {
#pragma unroll
          for (int k = 0; k < K; ++k) {
            const R lam = fma(Ac[i] * x[k], l1[k], -l2[k]);
            l2[k] = l1[k];
            l1[k] = lam;
            const T lt = (T)lam;
#pragma unroll
            for (int b = 0; b < NB; ++b)
              s[i][b] = cfma(lt, (i & 1) ? H1[k][b] : H0[k][b], s[i][b]);
          }
        }
      } else {
        // chunk holding a start multipole of some lane, or the last chunk
#pragma unroll
        for (int i = 0; i < CH; ++i) {
          if (i < nl) {
            const int l = l0 + i;
#pragma unroll
            for (int k = 0; k < K; ++k) {
              const R lam = fma(Ac[i] * x[k], l1[k], -l2[k]);
              const bool st = (l == lf[k]);
              l2[k] = st ? sa0[k] : l1[k];
              l1[k] = st ? sb0[k] : lam;
              const T lt = (T)l1[k];
#pragma unroll
              for (int b = 0; b < NB; ++b)
                s[i][b] = cfma(lt, (i & 1) ? H1[k][b] : H0[k][b], s[i][b]);
            }
          }
        }
      }
      wave_sync();
#pragma unroll
      for (int i = 0; i < CH; ++i)
#pragma unroll
        for (int b = 0; b < NB; ++b) red[i * NB + b][lane] = s[i][b];
      wave_sync();
      const cplx_t<T> tot = wave_rows<T, ROWS>(red, lane);
      if (writer && ri < nl && rb < nt) {
        const int l = l0 + ri;
        const T c = (T)cf[2 * l + 1];
        om[rb * out_ts + l] = {c * tot.x, c * tot.y};
      }
    }
  }
}

template <typename T, int K, int CH, int NB>
__global__ void __launch_bounds__(64 * kAdjWaves)
adj2_kernel(LegArgs g, AdjItems it, const R *__restrict__ xs, const int *__restrict__ lfps,
            const R *__restrict__ sps, const int *__restrict__ lfms, const R *__restrict__ sms,
            const R *__restrict__ coef, const cplx_t<T> *__restrict__ G, long long ntrans,
            cplx_t<T> *out, long long out_ts) {
  constexpr int ROWS = CH * NB;
  __shared__ cplx_t<T> redP_all[kAdjWaves][ROWS][64];
  __shared__ cplx_t<T> redM_all[kAdjWaves][ROWS][64];
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
  const long long item = (long long)blockIdx.x * kAdjWaves + wave;
  if (item >= (long long)it.nmpairs * it.ngroups) return;
  cplx_t<T>(*redP)[64] = redP_all[wave];
  cplx_t<T>(*redM)[64] = redM_all[wave];
  const int q = (int)(item / it.ngroups), grp = (int)(item % it.ngroups);
  const long long t0 = (long long)blockIdx.y * NB;
  const int nt = (int)min((long long)NB, ntrans - t0);
  const int lmax = g.lmax;
  const long long nelem = nelem_of(lmax);
  const long long stride = g.mmax + 1;
  constexpr int RL = 64 / ROWS;  // lanes per reduction row
  const int row = lane / RL, ri = row / NB, rb = row % NB;
  const bool writer = (lane % RL) == 0;

  for (int pass = 0; pass < 2; ++pass) {
    const int m = pass == 0 ? q : g.mmax - q;
    if (pass == 1 && m == q) break;
    const long long ms = mstart_of(m, lmax);
    const R *cf = coef + 3 * ms;
    cplx_t<T> *oE = out + grp * it.grp_ts + t0 * out_ts + ms;
    R x[K], p1[K], p2[K], q1[K], q2[K], pa[K], pb[K], qa[K], qb[K];
    int lfp[K], lfm[K];
    cplx_t<T> ZpN[K][NB], ZmN[K][NB], ZpS[K][NB], ZmS[K][NB];
    int mylf = INT_MAX;
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const long long ip = (long long)grp * 64 * K + lane + 64 * k;
      x[k] = p1[k] = p2[k] = q1[k] = q2[k] = pa[k] = pb[k] = qa[k] = qb[k] = 0;
      lfp[k] = lfm[k] = INT_MAX;
#pragma unroll
      for (int b = 0; b < NB; ++b) ZpN[k][b] = ZmN[k][b] = ZpS[k][b] = ZmS[k][b] = {0, 0};
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
#pragma unroll
          for (int b = 0; b < NB; ++b) {
            if (b < nt) {
              const cplx_t<T> *Gq = G + ((t0 + b) * 2 * g.nrings) * stride;
              const cplx_t<T> *Gu = Gq + g.nrings * stride;
              const cplx_t<T> qn = Gq[rn * stride + m], un = Gu[rn * stride + m];
              ZpN[k][b] = {qn.x - un.y, qn.y + un.x};
              ZmN[k][b] = {qn.x + un.y, qn.y - un.x};
              if (rs >= 0) {
                const cplx_t<T> qs = Gq[rs * stride + m], us = Gu[rs * stride + m];
                ZpS[k][b] = {qs.x - us.y, qs.y + us.x};
                ZmS[k][b] = {qs.x + us.y, qs.y - us.x};
              }
            }
          }
        }
        mylf = min(mylf, min(lfp[k], lfm[k]));
      }
    }
    const int lmin = __builtin_amdgcn_readfirstlane(wave_min(mylf));
    for (int b = 0; b < nt; ++b)
      for (int l = m + lane; l < min(lmin, lmax + 1); l += 64) {
        oE[b * out_ts + l] = {0, 0};
        oE[b * out_ts + nelem + l] = {0, 0};
      }

    for (int l0 = lmin; l0 <= lmax; l0 += CH) {
      const int nl = min(CH, lmax - l0 + 1);
      bool inj = false;
#pragma unroll
      for (int k = 0; k < K; ++k)
        inj |= (lfp[k] >= l0 && lfp[k] < l0 + nl) || (lfm[k] >= l0 && lfm[k] < l0 + nl);
      // wave-uniform coefficients of the chunk (clamped index: scalar loads)
      R Ac[CH], Bc[CH];
#pragma unroll
      for (int i = 0; i < CH; ++i) {
        const int l = min(l0 + i, lmax);
        Ac[i] = cf[3 * l];
        Bc[i] = cf[3 * l + 1];
      }
      // south weights with the sign of the chunk's first multipole; the
      // sign alternates with l
      const T sg0 = ((l0 - m) & 1) ? (T)-1 : (T)1;
      cplx_t<T> PS[K][NB], MS[K][NB];
#pragma unroll
      for (int k = 0; k < K; ++k)
#pragma unroll
        for (int b = 0; b < NB; ++b) {
          PS[k][b] = {sg0 * ZpS[k][b].x, sg0 * ZpS[k][b].y};
          MS[k][b] = {sg0 * ZmS[k][b].x, sg0 * ZmS[k][b].y};
        }
      cplx_t<T> sP[CH][NB], sM[CH][NB];
#pragma unroll
      for (int i = 0; i < CH; ++i)
#pragma unroll
        for (int b = 0; b < NB; ++b) sP[i][b] = sM[i][b] = {0, 0};
      if (nl == CH && !__any(inj)) {
        // steady state: plain recurrences over a full chunk
        // even l-m: S+ += lam+ ZpN + lam- ZpS, S- += lam- ZmN + lam+ ZmS
        // odd l-m : the south terms change sign
#pragma unroll
        for (int i = 0; i < CH; ++i) {
#pragma unroll
          for (int k = 0; k < K; ++k) {
            const R ax = Ac[i] * x[k];
            const R np_ = fma(ax + Bc[i], p1[k], -p2[k]);
            const R nq_ = fma(ax - Bc[i], q1[k], -q2[k]);
            p2[k] = p1[k];
            p1[k] = np_;
            q2[k] = q1[k];
            q1[k] = nq_;
            const T pt = (T)np_, qt = (T)nq_;
            const T nqs = (i & 1) ? -qt : qt, nps = (i & 1) ? -pt : pt;
#pragma unroll
            for (int b = 0; b < NB; ++b) {
              sP[i][b] = cfma(pt, ZpN[k][b], sP[i][b]);
              sP[i][b] = cfma(nqs, PS[k][b], sP[i][b]);
              sM[i][b] = cfma(qt, ZmN[k][b], sM[i][b]);
              sM[i][b] = cfma(nps, MS[k][b], sM[i][b]);
            }
          }
        }
      } else {
        // chunk holding a start multipole of some lane, or the last chunk
#pragma unroll
        for (int i = 0; i < CH; ++i) {
          if (i < nl) {
            const int l = l0 + i;
#pragma unroll
            for (int k = 0; k < K; ++k) {
              const R ax = Ac[i] * x[k];
              const R np_ = fma(ax + Bc[i], p1[k], -p2[k]);
              const R nq_ = fma(ax - Bc[i], q1[k], -q2[k]);
              const bool sp_ = (l == lfp[k]), sq_ = (l == lfm[k]);
              p2[k] = sp_ ? pa[k] : p1[k];
              p1[k] = sp_ ? pb[k] : np_;
              q2[k] = sq_ ? qa[k] : q1[k];
              q1[k] = sq_ ? qb[k] : nq_;
              const T pt = (T)p1[k], qt = (T)q1[k];
              const T nqs = (i & 1) ? -qt : qt, nps = (i & 1) ? -pt : pt;
#pragma unroll
              for (int b = 0; b < NB; ++b) {
                sP[i][b] = cfma(pt, ZpN[k][b], sP[i][b]);
                sP[i][b] = cfma(nqs, PS[k][b], sP[i][b]);
                sM[i][b] = cfma(qt, ZmN[k][b], sM[i][b]);
                sM[i][b] = cfma(nps, MS[k][b], sM[i][b]);
              }
            }
          }
        }
      }
      wave_sync();
#pragma unroll
      for (int i = 0; i < CH; ++i)
#pragma unroll
        for (int b = 0; b < NB; ++b) {
          redP[i * NB + b][lane] = sP[i][b];
          redM[i * NB + b][lane] = sM[i][b];
        }
      wave_sync();
      const cplx_t<T> tp = wave_rows<T, ROWS>(redP, lane);
      const cplx_t<T> tm = wave_rows<T, ROWS>(redM, lane);
      if (writer && ri < nl && rb < nt) {
        const int l = l0 + ri;
        const T c = (T)(0.5 * cf[3 * l + 2]);
        // aE = -(S+ + S-)/2, aB = (i/2)(S+ - S-)
        cplx_t<T> *o = oE + rb * out_ts;
        o[l] = {-c * (tp.x + tm.x), -c * (tp.y + tm.y)};
        o[nelem + l] = {-c * (tp.y - tm.y), c * (tp.x - tm.x)};
      }
    }
  }
}

// out[t, c, :] = sum over groups of partial[grp, t, c, :], fixed order.
template <typename T>
__global__ void sum_groups_kernel(const cplx_t<T> *part, long long grp_ts, int ngroups,
                                  cplx_t<T> *out, long long out_ts, long long part_ts,
                                  long long rowlen, long long nrows_per_t) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long t = blockIdx.y;
  if (i >= rowlen * nrows_per_t) return;
  cplx_t<T> acc = part[t * part_ts + i];
  for (int gi = 1; gi < ngroups; ++gi) {
    const cplx_t<T> v = part[gi * grp_ts + t * part_ts + i];
    acc.x += v.x;
    acc.y += v.y;
  }
  out[t * out_ts + i] = acc;
}

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

// Reduction rows per adjoint chunk (multipoles x transforms) and the
// widest transform batch per launch, bounded by LDS (4 waves x rows x 64
// lanes x complex, twice for spin 2) and registers.
template <typename T> struct AdjCfg;
template <> struct AdjCfg<double> {
  static constexpr int rows0 = 16, rows2 = 8, nb0 = 4, nb2 = 2;
};
template <> struct AdjCfg<float> {
  static constexpr int rows0 = 16, rows2 = 16, nb0 = 4, nb2 = 4;
};

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

template <typename T, int NB>
static void launch_synth_nb(owmi_plan_impl *p, int spin, const cplx_t<T> *alm, long long alm_ts,
                            cplx_t<T> *G, long long ntrans, hipStream_t st) {
  PlanTables<R> &tb = p->t64;  // recurrence tables (float64 for every precision)
  int bs = (int)std::min<long long>(256, ((p->npairs + 63) / 64) * 64);
  dim3 grid(ceil_div(p->npairs, bs), (unsigned)(p->mmax + 1), ceil_div(ntrans, NB));
  if (spin == 0)
    hipLaunchKernelGGL((synth0_kernel<T, NB>), grid, dim3(bs), 0, st, leg_args(p), tb.x, tb.lf0,
                       tb.s0, tb.coef0, alm, alm_ts, ntrans, G);
  else
    hipLaunchKernelGGL((synth2_kernel<T, NB>), grid, dim3(bs), 0, st, leg_args(p), tb.x, tb.lfp,
                       tb.sp, tb.lfm, tb.sm, tb.coef2, alm, alm_ts, ntrans, G);
}

// Transforms per block: the widest of 4, 2, 1 not exceeding ntrans.
template <typename T>
void launch_synthesis(owmi_plan_impl *p, int spin, const cplx_t<T> *alm, long long alm_ts,
                      cplx_t<T> *G, long long ntrans, hipStream_t st) {
  OWMI_REQUIRE(ntrans >= 1 && ntrans < 65536, "ntrans must be in [1, 65535]");
  if (ntrans >= 4)
    launch_synth_nb<T, 4>(p, spin, alm, alm_ts, G, ntrans, st);
  else if (ntrans >= 2)
    launch_synth_nb<T, 2>(p, spin, alm, alm_ts, G, ntrans, st);
  else
    launch_synth_nb<T, 1>(p, spin, alm, alm_ts, G, ntrans, st);
}

template <typename T, int SPIN, int K, int NB>
static void launch_adj_k(owmi_plan_impl *p, const cplx_t<T> *G, cplx_t<T> *alm,
                         long long alm_ts, long long ntrans, hipStream_t st) {
  PlanTables<R> &tb = p->t64;  // recurrence tables (float64 for every precision)
  AdjItems it;
  it.nmpairs = (p->mmax + 2) / 2;
  it.ngroups = (int)((p->npairs + 64LL * K - 1) / (64LL * K));
  const long long nelem = nelem_of(p->lmax);
  const int nc = SPIN == 0 ? 1 : 2;
  cplx_t<T> *dst = alm;
  long long dst_ts = alm_ts;
  if (it.ngroups > 1) {
    // partial sums [group][transform][comp][nelem] in the plan's second workspace
    const long long part_ts = nc * nelem;
    it.grp_ts = ntrans * part_ts;
    dst = static_cast<cplx_t<T> *>(
        p->ensure_work2((size_t)it.ngroups * it.grp_ts * sizeof(cplx_t<T>)));
    dst_ts = part_ts;
  } else {
    it.grp_ts = 0;
  }
  const long long nitems = (long long)it.nmpairs * it.ngroups;
  dim3 grid(ceil_div(nitems, kAdjWaves), ceil_div(ntrans, NB));
  if constexpr (SPIN == 0)
    hipLaunchKernelGGL((adj0_kernel<T, K, AdjCfg<T>::rows0 / NB, NB>), grid,
                       dim3(64 * kAdjWaves), 0, st, leg_args(p), it, tb.x, tb.lf0, tb.s0,
                       tb.coef0, G, ntrans, dst, dst_ts);
  else
    hipLaunchKernelGGL((adj2_kernel<T, K, AdjCfg<T>::rows2 / NB, NB>), grid,
                       dim3(64 * kAdjWaves), 0, st, leg_args(p), it, tb.x, tb.lfp, tb.sp, tb.lfm,
                       tb.sm, tb.coef2, G, ntrans, dst, dst_ts);
  if (it.ngroups > 1) {
    const long long n = nc * nelem;
    dim3 g2(ceil_div(n, 256), (unsigned)ntrans);
    hipLaunchKernelGGL(sum_groups_kernel<T>, g2, dim3(256), 0, st, dst, it.grp_ts, it.ngroups,
                       alm, alm_ts, dst_ts, nelem, (long long)nc);
  }
}

// Transforms per launch block: the widest of 4, 2, 1 allowed for the spin
// and precision and not exceeding ntrans.
template <typename T, int SPIN, int K>
static void launch_adj_nb(owmi_plan_impl *p, const cplx_t<T> *G, cplx_t<T> *alm,
                          long long alm_ts, long long ntrans, hipStream_t st) {
  constexpr int nbmax = SPIN == 0 ? AdjCfg<T>::nb0 : AdjCfg<T>::nb2;
  if constexpr (nbmax >= 4) {
    if (ntrans >= 4) return launch_adj_k<T, SPIN, K, 4>(p, G, alm, alm_ts, ntrans, st);
  }
  if (ntrans >= 2) return launch_adj_k<T, SPIN, K, 2>(p, G, alm, alm_ts, ntrans, st);
  launch_adj_k<T, SPIN, K, 1>(p, G, alm, alm_ts, ntrans, st);
}

template <typename T>
void launch_adjoint(owmi_plan_impl *p, int spin, const cplx_t<T> *G, cplx_t<T> *alm,
                    long long alm_ts, long long ntrans, hipStream_t st) {
  OWMI_REQUIRE(ntrans >= 1 && ntrans < 65536, "ntrans must be in [1, 65535]");
  // Ring pairs per lane: 2, or 1 on geometries of at most 64 ring pairs
  // (coarse or ring-subset plans), where two would leave most slots empty.
  // More pairs per lane amortise the cross-lane reduction further but cost
  // registers and waves (measured slower at lmax 1023).
  const long long np = p->npairs;
  if (spin == 0) {
    if (np <= 64) return launch_adj_nb<T, 0, 1>(p, G, alm, alm_ts, ntrans, st);
    if (np <= 128) return launch_adj_nb<T, 0, 2>(p, G, alm, alm_ts, ntrans, st);
    return launch_adj_nb<T, 0, 4>(p, G, alm, alm_ts, ntrans, st);
  }
  if (np <= 64) return launch_adj_nb<T, 2, 1>(p, G, alm, alm_ts, ntrans, st);
  launch_adj_nb<T, 2, 2>(p, G, alm, alm_ts, ntrans, st);
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
