// Legendre transforms between alm (triangular, m-major) and ring
// coefficients for spin 0 and spin 2 (HEALPix convention): the synthesis
// writes the pair-major layout of coef_syn_at, the adjoint reads the
// m-blocked pair-major layout of coef_pair_at.
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
#include <cstdlib>
#include <type_traits>

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

// The (north, south) ring coefficients of pair ip at multipole m in the
// adjoint's m-blocked pair-major layout (coef_pair_at).
template <typename T>
__device__ __forceinline__ const cplx_t<T> *gp_at(const cplx_t<T> *G, long long row, int m,
                                                  long long ip, const LegArgs &g) {
  return G + coef_pair_at<T>(row, m, ip, g.mmax, g.npairs);
}

// Stores the (north, south) values of one pair at Gp (coef_syn_at): one 16-
// or 32-byte store, or the north value alone for a pair without a south ring.
template <typename T>
__device__ __forceinline__ void store_pair(cplx_t<T> *Gp, cplx_t<T> n, cplx_t<T> s, bool has_s) {
  typedef T T4 __attribute__((ext_vector_type(4)));
  if (has_s)
    *reinterpret_cast<T4 *>(Gp) = T4{n.x, n.y, s.x, s.y};
  else
    Gp[0] = n;
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
  const bool has_s = g.pair_s[ip] >= 0;
#pragma unroll
  for (int b = 0; b < NB; ++b) {
    if (b < nt)
      store_pair<T>(G + coef_syn_at(t0 + b, m, ip, g.mmax, g.npairs),
                    {E[b].x + O[b].x, E[b].y + O[b].y}, {E[b].x - O[b].x, E[b].y - O[b].y},
                    has_s);
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
  const bool has_s = g.pair_s[ip] >= 0;
#pragma unroll
  for (int b = 0; b < NB; ++b) {
    if (b >= nt) continue;
    const long long t = t0 + b;
    // North: A+ = sum P 2lam, A- = sum Mm (-2)lam; Q = A+ + A-, U = -i (A+ - A-).
    const cplx_t<T> ApN = {PPe[b].x + PPo[b].x, PPe[b].y + PPo[b].y};
    const cplx_t<T> AmN = {MMe[b].x + MMo[b].x, MMe[b].y + MMo[b].y};
    // South: A+ = sum (-1)^(l+m) P (-2)lam, A- = sum (-1)^(l+m) Mm 2lam.
    const cplx_t<T> ApS = {PMe[b].x - PMo[b].x, PMe[b].y - PMo[b].y};
    const cplx_t<T> AmS = {MPe[b].x - MPo[b].x, MPe[b].y - MPo[b].y};
    store_pair<T>(G + coef_syn_at(2 * t, m, ip, g.mmax, g.npairs),
                  {ApN.x + AmN.x, ApN.y + AmN.y}, {ApS.x + AmS.x, ApS.y + AmS.y}, has_s);
    store_pair<T>(G + coef_syn_at(2 * t + 1, m, ip, g.mmax, g.npairs),
                  {ApN.y - AmN.y, -(ApN.x - AmN.x)}, {ApS.y - AmS.y, -(ApS.x - AmS.x)}, has_s);
  }
}

// ------------------------------------------------------- matrix-core synthesis
//
// float32 batches: the sums over l run on the matrix cores as
// D[i][j] += sum_k A[i][k] B[k][j] (v_mfma_f32_32x32x2f32) with
//   rows i    : 32 ring pairs (each wave owns 64 pairs: two row blocks),
//   depth k   : two multipoles (spin 0: l and l + 2 of one parity of l - m;
//               spin 2: l and l + 1),
//   columns j : 32 staged alm components.
// Operand layouts of this instruction: A[i][k] in lane i + 32k, B[k][j] in
// lane j + 32k, D[i][j] in lane j + 32 ((i / 4) % 2), register
// (i % 4) + 4 (i / 8).
// Every lane runs the float64 recurrence of its own ring pair; one
// v_permlane32_swap per function and multipole pair moves the values of 64
// pairs into the A operands of both row blocks.
//   spin 0 (16 transforms): j = 2t + (re, im) of c_l a_lm; the even and odd
//     sums over l - m accumulate separately, north = even + odd,
//     south = even - odd;
//   spin 2 (8 transforms): the south ring uses columns that carry
//     s = (-1)^(l-m): block 0 multiplies 2lam with (P, s Mm), block 1
//     multiplies (-2)lam with (Mm, s P) (P, Mm as in synth2_kernel), j = 4t + c,
//     so one lane ends with the terms of one Q/U component of one ring.

typedef float f32x16 __attribute__((ext_vector_type(16)));

// Workgroups go to the 8 XCDs (each with its own L2) round-robin by linear
// block index. Blocks of neighbouring m of the adjoint share the cache
// lines of the ring coefficients (4 m per 64-byte segment of a pair), so
// the (x, m) of a block are chosen such that every XCD runs 16 consecutive
// m in a row (all x of one m together), while each XCD still gets m from the
// whole range (the work per m falls with m). m beyond the last full group of
// 128 keep the plain order.
__device__ __forceinline__ void xcd_block(int gx, int nm, int &x, int &m) {
  const int lin = blockIdx.x + gx * blockIdx.y;
  const int full = (nm / 128) * 128;
  if (lin >= gx * full) {
    x = blockIdx.x;
    m = blockIdx.y;
    return;
  }
  const int xcd = lin & 7, r = lin >> 3;
  x = r % gx;
  const int t = r / gx;
  m = (t >> 4) * 128 + xcd * 16 + (t & 15);
}

constexpr int kMfCH = 64;  // multipoles staged per chunk (even)

// lo: lanes 0-31 v, lanes 32-63 w of lanes 0-31; hi: lanes 0-31 v of lanes
// 32-63, lanes 32-63 w.
__device__ __forceinline__ void swap_halves(float v, float w, float &lo, float &hi) {
  const auto r =
      __builtin_amdgcn_permlane32_swap(__float_as_uint(v), __float_as_uint(w), false, false);
  lo = __uint_as_float(r[0]);
  hi = __uint_as_float(r[1]);
}

__device__ __forceinline__ f32x16 mfma32(float a, float b, f32x16 c) {
  return __builtin_amdgcn_mfma_f32_32x32x2f32(a, b, c, 0, 0, 0);
}

// One recurrence step with the start-multipole injection.
__device__ __forceinline__ R rec_step(R coefx, R &l1, R &l2, int l, int lf, R s0, R s1) {
  const R lam = fma(coefx, l1, -l2);
  const bool st = (l == lf);
  l2 = st ? s0 : l1;
  l1 = st ? s1 : lam;
  return l1;
}

// PF: staging items per thread when the block has 256 threads, loaded one
// chunk ahead into registers so that no chunk waits on global memory; 0:
// stage each chunk directly (blocks of fewer threads).
template <int SPIN, int PF>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2)))
synth_mfma_kernel(LegArgs g, const R *__restrict__ xs, const int *__restrict__ lfas,
                  const R *__restrict__ sas, const int *__restrict__ lfbs,
                  const R *__restrict__ sbs, const R *__restrict__ coef,
                  const float2 *__restrict__ alm, long long alm_ts, long long ntrans,
                  float2 *G) {
  constexpr int NT = 16;                    // transforms per block
  constexpr int NCF = SPIN == 0 ? 2 : 3;    // recurrence coefficients per multipole
  constexpr int STEP = SPIN == 0 ? 4 : 2;   // multipoles per group of products
  constexpr int NCB = SPIN == 0 ? 1 : 4;    // column blocks of B
  // recurrence coefficients of the current and the next chunk (the
  // recurrence runs one group of multipoles ahead of the products)
  __shared__ R sA[2][kMfCH];
  __shared__ R sC[2][kMfCH];
  // 4 padding rows: the B reads of the group after a chunk's last
  __shared__ __align__(16) float sB[NCB][kMfCH + 4][32];
  __shared__ int smin;
  const int lane = threadIdx.x & 63;
  int bx, m;
  xcd_block(gridDim.x, gridDim.y, bx, m);
  const long long ip = (long long)bx * blockDim.x + threadIdx.x;
  const long long t0 = (long long)blockIdx.z * NT;
  const int nt = (int)min((long long)NT, ntrans - t0);
  const bool active = ip < g.npairs;
  const int lmax = g.lmax;
  const long long nelem = nelem_of(lmax);
  const long long ms = mstart_of(m, lmax);
  const long long si = (long long)m * g.npairs + ip;
  const int lfa = active ? lfas[si] : INT_MAX;
  const int lfb = (SPIN == 2 && active) ? lfbs[si] : INT_MAX;
  const int lmin = block_min(min(lfa, lfb), &smin);
  // acc[c][h]: row block h; spin 0: c = sums over l - lmin even / odd;
  // spin 2: c = 2 set + function for the transform sets 0-7, 8-15
  constexpr int NACC = SPIN == 0 ? 2 : 4;
  f32x16 acc[NACC][2];
#pragma unroll
  for (int c = 0; c < NACC; ++c) acc[c][0] = acc[c][1] = f32x16{};
  if (lmin <= lmax) {
    const R x = active ? xs[ip] : (R)0;
    const R a0s = active ? sas[2 * si] : (R)0, a1s = active ? sas[2 * si + 1] : (R)0;
    R b0s = 0, b1s = 0;
    if (SPIN == 2 && active) {
      b0s = sbs[2 * si];
      b1s = sbs[2 * si + 1];
    }
    const float2 *a = alm + t0 * alm_ts + ms;
    const R *cf = coef + NCF * ms;
    const int jc = lane & 31, kh = lane >> 5;
    R a1 = 0, a2 = 0, b1 = 0, b2 = 0;
    auto stage_coef = [&](int l0, int buf) {
      for (int j = threadIdx.x; j < kMfCH; j += blockDim.x) {
        const int l = l0 + j;
        const bool in = l <= lmax;
        sA[buf][j] = in ? cf[NCF * l] : (R)0;
        if (SPIN == 2) sC[buf][j] = in ? cf[NCF * l + 1] : (R)0;
      }
    };
    // staging item j of a chunk: (multipole j / NT, transform j % NT); the
    // raw values (alm, and for spin 2 the B alm) and the normalisation
    struct Item {
      float2 e, b;
      float c;
    };
    auto load_item = [&](int l0, int j) {
      const int b = j % NT, l = l0 + j / NT;
      Item it = {{0, 0}, {0, 0}, 0.0f};
      if (l <= lmax && b < nt) {
        it.e = a[b * alm_ts + l];
        if (SPIN == 2) it.b = a[b * alm_ts + nelem + l];
        it.c = (float)(SPIN == 0 ? cf[2 * l + 1] : -0.5 * cf[3 * l + 2]);
      }
      return it;
    };
    auto store_item = [&](int l0, int j, const Item &it) {
      const int b = j % NT, jl = j / NT;
      const float c = it.c;
      if (SPIN == 0) {
        *reinterpret_cast<float2 *>(&sB[0][jl][2 * b]) = float2{c * it.e.x, c * it.e.y};
      } else {
        const float s = ((l0 + jl - m) & 1) ? -1.0f : 1.0f;
        const float2 e = it.e, bb = it.b;
        // P = -(e + i b)/2, Mm = -(e - i b)/2 (the -1/2 is in c)
        const float2 P = {c * (e.x - bb.y), c * (e.y + bb.x)};
        const float2 Mm = {c * (e.x + bb.y), c * (e.y - bb.x)};
        const int set = b >> 3, bt = b & 7;
        *reinterpret_cast<float4 *>(&sB[(2 * set) % NCB][jl][4 * bt]) =
            float4{P.x, P.y, s * Mm.x, s * Mm.y};
        *reinterpret_cast<float4 *>(&sB[(2 * set + 1) % NCB][jl][4 * bt]) =
            float4{Mm.x, Mm.y, s * P.x, s * P.y};
      }
    };
    auto stage_alm = [&](int l0) {
      for (int j = threadIdx.x; j < kMfCH * NT; j += blockDim.x)
        store_item(l0, j, load_item(l0, j));
    };
    // prefetch registers: items threadIdx.x + 256 k of the next chunk and
    // the coefficients of the chunk after it
    Item pre[PF > 0 ? PF : 1];
    R pca = 0, pcc = 0;
    auto prefetch = [&](int l0n) {
#pragma unroll
      for (int k = 0; k < PF; ++k) pre[k] = load_item(l0n, (int)threadIdx.x + 256 * k);
      if (threadIdx.x < kMfCH) {
        const int l = l0n + kMfCH + (int)threadIdx.x;
        pca = l <= lmax ? cf[NCF * l] : (R)0;
        pcc = (SPIN == 2 && l <= lmax) ? cf[NCF * l + 1] : (R)0;
      }
    };
    auto commit = [&](int l0n, int cbuf) {
#pragma unroll
      for (int k = 0; k < PF; ++k) store_item(l0n, (int)threadIdx.x + 256 * k, pre[k]);
      if (threadIdx.x < kMfCH) {
        sA[cbuf][threadIdx.x] = pca;
        if (SPIN == 2) sC[cbuf][threadIdx.x] = pcc;
      }
    };
    // A operands of one group of STEP multipoles starting at l (coefficients
    // from kA / kC): spin 0: (alo, ahi) = rows l, l + 2 and (blo, bhi) =
    // rows l + 1, l + 3; spin 2: (alo, ahi) = 2lam and (blo, bhi) = (-2)lam
    // at l, l + 1
    // A operands of one group of STEP multipoles starting at l from its
    // recurrence coefficients kc (spin 0: A' at l .. l + 3; spin 2: A' at l,
    // l + 1, B' at l, l + 1). INJ: the group may hold a start multipole.
    auto operands = [&](auto inj_tag, const R *kc, int l, float &alo, float &ahi, float &blo,
                        float &bhi) {
      constexpr bool INJ = decltype(inj_tag)::value;
      auto step = [&](R cx, R &l1, R &l2, int ll, int lf, R s0, R s1) -> R {
        if (INJ) return rec_step(cx, l1, l2, ll, lf, s0, s1);
        const R lam = fma(cx, l1, -l2);
        l2 = l1;
        l1 = lam;
        return lam;
      };
      if (SPIN == 0) {
        const R u0 = step(kc[0] * x, a1, a2, l, lfa, a0s, a1s);
        const R u1 = step(kc[1] * x, a1, a2, l + 1, lfa, a0s, a1s);
        const R u2 = step(kc[2] * x, a1, a2, l + 2, lfa, a0s, a1s);
        const R u3 = step(kc[3] * x, a1, a2, l + 3, lfa, a0s, a1s);
        swap_halves((float)u0, (float)u2, alo, ahi);
        swap_halves((float)u1, (float)u3, blo, bhi);
      } else {
        const R ax0 = kc[0] * x, ax1 = kc[1] * x;
        const R p0 = step(ax0 + kc[2], a1, a2, l, lfa, a0s, a1s);
        const R q0 = step(ax0 - kc[2], b1, b2, l, lfb, b0s, b1s);
        const R p1 = step(ax1 + kc[3], a1, a2, l + 1, lfa, a0s, a1s);
        const R q1 = step(ax1 - kc[3], b1, b2, l + 1, lfb, b0s, b1s);
        swap_halves((float)p0, (float)p1, alo, ahi);
        swap_halves((float)q0, (float)q1, blo, bhi);
      }
    };
    // coefficients of the group at chunk row gi (rows from kMfCH on belong
    // to the next chunk, staged one chunk ahead)
    auto load_coefs = [&](int gi, int cb, R *kc) {
      const bool wrap = gi >= kMfCH;
      const int b = wrap ? cb ^ 1 : cb, r = wrap ? gi - kMfCH : gi;
      if (SPIN == 0) {
        kc[0] = sA[b][r];
        kc[1] = sA[b][r + 1];
        kc[2] = sA[b][r + 2];
        kc[3] = sA[b][r + 3];
      } else {
        kc[0] = sA[b][r];
        kc[1] = sA[b][r + 1];
        kc[2] = sC[b][r];
        kc[3] = sC[b][r + 1];
      }
    };
    // B operands of the group at chunk row gi
    auto load_bv = [&](int gi, float *bv) {
      if (SPIN == 0) {
        // depth k: multipoles l + 2k (acc[0]: l - l0 even) and l + 1 + 2k (acc[1])
        bv[0] = sB[0][gi + 2 * kh][jc];
        bv[1] = sB[0][gi + 1 + 2 * kh][jc];
      } else {
#pragma unroll
        for (int c = 0; c < NACC; ++c) bv[c] = sB[c % NCB][gi + kh][jc];
      }
    };
    stage_coef(lmin, 0);
    stage_coef(lmin + kMfCH, 1);
    stage_alm(lmin);
    if (PF > 0) prefetch(lmin + kMfCH);
    __syncthreads();
    // software pipeline: the products of group g run with the recurrence of
    // group g + 1, the coefficients of group g + 2 and the B operands of
    // group g + 1 in flight from LDS
    R kc[4];
    float av[4];  // A operands (alo, ahi, blo, bhi) of the current group
    load_coefs(0, 0, kc);
    operands(std::true_type{}, kc, lmin, av[0], av[1], av[2], av[3]);
    load_coefs(STEP, 0, kc);
    // start multipoles of this lane within [lo, lo + n)
    auto starts_in = [&](int lo, int n) {
      return active && ((lfa >= lo && lfa < lo + n) || (lfb >= lo && lfb < lo + n));
    };
    bool slow = __syncthreads_or(starts_in(lmin, kMfCH + STEP));
    for (int ch = 0, l0 = lmin; l0 <= lmax; ++ch, l0 += kMfCH) {
      const int cb = ch & 1;
      if (ch > 0) {
        __syncthreads();
        if (PF > 0) {
          commit(l0, cb ^ 1);
          if (l0 + kMfCH <= lmax) prefetch(l0 + kMfCH);
        } else {
          stage_coef(l0 + kMfCH, cb ^ 1);
          stage_alm(l0);
        }
        slow = __syncthreads_or(starts_in(l0, kMfCH + STEP));
      }
      const int nls = (min(kMfCH, lmax - l0 + 1) + STEP - 1) & ~(STEP - 1);
      float bv[NACC];
      load_bv(0, bv);
      // one group: the products of the current operands (a, b) while the
      // next group's operands (na, nb) are formed; the B rows of the group
      // after a chunk's last read the padding rows of sB (never used)
      auto group = [&](int i, const float(&a)[4], const float(&b)[NACC], float(&na)[4],
                       float(&nb)[NACC]) {
        if (slow)
          operands(std::true_type{}, kc, l0 + i + STEP, na[0], na[1], na[2], na[3]);
        else
          operands(std::false_type{}, kc, l0 + i + STEP, na[0], na[1], na[2], na[3]);
        load_coefs(i + 2 * STEP, cb, kc);
        load_bv(i + STEP, nb);
#pragma unroll
        for (int c = 0; c < NACC; c += 2) {
          acc[c][0] = mfma32(a[0], b[c], acc[c][0]);
          acc[c][1] = mfma32(a[1], b[c], acc[c][1]);
          acc[c + 1][0] = mfma32(a[2], b[c + 1], acc[c + 1][0]);
          acc[c + 1][1] = mfma32(a[3], b[c + 1], acc[c + 1][1]);
        }
      };
      // two groups per iteration, the operand sets alternating (no copies)
      float a2[4], b2[NACC];
      int i = 0;
      for (; i + 2 * STEP <= nls; i += 2 * STEP) {
        group(i, av, bv, a2, b2);
        group(i + STEP, a2, b2, av, bv);
      }
      if (i < nls) {
        group(i, av, bv, a2, b2);
#pragma unroll
        for (int k = 0; k < 4; ++k) av[k] = a2[k];
      }
    }
  }
  // D[i][j]: row i = 8 (r / 4) + 4 (lane / 32) + r % 4 of the row block
  const int jc = lane & 31, c = jc & 3, tt = jc >> 2;
  // spin 0: sums over l - lmin even / odd; the south ring takes their
  // difference with the sign of (-1)^(lmin - m)
  const float sgn0 = ((lmin - m) & 1) ? -1.0f : 1.0f;
  const long long wbase = (long long)bx * blockDim.x + (threadIdx.x & ~63);
  const bool odd = jc & 1;
#pragma unroll
  for (int h = 0; h < 2; ++h) {
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const long long pp = wbase + 32 * h + 8 * (r / 4) + 4 * (lane >> 5) + (r % 4);
      // lanes 2n and 2n + 1 hold the two components of one complex value;
      // after the exchange each lane stores one whole value
      if (SPIN == 0) {
        // column jc: transform jc / 2, component jc % 2; the even lane stores
        // the north value, the odd lane the south value
        const float v0 = acc[0][h][r], v1 = acc[1][h][r];
        const float vn = v0 + v1, vs = sgn0 * (v0 - v1);
        const float send = odd ? vn : vs;
        const float recv = __shfl_xor(send, 1);
        const int b = jc >> 1;
        if (pp >= g.npairs || b >= nt) continue;
        const long long t = t0 + b;
        const int ring = odd ? g.pair_s[pp] : g.pair_n[pp];
        if (ring < 0) continue;
        const float2 val = odd ? float2{recv, vs} : float2{vn, recv};
        G[coef_syn_at(t, m, pp, g.mmax, g.npairs) + odd] = val;
      } else {
        // column c = jc % 4 of transform tt of each set: q component c % 2
        // and u component 1 - c % 2 of ring (c < 2: north, else south); the
        // even lane stores Q, the odd lane U
        const int ring = c < 2 ? (pp < g.npairs ? g.pair_n[pp] : -1)
                               : (pp < g.npairs ? g.pair_s[pp] : -1);
#pragma unroll
        for (int set = 0; set < 2; ++set) {
          const float v0 = acc[2 * set][h][r], v1 = acc[2 * set + 1][h][r];
          const float q = v0 + v1;
          const float u = (c == 0 || c == 3) ? v1 - v0 : v0 - v1;
          const float send = odd ? q : u;
          const float recv = __shfl_xor(send, 1);
          const int b = 8 * set + tt;
          if (ring < 0 || b >= nt) continue;
          const long long t = t0 + b;
          // side c / 2 of the pair in row 2t (Q) or 2t + 1 (U)
          const long long qi = coef_syn_at(2 * t + odd, m, pp, g.mmax, g.npairs) + (c >> 1);
          G[qi] = odd ? float2{u, recv}    // U = (u of c + 1, u of c)
                      : float2{q, recv};   // Q = (q of c, q of c + 1)
        }
      }
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
              const cplx_t<T> *gp = gp_at<T>(G, t0 + b, m, ip, g);
              const cplx_t<T> gn = gp[0];
              if (rs >= 0) {
                const cplx_t<T> gs = gp[1];
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
              const cplx_t<T> *Gq = gp_at<T>(G, (t0 + b) * 2, m, ip, g);
              const cplx_t<T> *Gu = gp_at<T>(G, (t0 + b) * 2 + 1, m, ip, g);
              const cplx_t<T> qn = Gq[0], un = Gu[0];
              ZpN[k][b] = {qn.x - un.y, qn.y + un.x};
              ZmN[k][b] = {qn.x + un.y, qn.y - un.x};
              if (rs >= 0) {
                const cplx_t<T> qs = Gq[1], us = Gu[1];
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

// ---------------------------------------------------- matrix-core adjoint
//
// Batches: the sums over ring pairs run on the matrix cores as
// D[i][j] += sum_k A[i][k] B[k][j] (v_mfma_f32_16x16x4f32 for float32,
// v_mfma_f64_16x16x4f64 for float64; both take A[i][k] in lane i + 16k and
// B[k][j] in lane j + 16k; float32 returns D[i][j] in lane j + 16 (i / 4),
// register i % 4, float64 in lane j + 16 (i % 4), register i / 4) with
//   rows i    : 16 multipoles of one parity of l - m (two row tiles per
//               chunk of 32 multipoles),
//   depth k   : spin 0: 4 ring pairs; spin 2: 2 ring pairs x (2lam, s (-2)lam),
//   columns j : 16 per tile; spin 0: (re, im) of 8 transforms with the
//               weights N + S (even rows) or N - S (odd rows) of the north
//               and south ring coefficients; spin 2: 4 transforms x
//               (S+ re, S+ im, s S- re, s S- im), s = (-1)^(l-m).
// Spin 2 folds s into the (-2)lam operand: S+ = sum 2lam Zp(N) + s (-2)lam Zp(S),
// s S- = sum 2lam Zm(S) + s (-2)lam Zm(N), so its weights do not depend on the
// row parity.
// Each wave owns 64 ring pairs: its lanes run the float64 recurrences of
// their pairs over a chunk and write the values (float32 or float64) to a
// wave-private LDS tile [multipole][pair]; the matrix products read the A operands from
// that tile and keep the B operands (the weights, fixed per m) in registers.
// The W waves of a block (64 W ring pairs) add their partial tiles through
// LDS after every chunk; blocks of further ring pairs write partial sums
// that sum_groups_kernel adds in fixed order.

// Copies 16 bytes per lane from global memory (src, per lane) to LDS at
// dst + 16 lane (dst wave-uniform) without passing through registers
// (global_load_lds_dwordx4); complete after a wait for vmcnt 0.
__device__ __forceinline__ void lds_dma16(const void *src, void *dst) {
  __builtin_amdgcn_global_load_lds(src, (__attribute__((address_space(3))) void *)dst, 16, 0, 0);
}
// s_waitcnt vmcnt(0) / lgkmcnt(0) with the other counters left alone
constexpr int kWaitVm0 = 0x0F70;
constexpr int kWaitLgkm0 = 0xC07F;

// Minimum of v over the block with one LDS atomic per wave.
__device__ __forceinline__ int block_min_waves(int v, int *slot) {
  v = wave_min(v);
  if (threadIdx.x == 0) *slot = INT_MAX;
  __syncthreads();
  if ((threadIdx.x & 63) == 0) atomicMin(slot, v);
  __syncthreads();
  return *slot;
}

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef double f64x4 __attribute__((ext_vector_type(4)));

template <typename T> struct Mfma16;
template <> struct Mfma16<float> {
  using acc = f32x4;
  static __device__ __forceinline__ acc run(float a, float b, acc c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
};
template <> struct Mfma16<double> {
  using acc = f64x4;
  static __device__ __forceinline__ acc run(double a, double b, acc c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
};

// ------------------------------------------- float64 synthesis, matrix cores
//
// D[pair][column] += A[pair][depth] B[depth][column] on v_mfma_f64_16x16x4f64
// tiles: A holds the float64 Legendre values of 16 ring pairs at 4
// multipoles, B the normalised alm of those multipoles (spin 0: columns
// (transform, re/im) of 8 transforms per tile; spin 2: 4 transforms per tile
// with columns (P, s Mm) / (Mm, s P) as the float32 kernel). Every wave is
// independent (no block barrier): it runs the recurrence of its own 64 ring
// pairs, 8 multipoles at a time, into a wave-private LDS tile whose 16-pair
// column blocks are swizzled by depth (conflict-free reads of the 4 depths),
// and loads its B operands straight from global memory one group of
// multipoles ahead (the 4 waves of a block read the same alm, so the loads
// of 3 of them hit L1). Groups start at l - m even: the two depth sets of a
// group (l - m even / odd) are the spin-0 sums E and O (north E + O, south
// E - O); at spin 2 both sets accumulate into the tile of their function.
// Four row tiles (64 pairs) and NCT column tiles: 64 NCT accumulator
// registers per lane.
constexpr int kS64G = 8;  // multipoles per group

template <int SPIN, int NCT>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2)))
synth_mfma64_kernel(LegArgs g, const R *__restrict__ xs, const int *__restrict__ lfas,
                    const R *__restrict__ sas, const int *__restrict__ lfbs,
                    const R *__restrict__ sbs, const R *__restrict__ coef,
                    const double2 *__restrict__ alm, long long alm_ts, long long ntrans,
                    double2 *G) {
  using V4 = f64x4;
  constexpr int NF = SPIN == 0 ? 1 : 2;           // Legendre functions per pair
  constexpr int NT = (SPIN == 0 ? 8 : 4) * NCT;   // transforms per block
  constexpr int NCF = SPIN == 0 ? 2 : 3;          // recurrence coefficients per multipole
  constexpr int NACC = SPIN == 0 ? 2 : 2;         // spin 0: E / O; spin 2: functions
  __shared__ __align__(16) double Tall[4][NF][kS64G][64];
  const int lane = threadIdx.x & 63;
  const int wu = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  double(*T)[kS64G][64] = Tall[wu];
  int bx, m;
  xcd_block(gridDim.x, gridDim.y, bx, m);
  const long long wbase = (long long)bx * blockDim.x + 64 * wu;
  const long long ip = wbase + lane;
  const long long t0 = (long long)blockIdx.z * NT;
  const int nt = (int)min((long long)NT, ntrans - t0);
  const bool active = ip < g.npairs;
  const int lmax = g.lmax;
  const long long nelem = nelem_of(lmax);
  const long long ms = mstart_of(m, lmax);
  const long long si = (long long)m * g.npairs + ip;
  const int lfa = active ? lfas[si] : INT_MAX;
  const int lfb = (SPIN == 2 && active) ? lfbs[si] : INT_MAX;
  const int lmin = __builtin_amdgcn_readfirstlane(wave_min(min(lfa, lfb)));
  const int bj = lane & 15, bk = lane >> 4;
  V4 D[4][NCT][NACC];
#pragma unroll
  for (int r = 0; r < 4; ++r)
#pragma unroll
    for (int ct = 0; ct < NCT; ++ct) D[r][ct][0] = D[r][ct][1] = V4{0, 0, 0, 0};
  if (lmin <= lmax) {  // wave-uniform
    const R x = active ? xs[ip] : (R)0;
    const R a0s = active ? sas[2 * si] : (R)0, a1s = active ? sas[2 * si + 1] : (R)0;
    R b0s = 0, b1s = 0;
    if (SPIN == 2 && active) {
      b0s = sbs[2 * si];
      b1s = sbs[2 * si + 1];
    }
    const R *cf = coef + NCF * ms;
    const double2 *a = alm + t0 * alm_ts + ms;
    const int lb0 = lmin - ((lmin - m) & 1);
    // B-operand column of this lane: transform tcol[ct] of the block,
    // component comp of its alm (spin 2: of the E alm; the B alm's other
    // component pairs with it), ring side (spin 2) from the column
    const int comp = bj & 1;
    int tcol[NCT];
    bool tin[NCT];
#pragma unroll
    for (int ct = 0; ct < NCT; ++ct) {
      const int tl = SPIN == 0 ? 8 * ct + (bj >> 1) : 4 * ct + (bj >> 2);
      tin[ct] = tl < nt;
      tcol[ct] = min(tl, nt - 1);
    }
    // raw values of a group: per depth set the normalisation (rc) and per
    // column tile the alm component(s) (ra: E or the spin-0 alm, rb: B)
    R rc[2], ra[2][NCT], rb[2][NCT];
    auto load_raw = [&](int lg) {
#pragma unroll
      for (int par = 0; par < 2; ++par) {
        const int lc = min(lg + 2 * bk + par, lmax);
        rc[par] = cf[NCF * lc + NCF - 1];
#pragma unroll
        for (int ct = 0; ct < NCT; ++ct) {
          const double *ae = reinterpret_cast<const double *>(a + tcol[ct] * alm_ts + lc);
          ra[par][ct] = ae[comp];
          if (SPIN == 2) {
            const double *ab =
                reinterpret_cast<const double *>(a + tcol[ct] * alm_ts + nelem + lc);
            rb[par][ct] = ab[1 - comp];
          }
        }
      }
    };
    R a1 = 0, a2 = 0, b1 = 0, b2 = 0;
    load_raw(lb0);
    for (int lg = lb0; lg <= lmax; lg += kS64G) {
      // B operands of this group: Bv[f][par][ct]
      R Bv[NF][2][NCT];
#pragma unroll
      for (int par = 0; par < 2; ++par) {
        const bool lin = lg + 2 * bk + par <= lmax;
#pragma unroll
        for (int ct = 0; ct < NCT; ++ct) {
          const bool in = lin && tin[ct];
          if (SPIN == 0) {
            Bv[0][par][ct] = in ? rc[par] * ra[par][ct] : 0.0;
          } else {
            // P = -(e + i b) / 2, Mm = -(e - i b) / 2 (the -1/2 with the
            // normalisation): component comp of P is e_c - b_{1-c} (c = 0)
            // or e_c + b_{1-c} (c = 1); columns 2, 3 of a transform take the
            // south ring's sign s = (-1)^(l - m)
            const R cl = -0.5 * rc[par];
            const R sb = comp ? rb[par][ct] : -rb[par][ct];
            const R P = cl * (ra[par][ct] + sb), Mm = cl * (ra[par][ct] - sb);
            const bool south = (bj & 2) != 0;
            const R sg = par ? -1.0 : 1.0;
            Bv[0][par][ct] = in ? (south ? sg * Mm : P) : 0.0;
            Bv[SPIN == 2 ? 1 : 0][par][ct] = in ? (south ? sg * P : Mm) : 0.0;
          }
        }
      }
      if (lg + kS64G <= lmax) load_raw(lg + kS64G);
      // recurrence of the lane's ring pair over the group -> tile rows
      // (column block swizzled by depth: row i, lane L at L ^ 16 ((i / 2) % 4))
      const bool inj = __any((lfa >= lg && lfa < lg + kS64G) || (lfb >= lg && lfb < lg + kS64G));
#pragma unroll
      for (int i = 0; i < kS64G; ++i) {
        const int l = lg + i;
        const int lc = min(l, lmax);
        const int col = lane ^ (16 * ((i >> 1) & 3));
        const bool lin = l <= lmax;
        if (SPIN == 0) {
          const R ax = cf[2 * lc] * x;
          R v;
          if (inj) {
            v = rec_step(ax, a1, a2, l, lfa, a0s, a1s);
          } else {
            v = fma(ax, a1, -a2);
            a2 = a1;
            a1 = v;
          }
          T[0][i][col] = lin ? v : 0.0;
        } else {
          const R ax = cf[3 * lc] * x, bc = cf[3 * lc + 1];
          R p, q;
          if (inj) {
            p = rec_step(ax + bc, a1, a2, l, lfa, a0s, a1s);
            q = rec_step(ax - bc, b1, b2, l, lfb, b0s, b1s);
          } else {
            p = fma(ax + bc, a1, -a2);
            q = fma(ax - bc, b1, -b2);
            a2 = a1;
            a1 = p;
            b2 = b1;
            b1 = q;
          }
          T[0][i][col] = lin ? p : 0.0;
          T[SPIN == 2 ? 1 : 0][i][col] = lin ? q : 0.0;
        }
      }
      wave_sync();
      // products: depth k = bk of set par is row 2k + par
#pragma unroll
      for (int par = 0; par < 2; ++par) {
#pragma unroll
        for (int f = 0; f < NF; ++f) {
          double A[4];
#pragma unroll
          for (int r = 0; r < 4; ++r) A[r] = T[f][2 * bk + par][(16 * r + bj) ^ (16 * bk)];
#pragma unroll
          for (int r = 0; r < 4; ++r)
#pragma unroll
            for (int ct = 0; ct < NCT; ++ct) {
              const int c = SPIN == 0 ? par : f;
              D[r][ct][c] = __builtin_amdgcn_mfma_f64_16x16x4f64(A[r], Bv[f][par][ct],
                                                                 D[r][ct][c], 0, 0, 0);
            }
        }
      }
      wave_sync();
    }
  }
  // D[i][j] of row tile r: register i / 4 at lane j + 16 (i % 4); pair
  // wbase + 16 r + 4 reg + bk, column bj
  const bool odd = bj & 1;
#pragma unroll
  for (int r = 0; r < 4; ++r)
#pragma unroll
    for (int reg = 0; reg < 4; ++reg) {
      const long long pp = wbase + 16 * r + 4 * reg + bk;
#pragma unroll
      for (int ct = 0; ct < NCT; ++ct) {
        const double v0 = D[r][ct][0][reg], v1 = D[r][ct][1][reg];
        if (SPIN == 0) {
          // column bj: transform 8 ct + bj / 2, component bj % 2; the even
          // lane stores the north value E + O, the odd lane the south E - O
          const double vn = v0 + v1, vs = v0 - v1;
          const double recv = __shfl_xor(odd ? vn : vs, 1);
          const int b = 8 * ct + (bj >> 1);
          if (pp >= g.npairs || b >= nt) continue;
          const int ring = odd ? g.pair_s[pp] : g.pair_n[pp];
          if (ring < 0) continue;
          G[coef_syn_at(t0 + b, m, pp, g.mmax, g.npairs) + odd] =
              odd ? double2{recv, vs} : double2{vn, recv};
        } else {
          // column c = bj % 4 of transform 4 ct + bj / 4: q component c % 2
          // and u component 1 - c % 2 of ring side c / 2; the even lane
          // stores Q, the odd lane U
          const int c = bj & 3;
          const double q = v0 + v1;
          const double u = (c == 0 || c == 3) ? v1 - v0 : v0 - v1;
          const double recv = __shfl_xor(odd ? q : u, 1);
          const int b = 4 * ct + (bj >> 2);
          if (pp >= g.npairs || b >= nt) continue;
          const int ring = c < 2 ? g.pair_n[pp] : g.pair_s[pp];
          if (ring < 0) continue;
          const long long qi =
              coef_syn_at(2 * (t0 + b) + odd, m, pp, g.mmax, g.npairs) + (c >> 1);
          G[qi] = odd ? double2{u, recv} : double2{q, recv};
        }
      }
    }
}

constexpr int kMaCH = 32;  // multipoles per adjoint chunk

// E: element type (float: float32 alm and ring coefficients, float64
// recurrence; double: float64 throughout, float64 matrix cores). CT: column
// tiles of 16 per block (transforms per block: 8 CT for spin 0, 4 CT for
// spin 2). More tiles share one recurrence and one cross-wave reduction pass
// between more products, at 32 CT registers of weights.
template <typename E, int SPIN, int W, int CT>
__global__ void __launch_bounds__(64 * W)
adj_mfma_kernel(LegArgs g, long long grp_ts, const R *__restrict__ xs,
                const int *__restrict__ lfas, const R *__restrict__ sas,
                const int *__restrict__ lfbs, const R *__restrict__ sbs,
                const R *__restrict__ coef, const cplx_t<E> *__restrict__ G, long long ntrans,
                cplx_t<E> *out, long long out_ts) {
  using C = cplx_t<E>;
  using V4 = typename Mfma16<E>::acc;
  constexpr int NT = (SPIN == 0 ? 8 : 4) * CT;    // transforms per block
  constexpr int NCF = SPIN == 0 ? 2 : 3;    // recurrence coefficients per multipole
  constexpr int TS = SPIN == 0 ? 66 : 130;  // tile row stride in elements
  // tile elements per wave (spin 2: room for the staged weights of 8
  // transforms, 66 pair slots per row)
  constexpr int TW = SPIN == 0 ? kMaCH * TS : 2 * 2 * 8 * 66 * 2;
  static_assert(TW >= kMaCH * TS, "the tile fits");
  static_assert(TW >= 1024, "the reduction buffer overlays the tile");
  __shared__ __align__(32) E Tall[W * TW];
  // recurrence coefficients per chunk buffer and multipole: (A', B') for
  // spin 2, (A', c) for spin 0; (B', c) for spin 2 (c: the output's
  // normalisation), copied from the plan's table without registers
  __shared__ __align__(16) cplx_t<R> cAB[3][kMaCH];
  __shared__ __align__(16) cplx_t<R> cBC[SPIN == 2 ? 3 : 1][kMaCH];
  __shared__ int smin;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  E *T = Tall + wave * TW;
  int grp, m;
  xcd_block(gridDim.x, gridDim.y, grp, m);
  const long long wbase = (long long)grp * 64 * W + 64 * wave;
  const long long ip = wbase + lane;
  const long long t0 = (long long)blockIdx.z * NT;
  const int nt = (int)min((long long)NT, ntrans - t0);
  const bool active = ip < g.npairs;
  const int lmax = g.lmax;
  const long long nelem = nelem_of(lmax);
  const long long ms = mstart_of(m, lmax);
  const long long stride = g.mmax + 1;
  const long long si = (long long)m * g.npairs + ip;
  // Loads of inactive lanes read the last pair (their Legendre values stay 0).
  const long long ipc = active ? ip : g.npairs - 1;
  const long long sic = (long long)m * g.npairs + ipc;
  const int lfa = active ? lfas[si] : INT_MAX;
  const int lfb = (SPIN == 2 && active) ? lfbs[si] : INT_MAX;
  const R x = active ? xs[ip] : (R)0;
  const int wu = __builtin_amdgcn_readfirstlane(wave);
  // The start values (mu_{lf-1}, mu_lf) of the lane's functions go to LDS
  // once (no registers held through the kernel) and are read in the chunk
  // that holds the lane's start multipole.
  __shared__ __align__(16) cplx_t<R> sSeed[W][SPIN == 0 ? 1 : 2][64];
  lds_dma16(sas + 2 * sic, &sSeed[wu][0][0]);
  if (SPIN == 2) lds_dma16(sbs + 2 * sic, &sSeed[wu][SPIN == 2 ? 1 : 0][0]);

  // B operands: weights of the ring pairs of this wave (lane j + 16k).
  // The wave copies the ring coefficients of its pairs at this m (both rings
  // of a pair side by side in the m-blocked pair-major layout) from global
  // memory into its LDS tile, which the chunks use only later, 16 bytes per
  // lane and (transform[, Q/U][, ring]) without passing through registers,
  // and builds the operands from there. Staging rows of 64 lanes x 16 bytes
  // are SR bytes apart (32 bytes of padding: the lanes of one operand read
  // fall on distinct banks). Rows of transforms beyond ntrans repeat the
  // last transform (their columns are never stored).
  constexpr int SR = 1056;
  static_assert((SPIN == 0 ? 8 : 16) * (sizeof(E) == 4 ? 1 : 2) * SR <= TW * sizeof(E),
                "staging fits the tile");
  char *Sb = reinterpret_cast<char *>(Tall + wu * TW);
  auto stage = [&](int h) {
    // The copies write LDS outside the LDS queue: the reads of the previous
    // round's rows must have completed before they are issued.
    if (h > 0) __builtin_amdgcn_s_waitcnt(kWaitLgkm0);
    // float: row (qu * 8 + bb) holds (north, south) per lane; double: row
    // ((qu * 2 + side) * 8 + bb) holds one ring per lane
#pragma unroll
    for (int bb = 0; bb < 8; ++bb) {
      const long long t = t0 + min(8 * h + bb, nt - 1);
#pragma unroll
      for (int qu = 0; qu < (SPIN == 0 ? 1 : 2); ++qu) {
        const C *gp = gp_at<E>(G, SPIN == 0 ? t : 2 * t + qu, m, ipc, g);
        if (sizeof(E) == 4) {
          lds_dma16(gp, Sb + (qu * 8 + bb) * SR);
        } else {
          lds_dma16(gp, Sb + ((qu * 2 + 0) * 8 + bb) * SR);
          lds_dma16(gp + 1, Sb + ((qu * 2 + 1) * 8 + bb) * SR);
        }
      }
    }
  };
  stage(0);
  const int lmin = block_min_waves(min(lfa, lfb), &smin);
  const R *cf = coef + NCF * ms;
  C *om = out + grp * grp_ts + t0 * out_ts + ms;
  // chunks start at multipoles with l - m even
  const int lb0 = lmin <= lmax ? lmin - ((lmin - m) & 1) : lmax + 1;
  // coefficients of chunk k into buffer buf (wave 0, lanes below kMaCH);
  // multipoles past lmax repeat lmax's (their rows are never stored)
  auto fetch_coef = [&](int k, int buf) {
    const int l = min(lb0 + k * kMaCH + lane, lmax);
    lds_dma16(cf + NCF * l, &cAB[buf][0]);
    if (SPIN == 2) lds_dma16(cf + NCF * l + 1, &cBC[SPIN == 2 ? buf : 0][0]);
  };
  // the first two chunks' coefficients, copied together with the last
  // staged weights
  auto fetch_first_coefs = [&]() {
    if (threadIdx.x < kMaCH) {
      fetch_coef(0, 0);
      fetch_coef(1, 1);
    }
  };
  for (int b = 0; b < nt; ++b)
    for (int l = m + threadIdx.x; l < lb0; l += blockDim.x) {
      om[b * out_ts + l] = {0, 0};
      if (SPIN == 2) om[b * out_ts + nelem + l] = {0, 0};
    }
  if (lb0 > lmax) {  // block-uniform; no copy into LDS may outlive the block
    __builtin_amdgcn_s_waitcnt(kWaitVm0);
    return;
  }

  const int bj = lane & 15, bk = lane >> 4;
  E Bz[32 * CT];
  {
    // The staged slots of a lane beyond the last pair (both rings) and the
    // south slots of a pair without a south ring (the ring-coefficient
    // layout leaves them unwritten) are set to 0: a product of 0 with
    // unwritten memory could be NaN.
    const bool zero_n = !active, zero_s = !active || g.pair_s[ip] < 0;
    auto finish_stage = [&]() {
      __builtin_amdgcn_s_waitcnt(kWaitVm0);
      if (zero_s) {
#pragma unroll
        for (int r = 0; r < (SPIN == 0 ? 8 : 16); ++r) {
          // r: staging row of one ring side (float: qu * 8 + bb; double: the
          // (qu, bb) rows of the north ring are r + 8 (r / 8) + {0, 8})
          if (sizeof(E) == 4) {
            C *row = reinterpret_cast<C *>(Sb + r * SR);
            row[2 * lane + 1] = C{0, 0};
            if (zero_n) row[2 * lane] = C{0, 0};
          } else {
            const int rn = r + (r >= 8 ? 8 : 0);
            reinterpret_cast<C *>(Sb + (rn + 8) * SR)[lane] = C{0, 0};
            if (zero_n) reinterpret_cast<C *>(Sb + rn * SR)[lane] = C{0, 0};
          }
        }
      }
      wave_sync();
    };
    const E *Sf = reinterpret_cast<const E *>(Sb);
    // element of (staging row, pair, ring side, component)
    auto sv = [&](int row, int qu, int pl, int side, int comp) -> E {
      if (sizeof(E) == 4) return Sf[(qu * 8 + row) * (SR / 4) + pl * 4 + side * 2 + comp];
      return Sf[((qu * 2 + side) * 8 + row) * (SR / 8) + pl * 2 + comp];
    };
    if constexpr (SPIN == 0) {
      // one column tile (8 transforms) at a time
#pragma unroll
      for (int ct = 0; ct < CT; ++ct) {
        if (ct > 0) stage(ct);
        if (ct == CT - 1) fetch_first_coefs();
        finish_stage();
        // Bz[(par * 16 + s) * CT + ct]: pair 4s + k, transform 8ct + j / 2
#pragma unroll
        for (int sq = 0; sq < 16; ++sq) {
          const int pl = 4 * sq + bk, bb = bj >> 1, comp = bj & 1;
          const E gn = sv(bb, 0, pl, 0, comp);
          const E gs = sv(bb, 0, pl, 1, comp);
          Bz[(0 * 16 + sq) * CT + ct] = gn + gs;
          Bz[(1 * 16 + sq) * CT + ct] = gn - gs;
        }
        wave_sync();
      }
    } else {
      // two column tiles (8 transforms) at a time, qu = 0: Q, 1: U
      static_assert(CT % 2 == 0, "spin 2 stages column tiles in pairs");
      // f = 0: (Zp(N), Zm(S)), f = 1: (Zp(S), Zm(N)) for c < 2 / c >= 2, with
      // Zp = q + i u = (q.x - u.y, q.y + u.x), Zm = q - i u = (q.x + u.y, q.y - u.x)
      const int f = bk & 1, c = bj & 3, comp = c & 1;
      const int side = (f == 0) == (c < 2) ? 0 : 1;
      const bool minus = (c < 2) == (comp == 0);
#pragma unroll
      for (int h = 0; h < CT / 2; ++h) {
        if (h > 0) stage(h);
        if (h == CT / 2 - 1) fetch_first_coefs();
        finish_stage();
        // Bz[s * CT + ct]: pair 2s + k / 2, function k % 2, transform
        // 4ct + j / 4 (ct = 2h, 2h + 1)
#pragma unroll
        for (int sc = 0; sc < 32; ++sc) {
          const int pl = 2 * sc + (bk >> 1);
#pragma unroll
          for (int cq = 0; cq < 2; ++cq) {
            const int bb = 4 * cq + (bj >> 2);
            const E qv = sv(bb, 0, pl, side, comp);
            const E uv = sv(bb, 1, pl, side, 1 - comp);
            Bz[sc * CT + 2 * h + cq] = minus ? qv - uv : qv + uv;
          }
        }
        wave_sync();
      }
    }
  }

  R a1 = 0, a2 = 0, b1 = 0, b2 = 0;
  auto seeds = [&](int fn, int lf, int lb, R &s0, R &s1) {
    const bool in = active && lf >= lb && lf < lb + kMaCH;
    const cplx_t<R> v = sSeed[wu][fn][lane];
    s0 = in ? v.x : (R)0;
    s1 = in ? v.y : (R)0;
  };
  // The coefficients of chunk k + 2 are copied during chunk k (after its
  // recurrence) into the buffer chunk k - 1 used; the barriers of chunk k
  // complete the copy.
  const int nch = (lmax - lb0) / kMaCH + 1;
  __syncthreads();
  for (int k = 0; k < nch; ++k) {
    const int lb = lb0 + k * kMaCH;
    const int bcur = k % 3;
    // recurrences of the lane's ring pair over the chunk -> tile
    const cplx_t<R> *kab = cAB[bcur];
    const bool inj = __any((lfa >= lb && lfa < lb + kMaCH) || (lfb >= lb && lfb < lb + kMaCH));
    if (SPIN == 0) {
      if (inj) {
        R a0s, a1s;
        seeds(0, lfa, lb, a0s, a1s);
#pragma unroll 8
        for (int i = 0; i < kMaCH; ++i)
          T[i * TS + lane] = (E)rec_step(kab[i].x * x, a1, a2, lb + i, lfa, a0s, a1s);
      } else {
#pragma unroll 8
        for (int i = 0; i < kMaCH; ++i) {
          const R lam = fma(kab[i].x * x, a1, -a2);
          a2 = a1;
          a1 = lam;
          T[i * TS + lane] = (E)lam;
        }
      }
    } else {
      C *T2 = reinterpret_cast<C *>(T);
      if (inj) {
        R a0s, a1s, b0s, b1s;
        seeds(0, lfa, lb, a0s, a1s);
        seeds(SPIN == 2 ? 1 : 0, lfb, lb, b0s, b1s);
#pragma unroll 8
        for (int i = 0; i < kMaCH; ++i) {
          const cplx_t<R> k2 = kab[i];
          const R ax = k2.x * x;
          const R p = rec_step(ax + k2.y, a1, a2, lb + i, lfa, a0s, a1s);
          const R q = rec_step(ax - k2.y, b1, b2, lb + i, lfb, b0s, b1s);
          const E qs = (i & 1) ? -(E)q : (E)q;
          T2[i * (TS / 2) + lane] = {(E)p, qs};
        }
      } else {
#pragma unroll 8
        for (int i = 0; i < kMaCH; ++i) {
          const cplx_t<R> k2 = kab[i];
          const R ax = k2.x * x;
          const R p = fma(ax + k2.y, a1, -a2);
          const R q = fma(ax - k2.y, b1, -b2);
          a2 = a1;
          a1 = p;
          b2 = b1;
          b1 = q;
          const E qs = (i & 1) ? -(E)q : (E)q;
          T2[i * (TS / 2) + lane] = {(E)p, qs};
        }
      }
    }
    wave_sync();
    // coefficients of chunk k + 2, issued after the recurrence (whose start
    // values of an injection chunk are loaded first) and stored after the
    // products
    if (threadIdx.x < kMaCH && k + 2 < nch) fetch_coef(k + 2, (k + 2) % 3);
    // products: D[par][ct] over the wave's ring pairs
    V4 D[2][CT];
#pragma unroll
    for (int ct = 0; ct < CT; ++ct) D[0][ct] = D[1][ct] = V4{0, 0, 0, 0};
    if (SPIN == 0) {
#pragma unroll
      for (int s = 0; s < 16; ++s) {
        const E ae = T[(2 * bj) * TS + 4 * s + bk];
        const E ao = T[(2 * bj + 1) * TS + 4 * s + bk];
#pragma unroll
        for (int ct = 0; ct < CT; ++ct) {
          D[0][ct] = Mfma16<E>::run(ae, Bz[(0 * 16 + s) * CT + ct], D[0][ct]);
          D[1][ct] = Mfma16<E>::run(ao, Bz[(1 * 16 + s) * CT + ct], D[1][ct]);
        }
      }
    } else {
#pragma unroll
      for (int s = 0; s < 32; ++s) {
        const E ae = T[(2 * bj) * TS + 4 * s + bk];
        const E ao = T[(2 * bj + 1) * TS + 4 * s + bk];
#pragma unroll
        for (int ct = 0; ct < CT; ++ct) {
          D[0][ct] = Mfma16<E>::run(ae, Bz[s * CT + ct], D[0][ct]);
          D[1][ct] = Mfma16<E>::run(ao, Bz[s * CT + ct], D[1][ct]);
        }
      }
    }
    // partial tiles -> LDS (over this wave's tile), row-major with a padded
    // row stride: red[(par * CT + ct) * 16 + i][j] at ((par * CT + ct) * 16 + i) * RS + j,
    // so the reduction reads whole runs of a row (one vector load per wave)
    // and the stores of the 16 columns of a row spread over the banks
    constexpr int RS = 2 * CT * 16 * 20 <= TW ? 20 : 16;
    static_assert(2 * CT * 16 * RS <= TW, "partial tiles fit the tile");
#pragma unroll
    for (int par = 0; par < 2; ++par)
#pragma unroll
      for (int ct = 0; ct < CT; ++ct) {
        E *tile = &T[(par * CT + ct) * 16 * RS + bj];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          // float32: rows 4 bk + r; float64: rows 4 r + bk
          const int i = std::is_same<E, float>::value ? 4 * bk + r : 4 * r + bk;
          tile[i * RS] = D[par][ct][r];
        }
      }
    __syncthreads();
    if (SPIN == 0) {
      // items (par, ct, i, column pair)
      for (int it = threadIdx.x; it < 256 * CT; it += 64 * W) {
        const int pc = it >> 7, i = (it >> 3) & 15, jp = it & 7;
        E re = 0, im = 0;
#pragma unroll
        for (int w = 0; w < W; ++w) {
          const C v = *reinterpret_cast<const C *>(&Tall[w * TW + (pc * 16 + i) * RS + 2 * jp]);
          re += v.x;
          im += v.y;
        }
        const int row = pc / CT + 2 * i;
        const int l = lb + row, b = 8 * (pc % CT) + jp;
        if (l <= lmax && b < nt) {
          const E c = (E)cAB[bcur][row].y;
          om[b * out_ts + l] = {c * re, c * im};
        }
      }
    } else {
      // items (par, ct, i, transform of the tile)
      for (int it = threadIdx.x; it < 128 * CT; it += 64 * W) {
        const int pc = it >> 6, i = (it >> 2) & 15, tq = it & 3;
        E v[4] = {0, 0, 0, 0};
#pragma unroll
        for (int w = 0; w < W; ++w) {
          const C *src = reinterpret_cast<const C *>(&Tall[w * TW + (pc * 16 + i) * RS + 4 * tq]);
          const C lo = src[0], hi = src[1];
          v[0] += lo.x;
          v[1] += lo.y;
          v[2] += hi.x;
          v[3] += hi.y;
        }
        const int par = pc / CT;
        const int row = par + 2 * i;
        const int l = lb + row, b = 4 * (pc % CT) + tq;
        if (l <= lmax && b < nt) {
          const E sg = par ? (E)-1 : (E)1;
          const C sp = {v[0], v[1]}, sm = {sg * v[2], sg * v[3]};
          const E c = (E)(0.5 * cBC[SPIN == 2 ? bcur : 0][row].y);
          // aE = -(S+ + S-)/2, aB = (i/2)(S+ - S-)
          om[b * out_ts + l] = {-c * (sp.x + sm.x), -c * (sp.y + sm.y)};
          om[b * out_ts + nelem + l] = {-c * (sp.y - sm.y), c * (sp.x - sm.x)};
        }
      }
    }
    __syncthreads();
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

// Matrix-core Legendre kernels for float32 batches of at least 8 transforms
// (or the plan's tuned minimum), unless OWMI_LEGENDRE_VALU=1 selects the
// vector kernels (for comparisons).
static bool use_mfma(const owmi_plan_impl *p, long long ntrans) {
  static const bool off = [] {
    const char *e = std::getenv("OWMI_LEGENDRE_VALU");
    return e != nullptr && e[0] == '1';
  }();
  const int tuned = p->tuned(OWMI_TUNE_MFMA_MIN, ntrans);
  return !off && ntrans >= (tuned > 0 ? tuned : 8);
}

// Largest of 4, 2, 1 not above ntrans and the tuned cap (0: no cap).
static int nb_for(long long ntrans, int cap) {
  const int c = cap > 0 ? cap : 4;
  if (ntrans >= 4 && c >= 4) return 4;
  if (ntrans >= 2 && c >= 2) return 2;
  return 1;
}

static void launch_synth_mfma(owmi_plan_impl *p, int spin, const float2 *alm, long long alm_ts,
                              float2 *G, long long ntrans, hipStream_t st) {
  PlanTables<R> &tb = p->t64;
  const int bs = (int)std::min<long long>(256, ((p->npairs + 63) / 64) * 64);
  dim3 grid(ceil_div(p->npairs, bs), (unsigned)(p->mmax + 1), ceil_div(ntrans, 16));
  // 256-thread blocks prefetch 4 staging items per thread (64 x 16 items)
  if (spin == 0) {
    if (bs == 256)
      hipLaunchKernelGGL((synth_mfma_kernel<0, 4>), grid, dim3(bs), 0, st, leg_args(p), tb.x,
                         tb.lf0, tb.s0, nullptr, nullptr, tb.coef0, alm, alm_ts, ntrans, G);
    else
      hipLaunchKernelGGL((synth_mfma_kernel<0, 0>), grid, dim3(bs), 0, st, leg_args(p), tb.x,
                         tb.lf0, tb.s0, nullptr, nullptr, tb.coef0, alm, alm_ts, ntrans, G);
  } else {
    if (bs == 256)
      hipLaunchKernelGGL((synth_mfma_kernel<2, 4>), grid, dim3(bs), 0, st, leg_args(p), tb.x,
                         tb.lfp, tb.sp, tb.lfm, tb.sm, tb.coef2, alm, alm_ts, ntrans, G);
    else
      hipLaunchKernelGGL((synth_mfma_kernel<2, 0>), grid, dim3(bs), 0, st, leg_args(p), tb.x,
                         tb.lfp, tb.sp, tb.lfm, tb.sm, tb.coef2, alm, alm_ts, ntrans, G);
  }
}

// float64: NCT column tiles of 8 (spin 0) or 4 (spin 2) transforms; two
// once the batch fills them.
template <int SPIN>
static void launch_synth_mfma64(owmi_plan_impl *p, const double2 *alm, long long alm_ts,
                                double2 *G, long long ntrans, hipStream_t st) {
  PlanTables<R> &tb = p->t64;
  const int bs = (int)std::min<long long>(256, ((p->npairs + 63) / 64) * 64);
  const int per = SPIN == 0 ? 8 : 4;
  const int nct = ntrans >= 2 * per ? 2 : 1;
  dim3 grid(ceil_div(p->npairs, bs), (unsigned)(p->mmax + 1), ceil_div(ntrans, per * nct));
  const int *lfa = SPIN == 0 ? tb.lf0 : tb.lfp;
  const R *sa = SPIN == 0 ? tb.s0 : tb.sp;
  const int *lfb = SPIN == 0 ? nullptr : tb.lfm;
  const R *sb = SPIN == 0 ? nullptr : tb.sm;
  const R *cf = SPIN == 0 ? tb.coef0 : tb.coef2;
  if (nct == 2)
    hipLaunchKernelGGL((synth_mfma64_kernel<SPIN, 2>), grid, dim3(bs), 0, st, leg_args(p), tb.x,
                       lfa, sa, lfb, sb, cf, alm, alm_ts, ntrans, G);
  else
    hipLaunchKernelGGL((synth_mfma64_kernel<SPIN, 1>), grid, dim3(bs), 0, st, leg_args(p), tb.x,
                       lfa, sa, lfb, sb, cf, alm, alm_ts, ntrans, G);
}

static bool use_mfma64(const owmi_plan_impl *p, long long ntrans);

// Transforms per block: the widest of 4, 2, 1 not exceeding ntrans.
template <typename T>
void launch_synthesis(owmi_plan_impl *p, int spin, const cplx_t<T> *alm, long long alm_ts,
                      cplx_t<T> *G, long long ntrans, hipStream_t st) {
  OWMI_REQUIRE(ntrans >= 1 && ntrans < 65536, "ntrans must be in [1, 65535]");
  if constexpr (std::is_same<T, float>::value) {
    if (use_mfma(p, ntrans)) return launch_synth_mfma(p, spin, alm, alm_ts, G, ntrans, st);
  } else {
    if (use_mfma64(p, ntrans)) {
      if (spin == 0) return launch_synth_mfma64<0>(p, alm, alm_ts, G, ntrans, st);
      return launch_synth_mfma64<2>(p, alm, alm_ts, G, ntrans, st);
    }
  }
  const int nb = nb_for(ntrans, p->tuned(OWMI_TUNE_SYNTH_NB, ntrans));
  if (nb == 4)
    launch_synth_nb<T, 4>(p, spin, alm, alm_ts, G, ntrans, st);
  else if (nb == 2)
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
  const int nb = nb_for(ntrans, p->tuned(OWMI_TUNE_ADJ_NB, ntrans));
  if constexpr (nbmax >= 4) {
    if (nb == 4) return launch_adj_k<T, SPIN, K, 4>(p, G, alm, alm_ts, ntrans, st);
  }
  if (nb >= 2) return launch_adj_k<T, SPIN, K, 2>(p, G, alm, alm_ts, ntrans, st);
  launch_adj_k<T, SPIN, K, 1>(p, G, alm, alm_ts, ntrans, st);
}

template <typename E, int SPIN, int W, int CT>
static void launch_adj_mfma_w(owmi_plan_impl *p, const cplx_t<E> *G, cplx_t<E> *alm,
                              long long alm_ts, long long ntrans, hipStream_t st) {
  PlanTables<R> &tb = p->t64;
  const int ngroups = (int)ceil_div(p->npairs, 64LL * W);
  const long long nelem = nelem_of(p->lmax);
  const int nc = SPIN == 0 ? 1 : 2;
  cplx_t<E> *dst = alm;
  long long dst_ts = alm_ts, grp_ts = 0;
  if (ngroups > 1) {
    // partial sums [group][transform][comp][nelem] in the plan's second workspace
    dst_ts = nc * nelem;
    grp_ts = ntrans * dst_ts;
    dst = static_cast<cplx_t<E> *>(
        p->ensure_work2((size_t)ngroups * grp_ts * sizeof(cplx_t<E>)));
  }
  const int nt = (SPIN == 0 ? 8 : 4) * CT;
  dim3 grid(ngroups, (unsigned)(p->mmax + 1), ceil_div(ntrans, nt));
  if constexpr (SPIN == 0)
    hipLaunchKernelGGL((adj_mfma_kernel<E, 0, W, CT>), grid, dim3(64 * W), 0, st, leg_args(p),
                       grp_ts, tb.x, tb.lf0, tb.s0, nullptr, nullptr, tb.coef0, G, ntrans, dst,
                       dst_ts);
  else
    hipLaunchKernelGGL((adj_mfma_kernel<E, 2, W, CT>), grid, dim3(64 * W), 0, st, leg_args(p),
                       grp_ts, tb.x, tb.lfp, tb.sp, tb.lfm, tb.sm, tb.coef2, G, ntrans, dst,
                       dst_ts);
  if (ngroups > 1) {
    dim3 g2(ceil_div(nc * nelem, 256), (unsigned)ntrans);
    hipLaunchKernelGGL(sum_groups_kernel<E>, g2, dim3(256), 0, st, dst, grp_ts, ngroups, alm,
                       alm_ts, dst_ts, nelem, (long long)nc);
  }
}

// Column tiles per block of the float32 matrix-core adjoint: 4 for spin 2
// from 16 transforms (one recurrence per 16 transforms), else 2; OWMI_ADJ_CT
// (2 or 4) overrides for measurements.
static int adj_mfma_ct(const owmi_plan_impl *p, int spin, long long ntrans) {
  static const int forced = [] {
    const char *e = std::getenv("OWMI_ADJ_CT");
    return e != nullptr ? std::atoi(e) : 0;
  }();
  if (forced == 2 || forced == 4) return forced;
  const int tuned = p->tuned(spin == 0 ? OWMI_TUNE_ADJ_CT0 : OWMI_TUNE_ADJ_CT2, ntrans);
  if (tuned == 2 || tuned == 4) return tuned;
  return (spin == 2 && ntrans >= 16) ? 4 : 2;
}

// float32: CT from adj_mfma_ct; float64: 2 column tiles (the float64
// weights of 4 tiles would take 256 registers).
template <typename E, int SPIN, int W>
static void launch_adj_mfma_ct(owmi_plan_impl *p, const cplx_t<E> *G, cplx_t<E> *alm,
                               long long alm_ts, long long ntrans, hipStream_t st) {
  if constexpr (std::is_same<E, float>::value) {
    if (adj_mfma_ct(p, SPIN, ntrans) == 4)
      return launch_adj_mfma_w<E, SPIN, W, 4>(p, G, alm, alm_ts, ntrans, st);
  }
  launch_adj_mfma_w<E, SPIN, W, 2>(p, G, alm, alm_ts, ntrans, st);
}

// Waves per block: enough for every ring pair of the plan, at most 8 (at
// most 4 for float64 spin 2, whose tiles of 4 waves take 133 KB of LDS).
template <typename E, int SPIN>
static void launch_adj_mfma(owmi_plan_impl *p, const cplx_t<E> *G, cplx_t<E> *alm,
                            long long alm_ts, long long ntrans, hipStream_t st) {
  const long long np = p->npairs;
  if (np <= 64) return launch_adj_mfma_ct<E, SPIN, 1>(p, G, alm, alm_ts, ntrans, st);
  if (np <= 128) return launch_adj_mfma_ct<E, SPIN, 2>(p, G, alm, alm_ts, ntrans, st);
  if (np <= 256 || (std::is_same<E, double>::value && SPIN == 2))
    return launch_adj_mfma_ct<E, SPIN, 4>(p, G, alm, alm_ts, ntrans, st);
  if constexpr (!(std::is_same<E, double>::value && SPIN == 2))
    launch_adj_mfma_ct<E, SPIN, 8>(p, G, alm, alm_ts, ntrans, st);
}

// float64 batches from this size on the float64 matrix cores, synthesis and
// adjoint (tuned value, else 8: at 4 transforms the vector adjoint is still
// faster at lmax 1023; OWMI_LEGENDRE_VALU=1 keeps every precision on the
// vector kernels).
static bool use_mfma64(const owmi_plan_impl *p, long long ntrans) {
  static const bool off = [] {
    const char *e = std::getenv("OWMI_LEGENDRE_VALU");
    return e != nullptr && e[0] == '1';
  }();
  const int tuned = p->tuned(OWMI_TUNE_MFMA64_MIN, ntrans);
  return !off && ntrans >= (tuned > 0 ? tuned : 8);
}

template <typename T>
void launch_adjoint(owmi_plan_impl *p, int spin, const cplx_t<T> *G, cplx_t<T> *alm,
                    long long alm_ts, long long ntrans, hipStream_t st) {
  OWMI_REQUIRE(ntrans >= 1 && ntrans < 65536, "ntrans must be in [1, 65535]");
  if constexpr (std::is_same<T, float>::value) {
    if (use_mfma(p, ntrans)) {
      if (spin == 0) return launch_adj_mfma<float, 0>(p, G, alm, alm_ts, ntrans, st);
      return launch_adj_mfma<float, 2>(p, G, alm, alm_ts, ntrans, st);
    }
  } else {
    if (use_mfma64(p, ntrans)) {
      if (spin == 0) return launch_adj_mfma<double, 0>(p, G, alm, alm_ts, ntrans, st);
      return launch_adj_mfma<double, 2>(p, G, alm, alm_ts, ntrans, st);
    }
  }
  // Ring pairs per lane: 2, or 1 on geometries of at most 64 ring pairs
  // (coarse or ring-subset plans), where two would leave most slots empty.
  // More pairs per lane amortise the cross-lane reduction further but cost
  // registers and waves (measured slower at lmax 1023); a tuned value
  // (spin 0: 1, 2, 4; spin 2: 1, 2) replaces the rule.
  const long long np = p->npairs;
  if (spin == 0) {
    int k = p->tuned(OWMI_TUNE_ADJ_K0, ntrans);
    if (k != 1 && k != 2 && k != 4) k = np <= 64 ? 1 : np <= 128 ? 2 : 4;
    if (k == 1) return launch_adj_nb<T, 0, 1>(p, G, alm, alm_ts, ntrans, st);
    if (k == 2) return launch_adj_nb<T, 0, 2>(p, G, alm, alm_ts, ntrans, st);
    return launch_adj_nb<T, 0, 4>(p, G, alm, alm_ts, ntrans, st);
  }
  int k = p->tuned(OWMI_TUNE_ADJ_K2, ntrans);
  if (k != 1 && k != 2) k = np <= 64 ? 1 : 2;
  if (k == 1) return launch_adj_nb<T, 2, 1>(p, G, alm, alm_ts, ntrans, st);
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
