// Legendre transforms between alm (triangular, m-major) and ring
// coefficients G[row][m] for spin 0 and spin 2 (HEALPix convention).
//
// The normalised functions slam_lm(x) are generated on the fly by the
// three-term recurrence in l at fixed m,
//     slam_l = (A_l x + B_l) slam_{l-1} - C_l slam_{l-2},
// seeded at l = max(m, |s|) from per-(m, ring-pair) tables. Each thread owns
// one mirror pair of rings (theta, pi - theta); the south ring follows from
// slam_lm(-x) = (-1)^(l+m) slam_lm(x) (spin 0) and
// 2lam_lm(pi - theta) = (-1)^(l+m) (-2)lam_lm(theta) (spin 2), so every
// recurrence step serves both rings through even/odd (l - m) accumulators.
//
// Tiny seeds are carried as v * 2^(S k) (S = 128 for double, 32 for float):
// while k < 0 the recurrence rescales by 2^-S whenever |v| exceeds 2^S;
// values enter the sums with weight 1 (k = 0), 2^-S (k = -1) or 0
// (k <= -2, below 2^-2S).
//
// Spin-2 synthesis per m (P = -(aE + i aB)/2, Mm = -(aE - i aB)/2):
//     Q_m = sum_l (P 2lam + Mm (-2)lam),  U_m = -i sum_l (P 2lam - Mm (-2)lam)
// Spin-2 adjoint per m (Zp = q + i u, Zm = q - i u from ring coefficients):
//     S+ = sum_r 2lam Zp,  S- = sum_r (-2)lam Zm,
//     aE = -(S+ + S-)/2,  aB = (i/2)(S+ - S-).
#include "sht_internal.hpp"

namespace owmi {

namespace {

template <typename T>
struct Scaled {
  static constexpr int S = ScaleStep<T>::bits;
  __device__ static T big() { return (T)ldexp(1.0, S); }
  __device__ static T small() { return (T)ldexp(1.0, -S); }
  __device__ static T factor(int k) { return k == 0 ? (T)1 : (k == -1 ? small() : (T)0); }
};

struct LegArgs {
  const int *pair_n;
  const int *pair_s;
  long long npairs;
  long long nrings;
  int lmax;
  int mmax;
};

// ------------------------------------------------------------- synthesis

template <typename T>
__global__ void __launch_bounds__(256)
synth0_kernel(LegArgs g, const T *__restrict__ xs, const T *__restrict__ seed,
              const int *__restrict__ seedk, const T *__restrict__ coef,
              const cplx_t<T> *__restrict__ alm, long long alm_ts, cplx_t<T> *G) {
  const long long ip = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const int m = blockIdx.y;
  const long long t = blockIdx.z;
  if (ip >= g.npairs) return;
  const int lmax = g.lmax;
  const long long ms = mstart_of(m, lmax);
  const cplx_t<T> *a = alm + t * alm_ts + ms;
  const T *cf = coef + 2 * ms;
  const T x = xs[ip];
  T l1 = seed[(long long)m * g.npairs + ip];
  int k = seedk[(long long)m * g.npairs + ip];
  T l2 = 0;
  T fac = Scaled<T>::factor(k);
  const T big = Scaled<T>::big(), small = Scaled<T>::small();

  cplx_t<T> E, O = {0, 0};
  {
    cplx_t<T> av = a[m];
    T v = l1 * fac;
    E = {v * av.x, v * av.y};
  }
  for (int l = m + 1; l <= lmax; ++l) {
    const T A = cf[2 * l], C = cf[2 * l + 1];
    T lam = (A * x) * l1 - C * l2;
    l2 = l1;
    l1 = lam;
    if (k < 0 && fabs(l1) > big) {
      l1 *= small;
      l2 *= small;
      ++k;
      fac = Scaled<T>::factor(k);
    }
    const T v = l1 * fac;
    const cplx_t<T> av = a[l];
    if ((l - m) & 1) {
      O.x += v * av.x;
      O.y += v * av.y;
    } else {
      E.x += v * av.x;
      E.y += v * av.y;
    }
  }
  const int rn = g.pair_n[ip], rs = g.pair_s[ip];
  const long long stride = g.mmax + 1;
  G[(t * g.nrings + rn) * stride + m] = {E.x + O.x, E.y + O.y};
  if (rs >= 0) G[(t * g.nrings + rs) * stride + m] = {E.x - O.x, E.y - O.y};
}

template <typename T>
__global__ void __launch_bounds__(256)
synth2_kernel(LegArgs g, const T *__restrict__ xs, const T *__restrict__ seedp,
              const int *__restrict__ seedpk, const T *__restrict__ seedm,
              const int *__restrict__ seedmk, const T *__restrict__ coef,
              const cplx_t<T> *__restrict__ alm, long long alm_ts, cplx_t<T> *G) {
  const long long ip = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const int m = blockIdx.y;
  const long long t = blockIdx.z;
  if (ip >= g.npairs) return;
  const int lmax = g.lmax;
  const long long nelem = nelem_of(lmax);
  const long long ms = mstart_of(m, lmax);
  const cplx_t<T> *aE = alm + t * alm_ts + ms;
  const cplx_t<T> *aB = aE + nelem;
  const T *cf = coef + 3 * ms;
  const int lstart = m > 2 ? m : 2;
  const int rn = g.pair_n[ip], rs = g.pair_s[ip];
  const long long stride = g.mmax + 1;
  cplx_t<T> *Gq = G + (t * 2 * g.nrings) * stride;
  cplx_t<T> *Gu = Gq + g.nrings * stride;
  if (lstart > lmax) {
    Gq[rn * stride + m] = {0, 0};
    Gu[rn * stride + m] = {0, 0};
    if (rs >= 0) {
      Gq[rs * stride + m] = {0, 0};
      Gu[rs * stride + m] = {0, 0};
    }
    return;
  }
  const T x = xs[ip];
  const long long si = (long long)m * g.npairs + ip;
  T p1 = seedp[si], p2 = 0, q1 = seedm[si], q2 = 0;
  int kp = seedpk[si], kq = seedmk[si];
  T fp = Scaled<T>::factor(kp), fq = Scaled<T>::factor(kq);
  const T big = Scaled<T>::big(), small = Scaled<T>::small();

  // accumulators by parity of (l - m): P*2lam, P*(-2)lam, Mm*2lam, Mm*(-2)lam
  cplx_t<T> PPe = {0, 0}, PMe = {0, 0}, MPe = {0, 0}, MMe = {0, 0};
  cplx_t<T> PPo = {0, 0}, PMo = {0, 0}, MPo = {0, 0}, MMo = {0, 0};
  for (int l = lstart; l <= lmax; ++l) {
    if (l > lstart) {
      const T A = cf[3 * l], B = cf[3 * l + 1], C = cf[3 * l + 2];
      const T ax = A * x;
      T np = (ax + B) * p1 - C * p2;
      T nq = (ax - B) * q1 - C * q2;
      p2 = p1;
      p1 = np;
      q2 = q1;
      q1 = nq;
      if (kp < 0 && fabs(p1) > big) {
        p1 *= small;
        p2 *= small;
        ++kp;
        fp = Scaled<T>::factor(kp);
      }
      if (kq < 0 && fabs(q1) > big) {
        q1 *= small;
        q2 *= small;
        ++kq;
        fq = Scaled<T>::factor(kq);
      }
    }
    const T vp = p1 * fp, vq = q1 * fq;
    const cplx_t<T> e = aE[l], b = aB[l];
    // P = -(e + i b)/2, Mm = -(e - i b)/2
    const cplx_t<T> P = {(T)-0.5 * (e.x - b.y), (T)-0.5 * (e.y + b.x)};
    const cplx_t<T> Mm = {(T)-0.5 * (e.x + b.y), (T)-0.5 * (e.y - b.x)};
    if ((l - m) & 1) {
      PPo.x += P.x * vp; PPo.y += P.y * vp;
      PMo.x += P.x * vq; PMo.y += P.y * vq;
      MPo.x += Mm.x * vp; MPo.y += Mm.y * vp;
      MMo.x += Mm.x * vq; MMo.y += Mm.y * vq;
    } else {
      PPe.x += P.x * vp; PPe.y += P.y * vp;
      PMe.x += P.x * vq; PMe.y += P.y * vq;
      MPe.x += Mm.x * vp; MPe.y += Mm.y * vp;
      MMe.x += Mm.x * vq; MMe.y += Mm.y * vq;
    }
  }
  // North: A+ = sum P 2lam, A- = sum Mm (-2)lam.
  // South: A+ = sum (-1)^(l+m) P (-2)lam, A- = sum (-1)^(l+m) Mm 2lam.
  cplx_t<T> ApN = {PPe.x + PPo.x, PPe.y + PPo.y};
  cplx_t<T> AmN = {MMe.x + MMo.x, MMe.y + MMo.y};
  // Q = A+ + A-, U = -i (A+ - A-)
  Gq[rn * stride + m] = {ApN.x + AmN.x, ApN.y + AmN.y};
  Gu[rn * stride + m] = {ApN.y - AmN.y, -(ApN.x - AmN.x)};
  if (rs >= 0) {
    cplx_t<T> ApS = {PMe.x - PMo.x, PMe.y - PMo.y};
    cplx_t<T> AmS = {MPe.x - MPo.x, MPe.y - MPo.y};
    Gq[rs * stride + m] = {ApS.x + AmS.x, ApS.y + AmS.y};
    Gu[rs * stride + m] = {ApS.y - AmS.y, -(ApS.x - AmS.x)};
  }
}

// ---------------------------------------------------------------- adjoint

// Sum v[0..CH) over the 64 lanes of a wave. On return the lanes with
// (lane % (64/CH)) == 0 hold the wave total of index lane / (64/CH).
template <typename T, int CH>
__device__ __forceinline__ cplx_t<T> wave_transpose_reduce(cplx_t<T> (&v)[CH], int lane) {
  static_assert(CH == 8 || CH == 16, "chunk must be 8 or 16");
  // Halving steps: offsets 32, 16, ... keep half of the values each.
  cplx_t<T> cur[CH];
#pragma unroll
  for (int i = 0; i < CH; ++i) cur[i] = v[i];
  int n = CH;
#pragma unroll
  for (int off = 32; n > 1; off >>= 1) {
    const bool hi = (lane & off) != 0;
    const int h = n >> 1;
#pragma unroll
    for (int j = 0; j < CH / 2; ++j) {
      if (j < h) {
        cplx_t<T> send = hi ? cur[j] : cur[j + h];
        cplx_t<T> keep = hi ? cur[j + h] : cur[j];
        T rx = __shfl_xor(send.x, off);
        T ry = __shfl_xor(send.y, off);
        cur[j] = {keep.x + rx, keep.y + ry};
      }
    }
    n = h;
  }
  // Remaining lane bits below 64/CH: plain butterfly sums.
  cplx_t<T> r = cur[0];
#pragma unroll
  for (int off = (64 / CH) >> 1; off >= 1; off >>= 1) {
    r.x += __shfl_xor(r.x, off);
    r.y += __shfl_xor(r.y, off);
  }
  return r;
}

// Index carried by lane after wave_transpose_reduce.
template <int CH>
__device__ __forceinline__ int reduced_index(int lane) {
  // The halving steps fix index bits from the top: offset 32 picks the top
  // bit (h = CH/2), 16 the next, ...
  int idx = 0;
  int h = CH >> 1;
  for (int off = 32; h >= 1; off >>= 1, h >>= 1)
    if (lane & off) idx += h;
  return idx;
}

constexpr int kAdjThreads = 256;

template <typename T>
__global__ void __launch_bounds__(kAdjThreads)
adj0_kernel(LegArgs g, const T *__restrict__ xs, const T *__restrict__ seed,
            const int *__restrict__ seedk, const T *__restrict__ coef,
            const cplx_t<T> *__restrict__ G, cplx_t<T> *alm, long long alm_ts) {
  constexpr int CH = 16;
  __shared__ cplx_t<T> red[kAdjThreads / 64][CH];
  const int m = blockIdx.x;
  const long long t = blockIdx.y;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int nw = blockDim.x >> 6;
  const int lmax = g.lmax;
  const long long ms = mstart_of(m, lmax);
  const T *cf = coef + 2 * ms;
  cplx_t<T> *out = alm + t * alm_ts + ms;
  const long long stride = g.mmax + 1;
  const T big = Scaled<T>::big(), small = Scaled<T>::small();
  const int my_idx = reduced_index<CH>(lane);
  const bool holder = (lane & (64 / CH - 1)) == 0;

  for (long long base = 0; base < g.npairs; base += blockDim.x) {
    const long long ip = base + threadIdx.x;
    const bool active = ip < g.npairs;
    T x = 0, l1 = 0, l2 = 0;
    int k = -1000;
    cplx_t<T> He = {0, 0}, Ho = {0, 0};
    if (active) {
      x = xs[ip];
      l1 = seed[(long long)m * g.npairs + ip];
      k = seedk[(long long)m * g.npairs + ip];
      const int rn = g.pair_n[ip], rs = g.pair_s[ip];
      cplx_t<T> gn = G[(t * g.nrings + rn) * stride + m];
      if (rs >= 0) {
        cplx_t<T> gs = G[(t * g.nrings + rs) * stride + m];
        He = {gn.x + gs.x, gn.y + gs.y};
        Ho = {gn.x - gs.x, gn.y - gs.y};
      } else {
        He = gn;
        Ho = gn;
      }
    }
    T fac = Scaled<T>::factor(k);
    for (int l0 = m; l0 <= lmax; l0 += CH) {
      cplx_t<T> v[CH];
#pragma unroll
      for (int i = 0; i < CH; ++i) {
        const int l = l0 + i;
        T val = 0;
        if (l <= lmax) {
          if (l > m) {
            const T A = cf[2 * l], C = cf[2 * l + 1];
            T lam = (A * x) * l1 - C * l2;
            l2 = l1;
            l1 = lam;
            if (k < 0 && fabs(l1) > big) {
              l1 *= small;
              l2 *= small;
              ++k;
              fac = Scaled<T>::factor(k);
            }
          }
          val = l1 * fac;
        }
        const cplx_t<T> H = ((l - m) & 1) ? Ho : He;
        v[i] = {val * H.x, val * H.y};
      }
      cplx_t<T> r = wave_transpose_reduce<T, CH>(v, lane);
      if (holder) red[wave][my_idx] = r;
      __syncthreads();
      if (threadIdx.x < CH) {
        const int l = l0 + threadIdx.x;
        if (l <= lmax) {
          cplx_t<T> s = red[0][threadIdx.x];
          for (int w = 1; w < nw; ++w) {
            s.x += red[w][threadIdx.x].x;
            s.y += red[w][threadIdx.x].y;
          }
          if (base == 0) {
            out[l] = s;
          } else {
            cplx_t<T> o = out[l];
            out[l] = {o.x + s.x, o.y + s.y};
          }
        }
      }
      __syncthreads();
    }
  }
}

template <typename T>
__global__ void __launch_bounds__(kAdjThreads)
adj2_kernel(LegArgs g, const T *__restrict__ xs, const T *__restrict__ seedp,
            const int *__restrict__ seedpk, const T *__restrict__ seedm,
            const int *__restrict__ seedmk, const T *__restrict__ coef,
            const cplx_t<T> *__restrict__ G, cplx_t<T> *alm, long long alm_ts) {
  constexpr int CH = 8;
  __shared__ cplx_t<T> red[2][kAdjThreads / 64][CH];
  const int m = blockIdx.x;
  const long long t = blockIdx.y;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int nw = blockDim.x >> 6;
  const int lmax = g.lmax;
  const long long nelem = nelem_of(lmax);
  const long long ms = mstart_of(m, lmax);
  const T *cf = coef + 3 * ms;
  cplx_t<T> *outE = alm + t * alm_ts + ms;
  cplx_t<T> *outB = outE + nelem;
  const long long stride = g.mmax + 1;
  const cplx_t<T> *Gq = G + (t * 2 * g.nrings) * stride;
  const cplx_t<T> *Gu = Gq + g.nrings * stride;
  const T big = Scaled<T>::big(), small = Scaled<T>::small();
  const int lstart = m > 2 ? m : 2;
  const int my_idx = reduced_index<CH>(lane);
  const bool holder = (lane & (64 / CH - 1)) == 0;

  // l < 2 rows of spin-2 alm are zero.
  if (threadIdx.x == 0)
    for (int l = m; l < lstart && l <= lmax; ++l) {
      outE[l] = {0, 0};
      outB[l] = {0, 0};
    }
  if (lstart > lmax) return;

  for (long long base = 0; base < g.npairs; base += blockDim.x) {
    const long long ip = base + threadIdx.x;
    const bool active = ip < g.npairs;
    T x = 0, p1 = 0, p2 = 0, q1 = 0, q2 = 0;
    int kp = -1000, kq = -1000;
    // Zp = q + i u, Zm = q - i u for north (N) and south (S) rings.
    cplx_t<T> ZpN = {0, 0}, ZmN = {0, 0}, ZpS = {0, 0}, ZmS = {0, 0};
    if (active) {
      x = xs[ip];
      const long long si = (long long)m * g.npairs + ip;
      p1 = seedp[si];
      kp = seedpk[si];
      q1 = seedm[si];
      kq = seedmk[si];
      const int rn = g.pair_n[ip], rs = g.pair_s[ip];
      cplx_t<T> qn = Gq[rn * stride + m], un = Gu[rn * stride + m];
      ZpN = {qn.x - un.y, qn.y + un.x};
      ZmN = {qn.x + un.y, qn.y - un.x};
      if (rs >= 0) {
        cplx_t<T> qs = Gq[rs * stride + m], us = Gu[rs * stride + m];
        ZpS = {qs.x - us.y, qs.y + us.x};
        ZmS = {qs.x + us.y, qs.y - us.x};
      }
    }
    T fp = Scaled<T>::factor(kp), fq = Scaled<T>::factor(kq);
    for (int l0 = lstart; l0 <= lmax; l0 += CH) {
      cplx_t<T> vP[CH], vM[CH];
#pragma unroll
      for (int i = 0; i < CH; ++i) {
        const int l = l0 + i;
        T a = 0, b = 0;
        if (l <= lmax) {
          if (l > lstart) {
            const T A = cf[3 * l], B = cf[3 * l + 1], C = cf[3 * l + 2];
            const T ax = A * x;
            T np = (ax + B) * p1 - C * p2;
            T nq = (ax - B) * q1 - C * q2;
            p2 = p1;
            p1 = np;
            q2 = q1;
            q1 = nq;
            if (kp < 0 && fabs(p1) > big) {
              p1 *= small;
              p2 *= small;
              ++kp;
              fp = Scaled<T>::factor(kp);
            }
            if (kq < 0 && fabs(q1) > big) {
              q1 *= small;
              q2 *= small;
              ++kq;
              fq = Scaled<T>::factor(kq);
            }
          }
          a = p1 * fp;
          b = q1 * fq;
        }
        // even: S+ = 2lam ZpN + (-2)lam ZpS, S- = (-2)lam ZmN + 2lam ZmS
        // odd : S+ = 2lam ZpN - (-2)lam ZpS, S- = (-2)lam ZmN - 2lam ZmS
        const T sg = ((l - m) & 1) ? (T)-1 : (T)1;
        const T bs = b * sg, as = a * sg;
        vP[i] = {a * ZpN.x + bs * ZpS.x, a * ZpN.y + bs * ZpS.y};
        vM[i] = {b * ZmN.x + as * ZmS.x, b * ZmN.y + as * ZmS.y};
      }
      cplx_t<T> rP = wave_transpose_reduce<T, CH>(vP, lane);
      cplx_t<T> rM = wave_transpose_reduce<T, CH>(vM, lane);
      if (holder) {
        red[0][wave][my_idx] = rP;
        red[1][wave][my_idx] = rM;
      }
      __syncthreads();
      if (threadIdx.x < CH) {
        const int l = l0 + threadIdx.x;
        if (l <= lmax) {
          cplx_t<T> sp = red[0][0][threadIdx.x], sm = red[1][0][threadIdx.x];
          for (int w = 1; w < nw; ++w) {
            sp.x += red[0][w][threadIdx.x].x;
            sp.y += red[0][w][threadIdx.x].y;
            sm.x += red[1][w][threadIdx.x].x;
            sm.y += red[1][w][threadIdx.x].y;
          }
          // aE = -(S+ + S-)/2, aB = (i/2)(S+ - S-)
          cplx_t<T> e = {(T)-0.5 * (sp.x + sm.x), (T)-0.5 * (sp.y + sm.y)};
          cplx_t<T> bb = {(T)-0.5 * (sp.y - sm.y), (T)0.5 * (sp.x - sm.x)};
          if (base == 0) {
            outE[l] = e;
            outB[l] = bb;
          } else {
            cplx_t<T> oe = outE[l], ob = outB[l];
            outE[l] = {oe.x + e.x, oe.y + e.y};
            outB[l] = {ob.x + bb.x, ob.y + bb.y};
          }
        }
      }
      __syncthreads();
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

}  // namespace

template <typename T>
void launch_synthesis(owmi_plan_impl *p, int spin, const cplx_t<T> *alm, long long alm_ts,
                      cplx_t<T> *G, long long ntrans, hipStream_t st) {
  OWMI_REQUIRE(ntrans < 65536, "too many transforms for one launch");
  PlanTables<T> &tb = tab<T>(p);
  int bs = (int)std::min<long long>(256, ((p->npairs + 63) / 64) * 64);
  dim3 grid(ceil_div(p->npairs, bs), (unsigned)(p->mmax + 1), (unsigned)ntrans);
  if (spin == 0)
    hipLaunchKernelGGL(synth0_kernel<T>, grid, dim3(bs), 0, st, leg_args(p), tb.x, tb.seed0,
                       tb.seed0k, tb.coef0, alm, alm_ts, G);
  else
    hipLaunchKernelGGL(synth2_kernel<T>, grid, dim3(bs), 0, st, leg_args(p), tb.x, tb.seedp,
                       tb.seedpk, tb.seedm, tb.seedmk, tb.coef2, alm, alm_ts, G);
}

template <typename T>
void launch_adjoint(owmi_plan_impl *p, int spin, const cplx_t<T> *G, cplx_t<T> *alm,
                    long long alm_ts, long long ntrans, hipStream_t st) {
  OWMI_REQUIRE(ntrans < 65536, "too many transforms for one launch");
  PlanTables<T> &tb = tab<T>(p);
  int bs = (int)std::min<long long>(kAdjThreads, ((p->npairs + 63) / 64) * 64);
  dim3 grid((unsigned)(p->mmax + 1), (unsigned)ntrans);
  if (spin == 0)
    hipLaunchKernelGGL(adj0_kernel<T>, grid, dim3(bs), 0, st, leg_args(p), tb.x, tb.seed0,
                       tb.seed0k, tb.coef0, G, alm, alm_ts);
  else
    hipLaunchKernelGGL(adj2_kernel<T>, grid, dim3(bs), 0, st, leg_args(p), tb.x, tb.seedp,
                       tb.seedpk, tb.seedm, tb.seedmk, tb.coef2, G, alm, alm_ts);
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
