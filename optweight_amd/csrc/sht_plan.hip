// SHT plan: geometry analysis, Legendre seed/coefficient tables and
// Bluestein ring-FFT tables, built on the host in long double and uploaded
// once; plus the C-ABI transform entry points that chain the Legendre and
// ring-FFT kernels.
#include "sht_internal.hpp"

#include <algorithm>
#include <cmath>
#include <map>
#include <numeric>

namespace owmi {

namespace {

using LD = long double;
const LD kPi = 3.141592653589793238462643383279502884L;

// A long double with a separate power-of-two exponent: value = mant * 2^exp,
// 0.5 <= |mant| < 1 (or mant == 0). Used for the Legendre seeds, which can
// lie far below the long double range at high lmax.
struct XLD {
  LD mant = 0;
  long exp = 0;
  static XLD of(LD v) {
    XLD r;
    int e = 0;
    r.mant = frexpl(v, &e);
    r.exp = e;
    return r;
  }
  XLD mul(LD f) const {
    XLD r;
    int e = 0;
    r.mant = frexpl(mant * f, &e);
    r.exp = exp + e;
    if (r.mant == 0) r.exp = 0;
    return r;
  }
  // log2 |value| (-inf for zero)
  LD log2abs() const { return mant == 0 ? -INFINITY : log2l(fabsl(mant)) + (LD)exp; }
};

// Split a seed into (v, k) with value = v * 2^(S k): k = 0 when
// |value| >= 2^-S, else k = floor(log2|value| / S) and v in [1, 2^S).
template <typename T>
void split_seed(const XLD &s, T &v, int &k) {
  const int S = ScaleStep<T>::bits;
  if (s.mant == 0) {
    v = 0;
    k = -1000;
    return;
  }
  LD l2 = s.log2abs();
  if (l2 >= -(LD)S) {
    k = 0;
    v = (T)ldexpl(s.mant, (int)s.exp);
    return;
  }
  long kk = (long)floorl(l2 / S);
  if (kk < -1000) kk = -1000;
  k = (int)kk;
  v = (T)ldexpl(s.mant, (int)(s.exp - (long)S * kk));
}

template <typename T>
T *upload(const std::vector<T> &h) {
  T *d = nullptr;
  if (h.empty()) return nullptr;
  OWMI_HIP(hipMalloc(&d, h.size() * sizeof(T)));
  OWMI_HIP(hipMemcpy(d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
  return d;
}

template <typename T>
void free_tables(PlanTables<T> &t) {
  T **ptrs[] = {&t.x, &t.s0, &t.sp, &t.sm, &t.coef0, &t.coef2, &t.twiddle,
                &t.chirp, &t.bhat, &t.ringw, &t.phi0};
  for (T **p : ptrs)
    if (*p) {
      (void)hipFree(*p);
      *p = nullptr;
    }
  int **iptrs[] = {&t.lf0, &t.lfp, &t.lfm};
  for (int **p : iptrs)
    if (*p) {
      (void)hipFree(*p);
      *p = nullptr;
    }
  t.ready = false;
}

int ilog2_ceil(long long n) {
  int k = 0;
  while ((1LL << k) < n) ++k;
  return k;
}

}  // namespace

void *owmi_plan_impl::ensure_work(size_t bytes) {
  if (bytes <= work_bytes) return work;
  if (work) retired.push_back(work);
  work = nullptr;
  work_bytes = 0;
  OWMI_HIP(hipMalloc(&work, bytes));
  OWMI_HIP(hipMemset(work, 0, bytes));
  work_bytes = bytes;
  return work;
}

void *owmi_plan_impl::ensure_work2(size_t bytes) {
  if (bytes <= work2_bytes) return work2;
  if (work2) retired.push_back(work2);
  work2 = nullptr;
  work2_bytes = 0;
  OWMI_HIP(hipMalloc(&work2, bytes));
  work2_bytes = bytes;
  return work2;
}

void *owmi_plan_impl::ensure_work4(size_t bytes) {
  if (bytes <= work4_bytes) return work4;
  if (work4) retired.push_back(work4);
  work4 = nullptr;
  work4_bytes = 0;
  OWMI_HIP(hipMalloc(&work4, bytes));
  work4_bytes = bytes;
  return work4;
}

void *owmi_plan_impl::ensure_work3(size_t bytes) {
  if (bytes <= work3_bytes) return work3;
  if (work3) retired.push_back(work3);
  work3 = nullptr;
  work3_bytes = 0;
  OWMI_HIP(hipMalloc(&work3, bytes));
  work3_bytes = bytes;
  return work3;
}

// LDS budget of the device Bluestein-spectrum kernel (complex double).
constexpr size_t kInitLds = 160 * 1024;

// Bluestein kernel spectrum FFT_M(b) / M on the host in long double, for
// lengths whose FFT does not fit LDS: b_n = conj(w_n) for n < N,
// conj(w_{M-n}) for n > M - N, else 0, w_n = exp(-i pi n^2 / N); returns
// interleaved (re, im) doubles.
std::vector<double> host_bhat(long long N, int logm) {
  const long long M = 1LL << logm;
  std::vector<LD> re(M, 0), im(M, 0);
  auto wconj = [&](long long n, LD *r, LD *i) {
    const long long q = (n * n) % (2 * N);
    const LD ang = kPi * (LD)q / (LD)N;  // conj of exp(-i pi q / N)
    *r = cosl(ang);
    *i = sinl(ang);
  };
  for (long long n = 0; n < M; ++n) {
    if (n < N)
      wconj(n, &re[n], &im[n]);
    else if (n > M - N)
      wconj(M - n, &re[n], &im[n]);
  }
  // iterative radix-2 FFT, exp(-2 pi i jk / M), bit-reversed input order
  for (long long i = 1, j = 0; i < M; ++i) {
    long long bit = M >> 1;
    for (; j & bit; bit >>= 1) j ^= bit;
    j ^= bit;
    if (i < j) {
      std::swap(re[i], re[j]);
      std::swap(im[i], im[j]);
    }
  }
  for (long long len = 2; len <= M; len <<= 1) {
    const LD ang = -2 * kPi / (LD)len;
    for (long long i = 0; i < M; i += len) {
      for (long long k = 0; k < len / 2; ++k) {
        const LD wr = cosl(ang * k), wi = sinl(ang * k);
        const long long a = i + k, b = a + len / 2;
        const LD xr = re[b] * wr - im[b] * wi, xi = re[b] * wi + im[b] * wr;
        re[b] = re[a] - xr;
        im[b] = im[a] - xi;
        re[a] += xr;
        im[a] += xi;
      }
    }
  }
  std::vector<double> out(2 * M);
  for (long long k = 0; k < M; ++k) {
    out[2 * k] = (double)(re[k] / (LD)M);
    out[2 * k + 1] = (double)(im[k] / (LD)M);
  }
  return out;
}

// Host-side tables shared by both precisions (long double), built once.
struct HostTables {
  // seeds at l = max(m, |s|) as v * 2^(128 k) (float64 v)
  std::vector<double> v0, vp, vm;
  std::vector<int> k0, kp, km;
  // unnormalised recurrence coefficients (float64) for the seed search
  std::vector<double> U0, U2;
  // normalisation c_l and normalised coefficients per alm index
  std::vector<LD> c0, A0, c2, A2, B2;
  // search results: l_f and slam_{l_f - 1}, slam_{l_f} per (m, pair)
  std::vector<int> lf0, lfp, lfm;
  std::vector<double> s0, sp, sm;
};

static void build_host_tables(owmi_plan_impl &p, const std::vector<LD> &th_n,
                              const std::vector<LD> &xs, HostTables &h) {
  const int lmax = p.lmax, mmax = p.mmax;
  const long long np = p.npairs;
  const long long nelem = nelem_of(lmax);
  const size_t nseed = (size_t)(mmax + 1) * np;
  h.v0.resize(nseed); h.vp.resize(nseed); h.vm.resize(nseed);
  h.k0.resize(nseed); h.kp.resize(nseed); h.km.resize(nseed);
  const LD inv_sqrt_4pi = 1.0L / sqrtl(4.0L * kPi);
  const LD sqrt5_4pi = sqrtl(5.0L / (4.0L * kPi));
  for (long long i = 0; i < np; ++i) {
    LD th = th_n[i];
    LD st = sinl(th);
    LD ch = cosl(th / 2), sh = sinl(th / 2);
    // spin 0: slam_mm = -sqrt((2m+1)/(2m)) sin(theta) slam_{m-1,m-1}
    XLD a = XLD::of(inv_sqrt_4pi);
    for (int m = 0; m <= mmax; ++m) {
      if (m > 0) a = a.mul(-sqrtl((LD)(2 * m + 1) / (LD)(2 * m)) * st);
      split_seed<double>(a, h.v0[(size_t)m * np + i], h.k0[(size_t)m * np + i]);
    }
    // spin -2 at l = m >= 2: slam_22 = sqrt(5/4pi) cos^4(theta/2),
    // slam_mm = -sqrt((2m+1) 2m / ((m+2)(m-2))) (sin(theta)/2) slam_{m-1,m-1};
    // spin +2 = spin -2 * tan^4(theta/2). For m < 2 the l = 2 values of
    // (-1)^s sqrt(5/4pi) d^2_{m,-s}(theta).
    LD t4 = powl(sh / ch, 4);
    for (int m = 0; m <= mmax; ++m) {
      XLD vp, vm;
      if (m == 0) {
        LD v = sqrt5_4pi * sqrtl(6.0L) / 4.0L * st * st;
        vp = XLD::of(v);
        vm = XLD::of(v);
      } else if (m == 1) {
        vp = XLD::of(-sqrt5_4pi * sh * sh * st);
        vm = XLD::of(sqrt5_4pi * ch * ch * st);
      } else {
        if (m == 2) {
          a = XLD::of(sqrt5_4pi * powl(ch, 4));
        } else {
          a = a.mul(-sqrtl((LD)(2 * m + 1) * (LD)(2 * m) / ((LD)(m + 2) * (LD)(m - 2))) * st /
                    2.0L);
        }
        vm = a;
        vp = a.mul(t4);
      }
      split_seed<double>(vp, h.vp[(size_t)m * np + i], h.kp[(size_t)m * np + i]);
      split_seed<double>(vm, h.vm[(size_t)m * np + i], h.km[(size_t)m * np + i]);
    }
  }

  // slam_l = (A x + B) slam_{l-1} - C slam_{l-2}, (A, B, C) = (alpha, beta,
  // gamma)_{l-1} of the d^l_{m,m'} recurrence (m' = 0 or -2); normalised by
  // c_l = C c_{l-2}, c = 1 at the first two multipoles of each function.
  h.U0.assign(2 * (size_t)nelem, 0);
  h.U2.assign(3 * (size_t)nelem, 0);
  h.c0.assign((size_t)nelem, 1);
  h.A0.assign((size_t)nelem, 0);
  h.c2.assign((size_t)nelem, 1);
  h.A2.assign((size_t)nelem, 0);
  h.B2.assign((size_t)nelem, 0);
  for (int m = 0; m <= mmax; ++m) {
    long long ms = mstart_of(m, lmax);
    LD m2 = (LD)m * m;
    for (int s = 0; s <= 2; s += 2) {
      int lstart = std::max(m, s);
      LD mp2 = (LD)s * s;
      std::vector<LD> &cc = s == 0 ? h.c0 : h.c2;
      for (int l = lstart + 1; l <= lmax; ++l) {
        LD L = l - 1;
        LD r_L = sqrtl((L * L - m2) * (L * L - mp2));
        LD r_L1 = sqrtl(((L + 1) * (L + 1) - m2) * ((L + 1) * (L + 1) - mp2));
        LD alpha = sqrtl((2 * L + 3) * (2 * L + 1)) * (L + 1) / r_L1;
        LD gamma = (l - 1 == lstart || l - 1 == 0)
                       ? 0.0L
                       : sqrtl((2 * L + 3) / (2 * L - 1)) * (L + 1) * r_L / (L * r_L1);
        // s = +2 -> m' = -2: beta = +2 m sqrt((2L+3)(2L+1)) / (L r_{L+1})
        LD beta = (s == 0 || m == 0) ? 0.0L
                                     : 2.0L * m * sqrtl((2 * L + 3) * (2 * L + 1)) / (L * r_L1);
        size_t idx = (size_t)(ms + l);
        LD c = (l == lstart + 1) ? 1.0L : gamma * cc[idx - 2];
        cc[idx] = c;
        LD ratio = cc[idx - 1] / c;
        if (s == 0) {
          h.U0[2 * idx] = (double)alpha;
          h.U0[2 * idx + 1] = (double)gamma;
          h.A0[idx] = alpha * ratio;
        } else {
          h.U2[3 * idx] = (double)alpha;
          h.U2[3 * idx + 1] = (double)beta;
          h.U2[3 * idx + 2] = (double)gamma;
          h.A2[idx] = alpha * ratio;
          h.B2[idx] = beta * ratio;
        }
      }
    }
  }

  // First significant multipole and the seeds there (float64 device search).
  double *d_v0 = upload(h.v0), *d_vp = upload(h.vp), *d_vm = upload(h.vm);
  int *d_k0 = upload(h.k0), *d_kp = upload(h.kp), *d_km = upload(h.km);
  double *d_U0 = upload(h.U0), *d_U2 = upload(h.U2);
  std::vector<double> xd(np);
  for (long long i = 0; i < np; ++i) xd[i] = (double)xs[i];
  double *d_x = upload(xd);
  int *d_lf0 = nullptr, *d_lfp = nullptr, *d_lfm = nullptr;
  double *d_s0 = nullptr, *d_sp = nullptr, *d_sm = nullptr;
  OWMI_HIP(hipMalloc(&d_lf0, nseed * sizeof(int)));
  OWMI_HIP(hipMalloc(&d_lfp, nseed * sizeof(int)));
  OWMI_HIP(hipMalloc(&d_lfm, nseed * sizeof(int)));
  OWMI_HIP(hipMalloc(&d_s0, 2 * nseed * sizeof(double)));
  OWMI_HIP(hipMalloc(&d_sp, 2 * nseed * sizeof(double)));
  OWMI_HIP(hipMalloc(&d_sm, 2 * nseed * sizeof(double)));
  launch_seed_finder(&p, d_x, d_v0, d_k0, d_vp, d_kp, d_vm, d_km, d_U0, d_U2, d_lf0, d_s0,
                     d_lfp, d_sp, d_lfm, d_sm);
  OWMI_HIP(hipDeviceSynchronize());
  h.lf0.resize(nseed); h.lfp.resize(nseed); h.lfm.resize(nseed);
  h.s0.resize(2 * nseed); h.sp.resize(2 * nseed); h.sm.resize(2 * nseed);
  OWMI_HIP(hipMemcpy(h.lf0.data(), d_lf0, nseed * sizeof(int), hipMemcpyDeviceToHost));
  OWMI_HIP(hipMemcpy(h.lfp.data(), d_lfp, nseed * sizeof(int), hipMemcpyDeviceToHost));
  OWMI_HIP(hipMemcpy(h.lfm.data(), d_lfm, nseed * sizeof(int), hipMemcpyDeviceToHost));
  OWMI_HIP(hipMemcpy(h.s0.data(), d_s0, 2 * nseed * sizeof(double), hipMemcpyDeviceToHost));
  OWMI_HIP(hipMemcpy(h.sp.data(), d_sp, 2 * nseed * sizeof(double), hipMemcpyDeviceToHost));
  OWMI_HIP(hipMemcpy(h.sm.data(), d_sm, 2 * nseed * sizeof(double), hipMemcpyDeviceToHost));
  void *tmp[] = {d_x, d_v0, d_vp, d_vm, d_k0, d_kp, d_km, d_U0, d_U2,
                 d_lf0, d_lfp, d_lfm, d_s0, d_sp, d_sm};
  for (void *q : tmp) (void)hipFree(q);
}

// Per-precision device tables from the host tables.
template <typename T>
static void build_tables(owmi_plan_impl &p, PlanTables<T> &t, const HostTables &h,
                         const std::vector<LD> &xs, const std::vector<LD> &chirp) {
  const int lmax = p.lmax, mmax = p.mmax;
  const long long np = p.npairs;
  const long long nelem = nelem_of(lmax);
  const size_t nseed = (size_t)(mmax + 1) * np;

  std::vector<T> hx(np);
  for (long long i = 0; i < np; ++i) hx[i] = (T)xs[i];

  // seeds in normalised units mu = slam / c
  std::vector<T> s0(2 * nseed), sp(2 * nseed), sm(2 * nseed);
  for (int m = 0; m <= mmax; ++m) {
    long long ms = mstart_of(m, lmax);
    for (long long i = 0; i < np; ++i) {
      size_t k = (size_t)m * np + i;
      struct F { const std::vector<int> &lf; const std::vector<double> &sv;
                 const std::vector<LD> &c; std::vector<T> &out; int lstart; };
      F fs[3] = {{h.lf0, h.s0, h.c0, s0, m}, {h.lfp, h.sp, h.c2, sp, std::max(m, 2)},
                 {h.lfm, h.sm, h.c2, sm, std::max(m, 2)}};
      for (F &f : fs) {
        int lf = f.lf[k];
        T a = 0, b = 0;
        if (lf <= lmax) {
          b = (T)((LD)f.sv[2 * k + 1] / f.c[(size_t)(ms + lf)]);
          if (lf - 1 >= f.lstart) a = (T)((LD)f.sv[2 * k] / f.c[(size_t)(ms + lf - 1)]);
        }
        f.out[2 * k] = a;
        f.out[2 * k + 1] = b;
      }
    }
  }
  std::vector<T> c0(2 * (size_t)nelem), c2(3 * (size_t)nelem);
  for (size_t i = 0; i < (size_t)nelem; ++i) {
    c0[2 * i] = (T)h.A0[i];
    c0[2 * i + 1] = (T)h.c0[i];
    c2[3 * i] = (T)h.A2[i];
    c2[3 * i + 1] = (T)h.B2[i];
    c2[3 * i + 2] = (T)h.c2[i];
  }

  // Twiddles exp(-2 pi i k / Mmax), k < Mmax, for the largest FFT length.
  std::vector<T> tw(2 * (size_t)p.mmax_fft);
  for (long long k = 0; k < p.mmax_fft; ++k) {
    LD ang = -2.0L * kPi * (LD)k / (LD)p.mmax_fft;
    tw[2 * k] = (T)cosl(ang);
    tw[2 * k + 1] = (T)sinl(ang);
  }
  std::vector<T> chirpT(chirp.size());
  for (size_t i = 0; i < chirp.size(); ++i) chirpT[i] = (T)chirp[i];
  std::vector<T> rw(p.nrings), ph(p.nrings);
  for (long long r = 0; r < p.nrings; ++r) {
    rw[r] = (T)p.h_weight[r];
    ph[r] = (T)p.h_phi0[r];
  }

  t.x = upload(hx);
  t.lf0 = upload(h.lf0);
  t.s0 = upload(s0);
  t.lfp = upload(h.lfp);
  t.sp = upload(sp);
  t.lfm = upload(h.lfm);
  t.sm = upload(sm);
  t.coef0 = upload(c0);
  t.coef2 = upload(c2);
  t.twiddle = upload(tw);
  t.chirp = upload(chirpT);
  t.ringw = upload(rw);
  t.phi0 = upload(ph);
  OWMI_HIP(hipMalloc(&t.bhat, std::max<size_t>(1, p.nbhat) * 2 * sizeof(T)));
  t.ready = true;
}

}  // namespace owmi

using namespace owmi;

template <typename T>
static PlanTables<T> &tables(owmi_plan_impl &p);
template <>
PlanTables<double> &tables<double>(owmi_plan_impl &p) { return p.t64; }
template <>
PlanTables<float> &tables<float>(owmi_plan_impl &p) { return p.t32; }

// Bluestein kernel spectra are computed once in double on the device; the
// float tables are converted from them.
__global__ void bhat_to_float(const double *src, float *dst, long long n) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = (float)src[i];
}

namespace {

// Plan over the given rings of a map whose components are map_npix pixels
// apart (map_npix < 0: the rings make up the whole map, map_npix = sum nphi;
// map_npix == 0 is rejected).
int create_plan(owmi_sht_plan **out, long long nrings, const double *theta,
                const double *weight, const long long *nphi, const long long *offsets,
                const double *phi0, int lmax, long long map_npix) {
  return guard([&] {
    OWMI_REQUIRE(out != nullptr, "plan pointer is NULL");
    OWMI_REQUIRE(nrings >= 1, "need at least one ring");
    OWMI_REQUIRE(lmax >= 0, "lmax must be >= 0");
    OWMI_REQUIRE(map_npix != 0, "map_npix must be >= 1");
    *out = nullptr;
    auto *plan = new owmi_sht_plan();
    owmi_plan_impl &p = plan->impl;
    try {
      OWMI_HIP(hipGetDevice(&p.device));
      p.lmax = lmax;
      p.mmax = lmax;
      p.nrings = nrings;
      p.h_theta.assign(theta, theta + nrings);
      p.h_weight.assign(weight, weight + nrings);
      p.h_nphi.assign(nphi, nphi + nrings);
      p.h_offsets.assign(offsets, offsets + nrings);
      p.h_phi0.assign(phi0, phi0 + nrings);
      long long sum_nphi = 0;
      for (long long r = 0; r < nrings; ++r) {
        OWMI_REQUIRE(nphi[r] >= 1, "nphi must be >= 1");
        OWMI_REQUIRE(offsets[r] >= 0, "offsets must be >= 0");
        sum_nphi += nphi[r];
      }
      p.npix = map_npix < 0 ? sum_nphi : map_npix;
      for (long long r = 0; r < nrings; ++r)
        OWMI_REQUIRE(offsets[r] + nphi[r] <= p.npix, "ring extends past the map");

      // Mirror pairs: rings visited in ascending theta; a ring with
      // theta < pi/2 takes the lowest-index unused ring with
      // |theta_i + theta_j - pi| < 1e-12 as its south partner.
      std::vector<long long> order(nrings);
      std::iota(order.begin(), order.end(), 0);
      std::stable_sort(order.begin(), order.end(),
                       [&](long long a, long long b) { return theta[a] < theta[b]; });
      std::vector<char> used(nrings, 0);
      const double pi = 3.141592653589793;
      for (long long oi = 0; oi < nrings; ++oi) {
        long long i = order[oi];
        if (used[i]) continue;
        used[i] = 1;
        int j = -1;
        if (theta[i] < pi / 2) {
          // candidates lie near pi - theta_i in the sorted order
          double target = pi - theta[i];
          auto it = std::lower_bound(order.begin(), order.end(), target - 1e-11,
                                     [&](long long a, double v) { return theta[a] < v; });
          long long best = -1;
          for (; it != order.end() && theta[*it] <= target + 1e-11; ++it) {
            long long c = *it;
            if (!used[c] && std::fabs(theta[c] + theta[i] - pi) < 1e-12)
              if (best < 0 || c < best) best = c;
          }
          if (best >= 0) {
            j = (int)best;
            used[best] = 1;
          }
        }
        p.h_pair_n.push_back((int)i);
        p.h_pair_s.push_back(j);
      }
      p.npairs = (long long)p.h_pair_n.size();

      std::vector<LD> xs(p.npairs), thn(p.npairs);
      for (long long k = 0; k < p.npairs; ++k) {
        LD th = (LD)theta[p.h_pair_n[k]];
        thn[k] = th;
        xs[k] = (LD)(double)cosl(th);  // x rounded to double, shared by the pair
      }

      // Ring FFT tables: one Bluestein length per distinct nphi.
      std::map<long long, int> len_index;
      std::vector<long long> lens, chirp_off, bhat_off;
      std::vector<int> logms;
      std::vector<LD> chirp;
      size_t nbhat = 0;
      std::vector<int> ring_logm(nrings);
      std::vector<long long> ring_coff(nrings), ring_boff(nrings);
      p.mmax_fft = 16;
      for (long long r = 0; r < nrings; ++r) {
        long long n = nphi[r];
        auto it = len_index.find(n);
        int li;
        if (it == len_index.end()) {
          li = (int)lens.size();
          len_index[n] = li;
          lens.push_back(n);
          int lm = ilog2_ceil(2 * n - 1);
          if (lm < 4) lm = 4;  // the ring FFT works on 16 slots per thread
          logms.push_back(lm);
          chirp_off.push_back((long long)chirp.size() / 2);
          bhat_off.push_back((long long)nbhat);
          nbhat += (size_t)1 << lm;
          for (long long k = 0; k < n; ++k) {
            long long q = (k * k) % (2 * n);
            LD ang = -kPi * (LD)q / (LD)n;
            chirp.push_back(cosl(ang));
            chirp.push_back(sinl(ang));
          }
          p.mmax_fft = std::max(p.mmax_fft, 1 << lm);
        } else {
          li = it->second;
        }
        ring_logm[r] = logms[li];
        ring_coff[r] = chirp_off[li];
        ring_boff[r] = bhat_off[li];
      }
      p.nchirp = chirp.size() / 2;
      p.nbhat = nbhat;
      p.log2_mmax_fft = ilog2_ceil(p.mmax_fft);
      p.uniform_logm = logms.size() == 1 ? logms[0] : 0;

      // FFT jobs: mirror pairs share nphi on symmetric grids; other rings
      // are paired with the next ring of equal nphi, leftovers run alone.
      std::vector<int> ja, jb;
      std::vector<char> done(nrings, 0);
      for (long long k = 0; k < p.npairs; ++k) {
        int a = p.h_pair_n[k], b = p.h_pair_s[k];
        if (b >= 0 && nphi[a] == nphi[b]) {
          ja.push_back(a);
          jb.push_back(b);
          done[a] = done[b] = 1;
        }
      }
      std::map<long long, int> pending;
      for (long long r = 0; r < nrings; ++r) {
        if (done[r]) continue;
        auto it = pending.find(nphi[r]);
        if (it == pending.end()) {
          pending[nphi[r]] = (int)r;
        } else {
          ja.push_back(it->second);
          jb.push_back((int)r);
          pending.erase(it);
        }
      }
      for (auto &kv : pending) {
        ja.push_back(kv.second);
        jb.push_back(-1);
      }
      p.nfft_jobs = (long long)ja.size();

      p.pair_n = upload(p.h_pair_n);
      p.pair_s = upload(p.h_pair_s);
      p.nphi = upload(p.h_nphi);
      p.offsets = upload(p.h_offsets);
      p.ring_logm = upload(ring_logm);
      p.ring_chirp_off = upload(ring_coff);
      p.ring_bhat_off = upload(ring_boff);
      p.fft_ring_a = upload(ja);
      p.fft_ring_b = upload(jb);
      {
        // slot of each ring in the pair-major coefficients: every ring is
        // the north or the south ring of exactly one pair
        std::vector<int> slot(nrings, -1);
        for (long long k = 0; k < p.npairs; ++k) {
          slot[p.h_pair_n[k]] = (int)(2 * k);
          if (p.h_pair_s[k] >= 0) slot[p.h_pair_s[k]] = (int)(2 * k + 1);
        }
        for (long long r = 0; r < nrings; ++r)
          OWMI_REQUIRE(slot[r] >= 0, "internal: a ring belongs to no pair");
        p.ring_slot = upload(slot);
      }

      HostTables ht;
      build_host_tables(p, thn, xs, ht);
      build_tables<double>(p, p.t64, ht, xs, chirp);
      {
        // kernel spectra: on the device when the FFT fits LDS, else on the host
        std::vector<long long> dl, dc, db;
        std::vector<int> dm;
        for (size_t li = 0; li < lens.size(); ++li) {
          if (((size_t)16 << logms[li]) <= kInitLds) {
            dl.push_back(lens[li]);
            dm.push_back(logms[li]);
            dc.push_back(chirp_off[li]);
            db.push_back(bhat_off[li]);
          } else {
            std::vector<double> h = host_bhat(lens[li], logms[li]);
            OWMI_HIP(hipMemcpy(p.t64.bhat + 2 * bhat_off[li], h.data(), h.size() * sizeof(double),
                               hipMemcpyHostToDevice));
          }
        }
        launch_bluestein_init(&p, dl, dm, dc, db);
      }
      OWMI_HIP(hipDeviceSynchronize());
      build_tables<float>(p, p.t32, ht, xs, chirp);
      long long nb = 2 * (long long)p.nbhat;
      hipLaunchKernelGGL(bhat_to_float, dim3(ceil_div(nb, 256)), dim3(256), 0, 0,
                         p.t64.bhat, p.t32.bhat, nb);
      OWMI_HIP(hipDeviceSynchronize());
    } catch (...) {
      owmi_sht_plan_destroy(plan);
      throw;
    }
    *out = plan;
  });
}

}  // namespace

extern "C" {

int owmi_sht_plan_create(owmi_sht_plan **out, long long nrings, const double *theta,
                         const double *weight, const long long *nphi, const long long *offsets,
                         const double *phi0, int lmax) {
  return create_plan(out, nrings, theta, weight, nphi, offsets, phi0, lmax, -1);
}

int owmi_sht_plan_create_rings(owmi_sht_plan **out, long long nrings, const double *theta,
                               const double *weight, const long long *nphi,
                               const long long *offsets, const double *phi0, int lmax,
                               long long map_npix) {
  return create_plan(out, nrings, theta, weight, nphi, offsets, phi0, lmax,
                     map_npix < 1 ? 0 : map_npix);
}

int owmi_sht_plan_destroy(owmi_sht_plan *plan) {
  if (!plan) return 0;
  owmi_plan_impl &p = plan->impl;
  int dev = 0;
  (void)hipGetDevice(&dev);
  (void)hipSetDevice(p.device);
  free_tables(p.t64);
  free_tables(p.t32);
  void *ptrs[] = {p.pair_n, p.pair_s, p.nphi, p.offsets, p.ring_logm, p.ring_chirp_off,
                  p.ring_bhat_off, p.fft_ring_a, p.fft_ring_b, p.ring_slot, p.work, p.work2,
                  p.work3, p.work4};
  for (void *q : ptrs)
    if (q) (void)hipFree(q);
  for (void *q : p.retired) (void)hipFree(q);
  (void)hipSetDevice(dev);
  delete plan;
  return 0;
}

long long owmi_sht_plan_npix(const owmi_sht_plan *plan) { return plan ? plan->impl.npix : -1; }

int owmi_sht_plan_tune(owmi_sht_plan *plan, long long min_batch, int param, int value) {
  return guard([&] {
    OWMI_REQUIRE(plan != nullptr, "plan is NULL");
    OWMI_REQUIRE(min_batch >= 1, "min_batch must be >= 1");
    OWMI_REQUIRE(param >= OWMI_TUNE_MFMA_MIN && param <= OWMI_TUNE_MFMA64_MIN,
                 "unknown tuning parameter");
    bool ok = value == 0;
    switch (param) {
      case OWMI_TUNE_MFMA_MIN:
      case OWMI_TUNE_MFMA64_MIN: ok = ok || value >= 1; break;
      case OWMI_TUNE_ADJ_CT0:
      case OWMI_TUNE_ADJ_CT2: ok = ok || value == 2 || value == 4; break;
      case OWMI_TUNE_SYNTH_NB:
      case OWMI_TUNE_ADJ_NB:
      case OWMI_TUNE_ADJ_K0: ok = ok || value == 1 || value == 2 || value == 4; break;
      case OWMI_TUNE_ADJ_K2: ok = ok || value == 1 || value == 2; break;
    }
    OWMI_REQUIRE(ok, "value not allowed for this tuning parameter");
    auto &tune = plan->impl.tune;
    auto it = std::find_if(tune.begin(), tune.end(), [&](const owmi_plan_impl::TuneEntry &e) {
      return e.min_batch == min_batch;
    });
    if (it == tune.end()) {
      owmi_plan_impl::TuneEntry e;
      e.min_batch = min_batch;
      tune.push_back(e);
      std::sort(tune.begin(), tune.end(),
                [](const owmi_plan_impl::TuneEntry &a, const owmi_plan_impl::TuneEntry &b) {
                  return a.min_batch < b.min_batch;
                });
      it = std::find_if(tune.begin(), tune.end(), [&](const owmi_plan_impl::TuneEntry &e) {
        return e.min_batch == min_batch;
      });
    }
    it->v[param] = value;
  });
}

int owmi_sht_plan_tune_reset(owmi_sht_plan *plan) {
  return guard([&] {
    OWMI_REQUIRE(plan != nullptr, "plan is NULL");
    plan->impl.tune.clear();
  });
}

}  // extern "C"

// Bytes of ring coefficients for `rows` (transform, component) rows, large
// enough for the synthesis and the adjoint layouts (2 npairs slots per m,
// the adjoint's m rounded up to whole blocks).
template <typename T>
static size_t coef_bytes(const owmi_plan_impl &p, long long rows) {
  const long long per_row = 2 * p.npairs * coef_nmb<T>(p.mmax) * coef_mb<T>();
  return (size_t)rows * per_row * sizeof(cplx_t<T>);
}

template <typename T>
static void check_fft_fits(owmi_plan_impl &p) {
  // M / 16 threads per ring workgroup, at most 1024; the buffer lives in
  // LDS up to 160 KiB and in a global-memory workspace beyond
  OWMI_REQUIRE(p.mmax_fft <= 16384,
               "ring length too large for the ring FFT (Bluestein length " +
                   std::to_string(p.mmax_fft) + " > 16384)");
}

template <typename T>
static int do_synthesis(owmi_sht_plan *plan, int spin, const void *alm, long long alm_ts,
                        void *map, long long map_ts, long long ntrans, int weighted,
                        void *stream) {
  return guard([&] {
    OWMI_REQUIRE(plan != nullptr, "plan is NULL");
    owmi_plan_impl &p = plan->impl;
    OWMI_REQUIRE(spin == 0 || spin == 2, "spin must be 0 or 2");
    OWMI_REQUIRE(spin <= p.lmax, "Spin exceeds lmax");
    check_fft_fits<T>(p);
    if (ntrans == 0) return;
    int nc = spin == 0 ? 1 : 2;
    auto *G = static_cast<cplx_t<T> *>(p.ensure_work(coef_bytes<T>(p, ntrans * nc)));
    hipStream_t st = as_stream(stream);
    launch_synthesis<T>(&p, spin, static_cast<const cplx_t<T> *>(alm), alm_ts, G, ntrans, st);
    launch_ring_synth<T>(&p, G, static_cast<T *>(map), map_ts, nc, ntrans, weighted, st);
  });
}

template <typename T>
static int do_adjoint(owmi_sht_plan *plan, int spin, const void *map, long long map_ts,
                      void *alm, long long alm_ts, long long ntrans, int weighted,
                      void *stream) {
  return guard([&] {
    OWMI_REQUIRE(plan != nullptr, "plan is NULL");
    owmi_plan_impl &p = plan->impl;
    OWMI_REQUIRE(spin == 0 || spin == 2, "spin must be 0 or 2");
    OWMI_REQUIRE(spin <= p.lmax, "Spin exceeds lmax");
    check_fft_fits<T>(p);
    if (ntrans == 0) return;
    int nc = spin == 0 ? 1 : 2;
    auto *G = static_cast<cplx_t<T> *>(p.ensure_work(coef_bytes<T>(p, ntrans * nc)));
    hipStream_t st = as_stream(stream);
    launch_ring_adjoint<T>(&p, static_cast<const T *>(map), map_ts, G, nc, ntrans, weighted, st);
    launch_adjoint<T>(&p, spin, G, static_cast<cplx_t<T> *>(alm), alm_ts, ntrans, st);
  });
}

template <typename T>
static int do_sandwich(owmi_sht_plan *plan, int spin, const void *alm_in, long long in_ts,
                       void *alm_out, long long out_ts, const void *diag, long long diag_ts,
                       long long ntrans, void *stream) {
  return guard([&] {
    OWMI_REQUIRE(plan != nullptr, "plan is NULL");
    owmi_plan_impl &p = plan->impl;
    OWMI_REQUIRE(spin == 0 || spin == 2, "spin must be 0 or 2");
    OWMI_REQUIRE(spin <= p.lmax, "Spin exceeds lmax");
    check_fft_fits<T>(p);
    if (ntrans == 0) return;
    int nc = spin == 0 ? 1 : 2;
    auto *G = static_cast<cplx_t<T> *>(p.ensure_work(coef_bytes<T>(p, ntrans * nc)));
    auto *GP = static_cast<cplx_t<T> *>(p.ensure_work4(coef_bytes<T>(p, ntrans * nc)));
    hipStream_t st = as_stream(stream);
    launch_synthesis<T>(&p, spin, static_cast<const cplx_t<T> *>(alm_in), in_ts, G, ntrans, st);
    launch_ring_sandwich<T>(&p, G, GP, static_cast<const T *>(diag), diag_ts, nc, ntrans, st);
    launch_adjoint<T>(&p, spin, GP, static_cast<cplx_t<T> *>(alm_out), out_ts, ntrans, st);
  });
}

template <typename T>
static int do_sandwich_mult(owmi_sht_plan *plan, const void *diag, int ncomp, void *mult,
                            void *stream) {
  return guard([&] {
    OWMI_REQUIRE(plan != nullptr, "plan is NULL");
    owmi_plan_impl &p = plan->impl;
    OWMI_REQUIRE(ncomp >= 1, "ncomp must be >= 1");
    OWMI_REQUIRE(conv_logm(&p) > 0, "the spectral sandwich does not apply to this plan");
    const long long M = 1LL << conv_logm(&p);
    launch_conv_mult<T>(&p, static_cast<const T *>(diag), p.npix, ncomp, static_cast<T *>(mult),
                        p.nrings * M, as_stream(stream));
  });
}

template <typename T>
static int do_sandwich_spectral(owmi_sht_plan *plan, int spin, const void *alm_in,
                                long long in_ts, void *alm_out, long long out_ts,
                                const void *mult, long long ntrans, void *stream) {
  return guard([&] {
    OWMI_REQUIRE(plan != nullptr, "plan is NULL");
    owmi_plan_impl &p = plan->impl;
    OWMI_REQUIRE(spin == 0 || spin == 2, "spin must be 0 or 2");
    OWMI_REQUIRE(spin <= p.lmax, "Spin exceeds lmax");
    OWMI_REQUIRE(conv_logm(&p) > 0, "the spectral sandwich does not apply to this plan");
    check_fft_fits<T>(p);
    if (ntrans == 0) return;
    const int nc = spin == 0 ? 1 : 2;
    const long long M = 1LL << conv_logm(&p);
    auto *G = static_cast<cplx_t<T> *>(p.ensure_work(coef_bytes<T>(p, ntrans * nc)));
    auto *GP = static_cast<cplx_t<T> *>(p.ensure_work4(coef_bytes<T>(p, ntrans * nc)));
    hipStream_t st = as_stream(stream);
    launch_synthesis<T>(&p, spin, static_cast<const cplx_t<T> *>(alm_in), in_ts, G, ntrans, st);
    launch_ring_conv<T>(&p, G, GP, static_cast<const T *>(mult), p.nrings * M, nc, ntrans, st);
    launch_adjoint<T>(&p, spin, GP, static_cast<cplx_t<T> *>(alm_out), out_ts, ntrans, st);
  });
}

template <typename T>
static int do_legendre(owmi_sht_plan *plan, int spin, int direction, void *alm,
                       long long alm_ts, long long ntrans, void *stream) {
  return guard([&] {
    OWMI_REQUIRE(plan != nullptr, "plan is NULL");
    owmi_plan_impl &p = plan->impl;
    OWMI_REQUIRE(spin == 0 || spin == 2, "spin must be 0 or 2");
    OWMI_REQUIRE(direction == 0 || direction == 1, "direction must be 0 or 1");
    if (ntrans == 0) return;
    int nc = spin == 0 ? 1 : 2;
    auto *G = static_cast<cplx_t<T> *>(p.ensure_work(coef_bytes<T>(p, ntrans * nc)));
    hipStream_t st = as_stream(stream);
    if (direction == 0)
      launch_synthesis<T>(&p, spin, static_cast<const cplx_t<T> *>(alm), alm_ts, G, ntrans, st);
    else
      launch_adjoint<T>(&p, spin, G, static_cast<cplx_t<T> *>(alm), alm_ts, ntrans, st);
  });
}

extern "C" {

int owmi_sht_legendre(owmi_sht_plan *plan, int dtype, int spin, int direction, void *alm,
                      long long alm_tstride, long long ntrans, void *stream) {
  if (dtype == OWMI_F64)
    return do_legendre<double>(plan, spin, direction, alm, alm_tstride, ntrans, stream);
  if (dtype == OWMI_F32)
    return do_legendre<float>(plan, spin, direction, alm, alm_tstride, ntrans, stream);
  set_error("dtype must be OWMI_F32 or OWMI_F64");
  return 1;
}

int owmi_sht_synthesis(owmi_sht_plan *plan, int dtype, int spin, const void *alm,
                       long long alm_tstride, void *map, long long map_tstride, long long ntrans,
                       int weighted, void *stream) {
  if (dtype == OWMI_F64)
    return do_synthesis<double>(plan, spin, alm, alm_tstride, map, map_tstride, ntrans, weighted,
                                stream);
  if (dtype == OWMI_F32)
    return do_synthesis<float>(plan, spin, alm, alm_tstride, map, map_tstride, ntrans, weighted,
                               stream);
  set_error("dtype must be OWMI_F32 or OWMI_F64");
  return 1;
}

int owmi_sht_adjoint(owmi_sht_plan *plan, int dtype, int spin, const void *map,
                     long long map_tstride, void *alm, long long alm_tstride, long long ntrans,
                     int weighted, void *stream) {
  if (dtype == OWMI_F64)
    return do_adjoint<double>(plan, spin, map, map_tstride, alm, alm_tstride, ntrans, weighted,
                              stream);
  if (dtype == OWMI_F32)
    return do_adjoint<float>(plan, spin, map, map_tstride, alm, alm_tstride, ntrans, weighted,
                             stream);
  set_error("dtype must be OWMI_F32 or OWMI_F64");
  return 1;
}

long long owmi_sht_sandwich_length(owmi_sht_plan *plan) {
  if (plan == nullptr) return 0;
  const int lm = conv_logm(&plan->impl);
  return lm > 0 ? 1LL << lm : 0;
}

int owmi_sht_sandwich_multipliers(owmi_sht_plan *plan, int dtype, const void *pixdiag, int ncomp,
                                  void *mult, void *stream) {
  if (dtype == OWMI_F64) return do_sandwich_mult<double>(plan, pixdiag, ncomp, mult, stream);
  if (dtype == OWMI_F32) return do_sandwich_mult<float>(plan, pixdiag, ncomp, mult, stream);
  set_error("dtype must be OWMI_F32 or OWMI_F64");
  return 1;
}

int owmi_sht_pixel_sandwich_spectral(owmi_sht_plan *plan, int dtype, int spin,
                                     const void *alm_in, long long alm_in_tstride, void *alm_out,
                                     long long alm_out_tstride, const void *mult,
                                     long long ntrans, void *stream) {
  if (dtype == OWMI_F64)
    return do_sandwich_spectral<double>(plan, spin, alm_in, alm_in_tstride, alm_out,
                                        alm_out_tstride, mult, ntrans, stream);
  if (dtype == OWMI_F32)
    return do_sandwich_spectral<float>(plan, spin, alm_in, alm_in_tstride, alm_out,
                                       alm_out_tstride, mult, ntrans, stream);
  set_error("dtype must be OWMI_F32 or OWMI_F64");
  return 1;
}

int owmi_sht_pixel_sandwich(owmi_sht_plan *plan, int dtype, int spin, const void *alm_in,
                            long long alm_in_tstride, void *alm_out, long long alm_out_tstride,
                            const void *pixdiag, long long pix_tstride, long long ntrans,
                            void *stream) {
  if (dtype == OWMI_F64)
    return do_sandwich<double>(plan, spin, alm_in, alm_in_tstride, alm_out, alm_out_tstride,
                               pixdiag, pix_tstride, ntrans, stream);
  if (dtype == OWMI_F32)
    return do_sandwich<float>(plan, spin, alm_in, alm_in_tstride, alm_out, alm_out_tstride,
                              pixdiag, pix_tstride, ntrans, stream);
  set_error("dtype must be OWMI_F32 or OWMI_F64");
  return 1;
}

}  // extern "C"
