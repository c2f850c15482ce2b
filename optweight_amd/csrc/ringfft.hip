// Ring Fourier transforms between the ring coefficients G[row][m]
// (row = (transform * ncomp + comp) * nrings + ring, m = 0..mmax) and
// ring-major real maps.
//
// One workgroup transforms two real rings of equal length N packed as one
// complex sequence (z = a + i b). Any N is handled with Bluestein's
// algorithm: DFT_N(x)_k = w_k sum_j (x_j w_j) conj(w_{k-j}), w_n =
// exp(-i pi n^2 / N), evaluated as a cyclic convolution of length
// M = 2^ceil(log2(2N - 1)) with a radix-2 decimation-in-time FFT
// (bit-reversed input, natural output) and a decimation-in-frequency
// inverse (natural input, bit-reversed output), both in LDS. The kernel
// spectrum FFT_M(conj w) / M is precomputed per distinct N.
//
// Frequencies m > N/2 alias onto bin m mod N (reduced rings of equal-area
// grids), and the first-pixel longitude phi0 enters as exp(i m phi0).
#include "sht_internal.hpp"

namespace owmi {

namespace {

template <typename T>
__device__ __forceinline__ cplx_t<T> cmul(cplx_t<T> a, cplx_t<T> b) {
  return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
template <typename T>
__device__ __forceinline__ cplx_t<T> cconj(cplx_t<T> a) {
  return {a.x, -a.y};
}
template <typename T>
__device__ __forceinline__ cplx_t<T> cadd(cplx_t<T> a, cplx_t<T> b) {
  return {a.x + b.x, a.y + b.y};
}
template <typename T>
__device__ __forceinline__ cplx_t<T> csub(cplx_t<T> a, cplx_t<T> b) {
  return {a.x - b.x, a.y - b.y};
}

__device__ __forceinline__ void sincos_t(double x, double *s, double *c) { sincos(x, s, c); }
__device__ __forceinline__ void sincos_t(float x, float *s, float *c) { sincosf(x, s, c); }

__device__ __forceinline__ unsigned bitrev(unsigned k, int logm) {
  return __brev(k) >> (32 - logm);
}

// In-place radix-2 DIT FFT (forward, exp(-2 pi i / M)) on bit-reversed
// input; output in natural order.
template <typename T>
__device__ void fft_dit(cplx_t<T> *buf, int logm, const cplx_t<T> *tw, int log_mmax) {
  const int M = 1 << logm;
  const int half = M >> 1;
  for (int s = 0; s < logm; ++s) {
    const int h = 1 << s;
    const int twshift = log_mmax - s - 1;  // twiddle index j * (Mmax / 2h)
    for (int b = threadIdx.x; b < half; b += blockDim.x) {
      int j = b & (h - 1);
      int i0 = ((b >> s) << (s + 1)) + j;
      int i1 = i0 + h;
      cplx_t<T> w = tw[j << twshift];
      cplx_t<T> u = buf[i0];
      cplx_t<T> v = cmul<T>(buf[i1], w);
      buf[i0] = cadd<T>(u, v);
      buf[i1] = csub<T>(u, v);
    }
    __syncthreads();
  }
}

// In-place radix-2 DIF inverse FFT (exp(+2 pi i / M), unnormalised) on
// natural-order input; output in bit-reversed order.
template <typename T>
__device__ void ifft_dif(cplx_t<T> *buf, int logm, const cplx_t<T> *tw, int log_mmax) {
  const int M = 1 << logm;
  const int half = M >> 1;
  for (int s = logm - 1; s >= 0; --s) {
    const int h = 1 << s;
    const int twshift = log_mmax - s - 1;
    for (int b = threadIdx.x; b < half; b += blockDim.x) {
      int j = b & (h - 1);
      int i0 = ((b >> s) << (s + 1)) + j;
      int i1 = i0 + h;
      cplx_t<T> w = cconj<T>(tw[j << twshift]);
      cplx_t<T> u = buf[i0];
      cplx_t<T> v = buf[i1];
      buf[i0] = cadd<T>(u, v);
      buf[i1] = cmul<T>(csub<T>(u, v), w);
    }
    __syncthreads();
  }
}

// After loading buf[bitrev(j)] = z_j w_j (zero for j >= N): leaves
// DFT_N(z)_k / w_k ... i.e. the Bluestein convolution c_k at buf[bitrev(k)].
template <typename T>
__device__ void bluestein_core(cplx_t<T> *buf, int logm, const cplx_t<T> *tw, int log_mmax,
                               const cplx_t<T> *bhat) {
  fft_dit<T>(buf, logm, tw, log_mmax);
  const int M = 1 << logm;
  for (int k = threadIdx.x; k < M; k += blockDim.x) buf[k] = cmul<T>(buf[k], bhat[k]);
  __syncthreads();
  ifft_dif<T>(buf, logm, tw, log_mmax);
}

// Folded, phase-shifted Hermitian spectrum bin k of one ring:
// Z_k = sum_{m = k mod N} C_m + sum_{m >= 1, m = -k mod N} conj(C_m),
// C_0 = Re F_0, C_m = F_m exp(i m phi0).
template <typename T>
__device__ cplx_t<T> fold_bin(const cplx_t<T> *F, long long k, long long N, int mmax, T phi0) {
  cplx_t<T> acc = {0, 0};
  for (long long m = k; m <= mmax; m += N) {
    cplx_t<T> c = F[m];
    if (m == 0) {
      c.y = 0;
    } else if (phi0 != (T)0) {
      T sn, cs;
      sincos_t((T)m * phi0, &sn, &cs);
      c = cmul<T>(c, cplx_t<T>{cs, sn});
    }
    acc = cadd<T>(acc, c);
  }
  for (long long m = (k == 0 ? N : N - k); m <= mmax; m += N) {
    cplx_t<T> c = F[m];
    if (phi0 != (T)0) {
      T sn, cs;
      sincos_t((T)m * phi0, &sn, &cs);
      c = cmul<T>(c, cplx_t<T>{cs, sn});
    }
    acc = cadd<T>(acc, cconj<T>(c));
  }
  return acc;
}

enum Mode { kSynth = 0, kAdjoint = 1, kSandwich = 2 };

struct RingArgs {
  const int *fft_a;
  const int *fft_b;
  const long long *nphi;
  const long long *offsets;
  const int *ring_logm;
  const long long *chirp_off;
  const long long *bhat_off;
  long long nrings;
  long long npix;
  int mmax;
  int log_mmax_fft;
  int ncomp;
  int weighted;
};

template <typename T, int MODE>
__global__ void __launch_bounds__(256)
ring_kernel(RingArgs a, const cplx_t<T> *twiddle, const cplx_t<T> *chirp,
            const cplx_t<T> *bhat_all, const T *ringw, const T *phi0s, const cplx_t<T> *Gin,
            cplx_t<T> *Gout, const T *map_in, T *map_out, long long map_ts, const T *diag,
            long long diag_ts) {
  extern __shared__ unsigned char smem_raw[];
  cplx_t<T> *buf = reinterpret_cast<cplx_t<T> *>(smem_raw);

  const long long job = blockIdx.x;
  const long long tc = blockIdx.y;  // transform * ncomp + comp
  const long long t = tc / a.ncomp;
  const int c = (int)(tc % a.ncomp);
  const int ra = a.fft_a[job];
  const int rb = a.fft_b[job];
  const long long N = a.nphi[ra];
  const int logm = a.ring_logm[ra];
  const int M = 1 << logm;
  const cplx_t<T> *w = chirp + a.chirp_off[ra];
  const cplx_t<T> *bhat = bhat_all + a.bhat_off[ra];
  const int mmax = a.mmax;
  const long long G_row_a = tc * a.nrings + ra;
  const long long G_row_b = tc * a.nrings + (rb >= 0 ? rb : 0);
  const T pa = phi0s[ra];
  const T pb = rb >= 0 ? phi0s[rb] : (T)0;

  // Pointers of the two rings in map space.
  const long long off_a = a.offsets[ra];
  const long long off_b = rb >= 0 ? a.offsets[rb] : 0;

  if (MODE == kSynth || MODE == kSandwich) {
    const cplx_t<T> *Fa = Gin + G_row_a * (mmax + 1);
    const cplx_t<T> *Fb = Gin + G_row_b * (mmax + 1);
    // Inverse DFT of Z = Za + i Zb through conj: x = conj(DFT(conj Z)).
    for (int k = threadIdx.x; k < M; k += blockDim.x) {
      cplx_t<T> v = {0, 0};
      if (k < N) {
        cplx_t<T> za = fold_bin<T>(Fa, k, N, mmax, pa);
        cplx_t<T> zb = rb >= 0 ? fold_bin<T>(Fb, k, N, mmax, pb) : cplx_t<T>{0, 0};
        cplx_t<T> Z = {za.x - zb.y, za.y + zb.x};
        v = cmul<T>(cconj<T>(Z), w[k]);
      }
      buf[bitrev(k, logm)] = v;
    }
    __syncthreads();
    bluestein_core<T>(buf, logm, twiddle, a.log_mmax_fft, bhat);
    // c_j sits at buf[bitrev(j)]; DFT(conj Z)_j = w_j c_j and x = conj of
    // it, so ring a gets Re(w_j c_j) and ring b gets -Im(w_j c_j).
    for (int j = threadIdx.x; j < N; j += blockDim.x) {
      unsigned bj = bitrev(j, logm);
      cplx_t<T> o = cmul<T>(w[j], buf[bj]);
      T va = o.x;
      T vb = -o.y;
      if (MODE == kSynth) {
        T *row = map_out + t * map_ts + (long long)c * a.npix;
        if (a.weighted) {
          va *= ringw[ra];
          if (rb >= 0) vb *= ringw[rb];
        }
        row[off_a + j] = va;
        if (rb >= 0) row[off_b + j] = vb;
      } else {
        // Pixel multiply; each thread rewrites only the slot it read.
        const T *drow = diag + t * diag_ts + (long long)c * a.npix;
        va *= drow[off_a + j];
        vb = rb >= 0 ? vb * drow[off_b + j] : (T)0;
        buf[bj] = cplx_t<T>{va, vb};
      }
    }
    if (MODE == kSandwich) {
      __syncthreads();
      // Forward-DFT input: buf[bitrev(j)] = z'_j w_j, zero for j >= N.
      for (int j = threadIdx.x; j < M; j += blockDim.x) {
        unsigned bj = bitrev(j, logm);
        buf[bj] = j < N ? cmul<T>(buf[bj], w[j]) : cplx_t<T>{0, 0};
      }
      __syncthreads();
    }
  }

  if (MODE == kAdjoint) {
    const T *row = map_in + t * map_ts + (long long)c * a.npix;
    for (int j = threadIdx.x; j < M; j += blockDim.x) {
      cplx_t<T> v = {0, 0};
      if (j < N) {
        T va = row[off_a + j];
        T vb = rb >= 0 ? row[off_b + j] : (T)0;
        if (a.weighted) {
          va *= ringw[ra];
          if (rb >= 0) vb *= ringw[rb];
        }
        v = cmul<T>(cplx_t<T>{va, vb}, w[j]);
      }
      buf[bitrev(j, logm)] = v;
    }
    __syncthreads();
  }

  if (MODE == kAdjoint || MODE == kSandwich) {
    bluestein_core<T>(buf, logm, twiddle, a.log_mmax_fft, bhat);
    // X_k = w_k c_k with c_k at buf[bitrev(k)]; finish in place.
    for (int k = threadIdx.x; k < N; k += blockDim.x) {
      unsigned bk = bitrev(k, logm);
      buf[bk] = cmul<T>(w[k], buf[bk]);
    }
    __syncthreads();
    cplx_t<T> *Ga = Gout + G_row_a * (mmax + 1);
    cplx_t<T> *Gb = Gout + G_row_b * (mmax + 1);
    for (int m = threadIdx.x; m <= mmax; m += blockDim.x) {
      long long k = m % N;
      long long kk = (N - k) % N;
      cplx_t<T> X = buf[bitrev((unsigned)k, logm)];
      cplx_t<T> Xc = cconj<T>(buf[bitrev((unsigned)kk, logm)]);
      cplx_t<T> xa = {(T)0.5 * (X.x + Xc.x), (T)0.5 * (X.y + Xc.y)};
      cplx_t<T> d = {(T)0.5 * (X.x - Xc.x), (T)0.5 * (X.y - Xc.y)};
      cplx_t<T> xb = {d.y, -d.x};  // d / i
      if (k == 0) {
        xa.y = 0;
        xb.y = 0;
      }
      if (m > 0 && pa != (T)0) {
        T sn, cs;
        sincos_t(-(T)m * pa, &sn, &cs);
        xa = cmul<T>(xa, cplx_t<T>{cs, sn});
      }
      Ga[m] = xa;
      if (rb >= 0) {
        if (m > 0 && pb != (T)0) {
          T sn, cs;
          sincos_t(-(T)m * pb, &sn, &cs);
          xb = cmul<T>(xb, cplx_t<T>{cs, sn});
        }
        Gb[m] = xb;
      }
    }
  }
}

// bhat for one distinct ring length per block.
__global__ void __launch_bounds__(256)
bluestein_init_kernel(const long long *lens, const int *logms, const long long *chirp_off,
                      const long long *bhat_off, const double2 *chirp, const double2 *twiddle,
                      int log_mmax, double2 *bhat) {
  extern __shared__ unsigned char smem_raw[];
  double2 *buf = reinterpret_cast<double2 *>(smem_raw);
  const int li = blockIdx.x;
  const long long N = lens[li];
  const int logm = logms[li];
  const int M = 1 << logm;
  const double2 *w = chirp + chirp_off[li];
  for (int n = threadIdx.x; n < M; n += blockDim.x) {
    double2 v = {0, 0};
    if (n < N)
      v = cconj<double>(w[n]);
    else if (n > M - N)
      v = cconj<double>(w[M - n]);
    buf[bitrev(n, logm)] = v;
  }
  __syncthreads();
  fft_dit<double>(buf, logm, twiddle, log_mmax);
  const double inv = 1.0 / (double)M;
  double2 *out = bhat + bhat_off[li];
  for (int k = threadIdx.x; k < M; k += blockDim.x) out[k] = {buf[k].x * inv, buf[k].y * inv};
}

template <typename K>
void allow_lds(K kernel, size_t bytes) {
  if (bytes > 64 * 1024)
    OWMI_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(kernel),
                                 hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
}

template <typename T>
RingArgs ring_args(owmi_plan_impl *p, int ncomp, int weighted) {
  RingArgs a;
  a.fft_a = p->fft_ring_a;
  a.fft_b = p->fft_ring_b;
  a.nphi = p->nphi;
  a.offsets = p->offsets;
  a.ring_logm = p->ring_logm;
  a.chirp_off = p->ring_chirp_off;
  a.bhat_off = p->ring_bhat_off;
  a.nrings = p->nrings;
  a.npix = p->npix;
  a.mmax = p->mmax;
  a.log_mmax_fft = p->log2_mmax_fft;
  a.ncomp = ncomp;
  a.weighted = weighted;
  return a;
}

template <typename T>
PlanTables<T> &tab(owmi_plan_impl *p);
template <>
PlanTables<double> &tab<double>(owmi_plan_impl *p) { return p->t64; }
template <>
PlanTables<float> &tab<float>(owmi_plan_impl *p) { return p->t32; }

template <typename T, int MODE>
void launch_ring(owmi_plan_impl *p, const cplx_t<T> *Gin, cplx_t<T> *Gout, const T *map_in,
                 T *map_out, long long map_ts, const T *diag, long long diag_ts, int ncomp,
                 long long ntrans, int weighted, hipStream_t st) {
  long long nrows = ntrans * ncomp;
  OWMI_REQUIRE(nrows < 65536, "too many rows for one ring-FFT launch");
  OWMI_REQUIRE(p->nfft_jobs < (1LL << 31), "too many rings");
  size_t lds = (size_t)p->mmax_fft * sizeof(cplx_t<T>);
  auto kern = ring_kernel<T, MODE>;
  allow_lds(kern, lds);
  PlanTables<T> &tb = tab<T>(p);
  dim3 grid((unsigned)p->nfft_jobs, (unsigned)nrows);
  hipLaunchKernelGGL(kern, grid, dim3(256), lds, st, ring_args<T>(p, ncomp, weighted),
                     reinterpret_cast<const cplx_t<T> *>(tb.twiddle),
                     reinterpret_cast<const cplx_t<T> *>(tb.chirp),
                     reinterpret_cast<const cplx_t<T> *>(tb.bhat), tb.ringw, tb.phi0, Gin, Gout,
                     map_in, map_out, map_ts, diag, diag_ts);
}

}  // namespace

template <typename T>
void launch_ring_synth(owmi_plan_impl *p, const cplx_t<T> *G, T *map, long long map_ts, int ncomp,
                       long long ntrans, int weighted, hipStream_t st) {
  launch_ring<T, kSynth>(p, G, nullptr, nullptr, map, map_ts, nullptr, 0, ncomp, ntrans,
                         weighted, st);
}

template <typename T>
void launch_ring_adjoint(owmi_plan_impl *p, const T *map, long long map_ts, cplx_t<T> *G,
                         int ncomp, long long ntrans, int weighted, hipStream_t st) {
  launch_ring<T, kAdjoint>(p, nullptr, G, map, nullptr, map_ts, nullptr, 0, ncomp, ntrans,
                           weighted, st);
}

template <typename T>
void launch_ring_sandwich(owmi_plan_impl *p, const cplx_t<T> *G, cplx_t<T> *Gout, const T *diag,
                          long long diag_tstride, int ncomp, long long ntrans, hipStream_t st) {
  launch_ring<T, kSandwich>(p, G, Gout, nullptr, nullptr, 0, diag, diag_tstride, ncomp,
                            ntrans, 0, st);
}

void launch_bluestein_init(owmi_plan_impl *p, const std::vector<long long> &lens,
                           const std::vector<int> &logms, const std::vector<long long> &chirp_off,
                           const std::vector<long long> &bhat_off) {
  if (lens.empty()) return;
  long long *d_lens = nullptr, *d_coff = nullptr, *d_boff = nullptr;
  int *d_logm = nullptr;
  OWMI_HIP(hipMalloc(&d_lens, lens.size() * sizeof(long long)));
  OWMI_HIP(hipMalloc(&d_coff, lens.size() * sizeof(long long)));
  OWMI_HIP(hipMalloc(&d_boff, lens.size() * sizeof(long long)));
  OWMI_HIP(hipMalloc(&d_logm, lens.size() * sizeof(int)));
  OWMI_HIP(hipMemcpy(d_lens, lens.data(), lens.size() * sizeof(long long), hipMemcpyHostToDevice));
  OWMI_HIP(hipMemcpy(d_coff, chirp_off.data(), lens.size() * sizeof(long long),
                     hipMemcpyHostToDevice));
  OWMI_HIP(hipMemcpy(d_boff, bhat_off.data(), lens.size() * sizeof(long long),
                     hipMemcpyHostToDevice));
  OWMI_HIP(hipMemcpy(d_logm, logms.data(), lens.size() * sizeof(int), hipMemcpyHostToDevice));
  size_t lds = (size_t)p->mmax_fft * sizeof(double2);
  OWMI_REQUIRE(lds <= 160 * 1024, "Bluestein length exceeds LDS");
  allow_lds(bluestein_init_kernel, lds);
  hipLaunchKernelGGL(bluestein_init_kernel, dim3((unsigned)lens.size()), dim3(256), lds, 0, d_lens,
                     d_logm, d_coff, d_boff, reinterpret_cast<const double2 *>(p->t64.chirp),
                     reinterpret_cast<const double2 *>(p->t64.twiddle), p->log2_mmax_fft,
                     reinterpret_cast<double2 *>(p->t64.bhat));
  OWMI_HIP(hipGetLastError());
  OWMI_HIP(hipDeviceSynchronize());
  (void)hipFree(d_lens);
  (void)hipFree(d_coff);
  (void)hipFree(d_boff);
  (void)hipFree(d_logm);
}

template void launch_ring_synth<double>(owmi_plan_impl *, const double2 *, double *, long long,
                                        int, long long, int, hipStream_t);
template void launch_ring_synth<float>(owmi_plan_impl *, const float2 *, float *, long long, int,
                                       long long, int, hipStream_t);
template void launch_ring_adjoint<double>(owmi_plan_impl *, const double *, long long, double2 *,
                                          int, long long, int, hipStream_t);
template void launch_ring_adjoint<float>(owmi_plan_impl *, const float *, long long, float2 *, int,
                                         long long, int, hipStream_t);
template void launch_ring_sandwich<double>(owmi_plan_impl *, const double2 *, double2 *,
                                           const double *, long long, int, long long, hipStream_t);
template void launch_ring_sandwich<float>(owmi_plan_impl *, const float2 *, float2 *,
                                          const float *, long long, int, long long, hipStream_t);

}  // namespace owmi
