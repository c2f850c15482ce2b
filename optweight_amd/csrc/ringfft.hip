// Ring Fourier transforms between ring coefficients and ring-major real
// maps. The synthesis reads the coefficients in the pair-major order of
// coef_syn_at (as the Legendre synthesis writes them); the analysis
// (adjoint and sandwich modes) writes them for the Legendre adjoint in
// m-blocked pair-major order (coef_pair_at): a workgroup holding the two
// rings of one pair writes whole 64-byte segments, and the adjoint's lanes
// (one pair each) read both rings of their pair with one access.
//
// One workgroup transforms two real rings of equal length N packed as one
// complex sequence (z = a + i b). Any N is handled with Bluestein's
// algorithm: DFT_N(x)_k = w_k sum_j (x_j w_j) conj(w_{k-j}), w_n =
// exp(-i pi n^2 / N), evaluated as a cyclic convolution of length
// M = 2^ceil(log2(2N - 1)) (at least 16) with radix-16 Stockham FFTs whose
// data stay in registers (16 elements per thread) between passes through
// LDS. The kernel spectrum FFT_M(conj w) / M is precomputed per distinct N.
// The sandwich mode runs synthesis, the pixel multiply and the analysis of
// one ring pair without leaving the workgroup.
//
// Frequencies m > N/2 alias onto bin m mod N (reduced rings of equal-area
// grids), and the first-pixel longitude phi0 enters as exp(i m phi0).
#include "sht_internal.hpp"

#include <algorithm>
#include <cstdlib>

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

// Folded, phase-shifted Hermitian spectrum bin k of one ring:
// Z_k = sum_{m = k mod N} C_m + sum_{m >= 1, m = -k mod N} conj(C_m),
// C_0 = Re F_0, C_m = F_m exp(i m phi0).
template <typename T>
__device__ __attribute__((noinline)) cplx_t<T> fold_bin(const cplx_t<T> *F, long long fs, long long k, long long N, int mmax, T phi0) {
  cplx_t<T> acc = {0, 0};
  for (long long m = k; m <= mmax; m += N) {
    cplx_t<T> c = F[m * fs];
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
    cplx_t<T> c = F[m * fs];
    if (phi0 != (T)0) {
      T sn, cs;
      sincos_t((T)m * phi0, &sn, &cs);
      c = cmul<T>(c, cplx_t<T>{cs, sn});
    }
    acc = cadd<T>(acc, cconj<T>(c));
  }
  return acc;
}

// fold_bin of a ring that resolves every m (N > 2 mmax) with phi0 = 0 (the
// Gauss-Legendre and Clenshaw-Curtis grids at nphi >= 2 lmax + 1): bin k
// holds C_k for k <= mmax, conj C_{N-k} for k >= N - mmax, zero between.
template <typename T>
__device__ __forceinline__ cplx_t<T> plain_bin(const cplx_t<T> *F, long long fs, long long k,
                                               long long N, int mmax) {
  if (k <= mmax) {
    cplx_t<T> c = F[k * fs];
    if (k == 0) c.y = 0;
    return c;
  }
  if (N - k <= mmax) return cconj<T>(F[(N - k) * fs]);
  return {0, 0};
}

// plain_bin of the two rings of one pair, side by side at F (one 16- or
// 32-byte load per bin).
template <typename T>
__device__ __forceinline__ void plain_bin2(const cplx_t<T> *F, long long fs, long long k,
                                           long long N, int mmax, cplx_t<T> &za,
                                           cplx_t<T> &zb) {
  typedef T T4 __attribute__((ext_vector_type(4)));
  za = zb = {0, 0};
  if (k <= mmax) {
    const T4 v = *reinterpret_cast<const T4 *>(F + k * fs);
    za = {v.x, k == 0 ? (T)0 : v.y};
    zb = {v.z, k == 0 ? (T)0 : v.w};
  } else if (N - k <= mmax) {
    const T4 v = *reinterpret_cast<const T4 *>(F + (N - k) * fs);
    za = {v.x, -v.y};
    zb = {v.z, -v.w};
  }
}

// Workgroup barrier for the FFT buffer: LDS, or (GB) a global-memory
// region private to the workgroup for rings too long for LDS, where the
// barrier is wrapped in device-scope release / acquire fences so that the
// other waves' stores are visible.
template <bool GB>
__device__ __forceinline__ void block_sync() {
  if (GB) __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
  __syncthreads();
  if (GB) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
}

// ------------------------------------------------------------------ FFT
//
// Stockham autosort FFT of length M = 2^logm (logm >= 4) held by Q = M/16
// threads, thread q owning the 16 elements q + s Q (s < 16) in registers
// on entry and on exit. Passes: a first pass of radix 2^(logm mod 4) (if
// nonzero) then radix-16 passes; pass p with radix R and span Ns applies
// the twiddles W_{Ns R}^{(j mod Ns) r} to the inputs of butterfly j, a
// radix-R DFT in registers, and stores to (j / Ns) Ns R + (j mod Ns) + r Ns
// (Govindaraju et al. 2008). Between passes the data go through LDS
// (padded by one element per 16 against bank conflicts); the last
// radix-16 pass leaves its outputs in the thread's own slots.
// DIR = -1: exp(-2 pi i jk / M); DIR = +1: exp(+2 pi i jk / M), unscaled.

__device__ __forceinline__ int pad16(int i) { return i + (i >> 4); }

template <typename T, int DIR>
__device__ __forceinline__ cplx_t<T> rot_mi(cplx_t<T> a) {
  // a * (DIR i): forward -i, inverse +i
  return DIR < 0 ? cplx_t<T>{a.y, -a.x} : cplx_t<T>{-a.y, a.x};
}

// In-register DFT of the R values v[o], v[o + st], ..., v[o + (R-1) st].
template <typename T, int DIR, int R>
struct Dft;

template <typename T, int DIR>
struct Dft<T, DIR, 2> {
  __device__ static __forceinline__ void run(cplx_t<T> *v, int o, int st) {
    const cplx_t<T> a = v[o], b = v[o + st];
    v[o] = cadd<T>(a, b);
    v[o + st] = csub<T>(a, b);
  }
};

template <typename T, int DIR>
struct Dft<T, DIR, 4> {
  __device__ static __forceinline__ void run(cplx_t<T> *v, int o, int st) {
    const cplx_t<T> x0 = v[o], x1 = v[o + st], x2 = v[o + 2 * st], x3 = v[o + 3 * st];
    const cplx_t<T> a0 = cadd<T>(x0, x2), a1 = csub<T>(x0, x2);
    const cplx_t<T> a2 = cadd<T>(x1, x3), a3 = rot_mi<T, DIR>(csub<T>(x1, x3));
    v[o] = cadd<T>(a0, a2);
    v[o + 2 * st] = csub<T>(a0, a2);
    v[o + st] = cadd<T>(a1, a3);
    v[o + 3 * st] = csub<T>(a1, a3);
  }
};

// W_R^q for R in {8, 16} with sign DIR (q < R).
template <typename T, int DIR>
__device__ __forceinline__ cplx_t<T> wr16(int q) {
  // cos / sin of 2 pi q / 16
  constexpr T c1 = (T)0.92387953251128675613, s1 = (T)0.38268343236508977173;
  constexpr T h = (T)0.70710678118654752440;
  const T cs[16] = {1, c1, h, s1, 0, -s1, -h, -c1, -1, -c1, -h, -s1, 0, s1, h, c1};
  const T sn[16] = {0, s1, h, c1, 1, c1, h, s1, 0, -s1, -h, -c1, -1, -c1, -h, -s1};
  return {cs[q & 15], (T)DIR * sn[q & 15]};
}

// R = R1 R2 with R1 = 4: input n = R2 n1 + n2, output k = k1 + 4 k2.
template <typename T, int DIR, int R2>
__device__ __forceinline__ void dft4xr2(cplx_t<T> *v, int o, int st) {
  constexpr int R = 4 * R2;
  cplx_t<T> y[R];
#pragma unroll
  for (int n = 0; n < R; ++n) y[n] = v[o + n * st];
  // inner radix-4 DFTs over n1 for each n2 (stride R2 in y)
#pragma unroll
  for (int n2 = 0; n2 < R2; ++n2) Dft<T, DIR, 4>::run(y, n2, R2);
  // now y[R2 k1 + n2]; twiddle W_R^{n2 k1}
#pragma unroll
  for (int n2 = 1; n2 < R2; ++n2)
#pragma unroll
    for (int k1 = 1; k1 < 4; ++k1)
      y[R2 * k1 + n2] = cmul<T>(y[R2 * k1 + n2], wr16<T, DIR>((16 / R) * n2 * k1));
  // outer radix-R2 DFTs over n2 for each k1 (contiguous in y)
#pragma unroll
  for (int k1 = 0; k1 < 4; ++k1) Dft<T, DIR, R2>::run(y, R2 * k1, 1);
  // y[R2 k1 + k2] holds X[k1 + 4 k2]
#pragma unroll
  for (int k1 = 0; k1 < 4; ++k1)
#pragma unroll
    for (int k2 = 0; k2 < R2; ++k2) v[o + (k1 + 4 * k2) * st] = y[R2 * k1 + k2];
}

template <typename T, int DIR>
struct Dft<T, DIR, 8> {
  __device__ static __forceinline__ void run(cplx_t<T> *v, int o, int st) {
    dft4xr2<T, DIR, 2>(v, o, st);
  }
};

template <typename T, int DIR>
struct Dft<T, DIR, 16> {
  __device__ static __forceinline__ void run(cplx_t<T> *v, int o, int st) {
    dft4xr2<T, DIR, 4>(v, o, st);
  }
};

// First pass of radix R (< 16): butterfly j = q + u Q (u < 16 / R) takes
// v[u + r (16 / R)], Ns = 1, stores at j R + r.
template <typename T, int DIR, int R>
__device__ __forceinline__ void first_pass_small(cplx_t<T> *v, cplx_t<T> *buf, int q, int Q) {
  constexpr int U = 16 / R;
#pragma unroll
  for (int u = 0; u < U; ++u) {
    cplx_t<T> y[R];
#pragma unroll
    for (int r = 0; r < R; ++r) y[r] = v[u + r * U];
    Dft<T, DIR, R>::run(y, 0, 1);
    const int j = q + u * Q;
#pragma unroll
    for (int r = 0; r < R; ++r) buf[pad16(j * R + r)] = y[r];
  }
}

// Twiddles W^j = exp(-2 pi i j / Mmax), j < Mmax, from two workgroup
// tables in LDS: lo[j mod 2^hb] * hi[j >> hb] (one complex product of two
// correctly rounded values). Per-lane reads of a global table of Mmax
// entries at scattered indices measured 35-40 % of the sandwich kernel's
// time at Mmax 4096.
constexpr int kTwTab = 128;  // entries per table: Mmax <= 2^14

template <typename T>
struct Twiddles {
  const cplx_t<T> *lo;
  const cplx_t<T> *hi;
  int hb;
  __device__ __forceinline__ cplx_t<T> operator()(int j) const {
    return cmul<T>(lo[j & ((1 << hb) - 1)], hi[j >> hb]);
  }
};

// Twiddle W_{16 Ns}^{r k} of the radix-16 pass with Ns = 2^lns.
template <typename T>
__device__ __forceinline__ cplx_t<T> pass_twiddle(const Twiddles<T> &tw, int lns, int log_mmax,
                                                  int r, int k) {
  return tw(r * (k << (log_mmax - lns - 4)));  // W_Mmax^(r k Mmax / (16 Ns))
}

// Direct tables for FFTs of length 4096 (passes Ns = 16 and 256): entries
// of W_256 and W_4096 read as they are, no product.
template <typename T>
struct Twiddles4096 {
  const cplx_t<T> *t256;
  const cplx_t<T> *t4096;
};

template <typename T>
__device__ __forceinline__ cplx_t<T> pass_twiddle(const Twiddles4096<T> &tw, int lns, int, int r,
                                                  int k) {
  return lns == 4 ? tw.t256[r * k] : tw.t4096[r * k];
}

// Fill the tables of a Twiddles from the plan's table tw (Mmax entries);
// the caller's next barrier makes them visible.
template <typename T>
__device__ __forceinline__ Twiddles<T> load_twiddles(const cplx_t<T> *tw, int log_mmax,
                                                     cplx_t<T> *lo, cplx_t<T> *hi) {
  const int hb = (log_mmax + 1) >> 1;
  for (int i = threadIdx.x; i < (1 << hb); i += blockDim.x) lo[i] = tw[i];
  for (int i = threadIdx.x; i < (1 << (log_mmax - hb)); i += blockDim.x) hi[i] = tw[i << hb];
  return Twiddles<T>{lo, hi, hb};
}

// Full FFT of the thread's 16 slots (see above). All threads of the block
// call it (barriers); threads with q >= Q pass through. LOGM > 0 fixes the
// length at compile time (logm is then ignored): the passes unroll and the
// slot offsets fold into the LDS instructions. For a stride that is a
// multiple of 16 the padded index is linear: pad16(b + r s) = pad16(b) +
// r (s + s / 16).
template <typename T, int DIR, bool GB, int LOGM = 0, typename TW = Twiddles<T>>
__device__ __forceinline__ void fft16(cplx_t<T> *v, cplx_t<T> *buf, int logm_rt, int q,
                                      const TW &tw, int log_mmax) {
  const int logm = LOGM > 0 ? LOGM : logm_rt;
  const int Q = 1 << (logm - 4);
  const bool on = q < Q;
  const int r0 = logm & 3;
  block_sync<GB>();  // the previous user of buf has finished reading it
  if (r0 == 0) {
    if (logm == 4) {
      if (on) Dft<T, DIR, 16>::run(v, 0, 1);
      return;
    }
    if (on) {
      Dft<T, DIR, 16>::run(v, 0, 1);
#pragma unroll
      for (int r = 0; r < 16; ++r) buf[17 * q + r] = v[r];  // pad16(16 q + r)
    }
  } else {
    if (on) {
      if (r0 == 1)
        first_pass_small<T, DIR, 2>(v, buf, q, Q);
      else if (r0 == 2)
        first_pass_small<T, DIR, 4>(v, buf, q, Q);
      else
        first_pass_small<T, DIR, 8>(v, buf, q, Q);
    }
  }
  // radix-16 passes; the last one keeps its outputs in registers
  const int lns0 = r0 == 0 ? 4 : r0;  // log2 Ns of the first of them
  const int npass = (logm - lns0) >> 2;
#pragma unroll
  for (int pass = 0; pass < npass; ++pass) {
    const int lns = lns0 + 4 * pass;
    block_sync<GB>();
    const bool last = pass == npass - 1;
    const int k = q & ((1 << lns) - 1);
    if (on) {
      if (Q >= 16) {
        const cplx_t<T> *src = buf + pad16(q);
#pragma unroll
        for (int r = 0; r < 16; ++r) v[r] = src[r * (Q + (Q >> 4))];
      } else {
#pragma unroll
        for (int r = 0; r < 16; ++r) v[r] = buf[pad16(q + r * Q)];
      }
      if (k != 0) {
#pragma unroll
        for (int r = 1; r < 16; ++r) {
          cplx_t<T> w = pass_twiddle(tw, lns, log_mmax, r, k);
          if (DIR > 0) w.y = -w.y;
          v[r] = cmul<T>(v[r], w);
        }
      }
      Dft<T, DIR, 16>::run(v, 0, 1);
    }
    if (last) return;
    block_sync<GB>();
    if (on) {
      const int base = ((q >> lns) << (lns + 4)) + k;
      if (lns >= 4) {
        cplx_t<T> *dst = buf + pad16(base);
        const int st = (1 << lns) + (1 << (lns - 4));
#pragma unroll
        for (int r = 0; r < 16; ++r) dst[r * st] = v[r];
      } else {
#pragma unroll
        for (int r = 0; r < 16; ++r) buf[pad16(base + (r << lns))] = v[r];
      }
    }
  }
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
  const int *ring_slot;  // 2 pair + side of each ring
  long long nrings;
  long long npairs;
  long long npix;
  int mmax;
  int log_mmax_fft;
  int ncomp;
  int weighted;
};

// Cyclic convolution with the ring's Bluestein kernel, in the thread's
// slots: v <- IDFT_M(DFT_M(v) * bhat) (bhat includes 1/M).
template <typename T, bool GB, int LOGM = 0>
__device__ __forceinline__ void bluestein_conv(cplx_t<T> *v, cplx_t<T> *buf, int logm, int q,
                                               const Twiddles<T> &tw, int log_mmax,
                                               const cplx_t<T> *bhat) {
  const int Q = 1 << (logm - 4);
  fft16<T, -1, GB, LOGM>(v, buf, logm, q, tw, log_mmax);
  if (q < Q) {
#pragma unroll
    for (int s = 0; s < 16; ++s) v[s] = cmul<T>(v[s], bhat[q + s * Q]);
  }
  fft16<T, 1, GB, LOGM>(v, buf, logm, q, tw, log_mmax);
}

// MAXT: the launch's workgroup size bound (M / 16 threads rounded up to
// 256, 512 or 1024), which sets the register budget per thread. GB: the
// FFT buffer of each workgroup is a region of gbuf (region elements) in
// global memory instead of LDS. LOGM > 0: every ring of the plan has the
// Bluestein length 2^LOGM, fixed at compile time.
// The float32 synthesis mode in 256-thread workgroups is held to 128
// registers: four workgroups per CU (the LDS bound) instead of three, 9 %
// faster at Bluestein length 4096 (more waves to fill the issue slots).
template <typename T, int MODE, int MAXT, bool GB, int LOGM = 0>
__global__ void __launch_bounds__(MAXT)
__attribute__((amdgpu_waves_per_eu(MAXT == 256 && sizeof(T) == 4 && MODE == 0 ? 4 : 1)))
ring_kernel(RingArgs a, const cplx_t<T> *twiddle, const cplx_t<T> *chirp,
            const cplx_t<T> *bhat_all, const T *ringw, const T *phi0s, const cplx_t<T> *Gin,
            cplx_t<T> *Gout, const T *map_in, T *map_out, long long map_ts, const T *diag,
            long long diag_ts, cplx_t<T> *gbuf, long long region) {
  extern __shared__ unsigned char smem_raw[];
  cplx_t<T> *buf = GB ? gbuf + ((long long)blockIdx.y * gridDim.x + blockIdx.x) * region
                      : reinterpret_cast<cplx_t<T> *>(smem_raw);
  __shared__ cplx_t<T> tw_lo[kTwTab], tw_hi[kTwTab];
  const Twiddles<T> tws = load_twiddles<T>(twiddle, a.log_mmax_fft, tw_lo, tw_hi);

  // modes that read the pair-major synthesis coefficients: workgroups that
  // run on one XCD (dispatched round-robin by linear index) take a
  // contiguous range of jobs, so neighbouring pairs share cache lines in
  // one L2 (measured: alm2map faster by 12 % at TQU lmax 1023, batch 128;
  // the adjoint mode, which reads maps only, slower by 7 % with it)
  const long long gx = gridDim.x;
  const long long job = (MODE != kAdjoint && (gx & 7) == 0)
                            ? (long long)(blockIdx.x & 7) * (gx >> 3) + (blockIdx.x >> 3)
                            : (long long)blockIdx.x;
  const long long tc = blockIdx.y;  // transform * ncomp + comp
  const long long t = tc / a.ncomp;
  const int c = (int)(tc % a.ncomp);
  const int ra = a.fft_a[job];
  const int rb = a.fft_b[job];
  const long long N = a.nphi[ra];
  const int logm = LOGM > 0 ? LOGM : a.ring_logm[ra];
  const int Q = 1 << (logm - 4);
  const int q = threadIdx.x;
  const bool on = q < Q;
  const cplx_t<T> *w = chirp + a.chirp_off[ra];
  const cplx_t<T> *bhat = bhat_all + a.bhat_off[ra];
  const int mmax = a.mmax;
  // slots of the two rings in the pair-major coefficients; the two rings
  // of a mirror pair sit side by side
  const int sa = a.ring_slot[ra], sb = rb >= 0 ? a.ring_slot[rb] : 0;
  const bool joint = rb >= 0 && sb == sa + 1 && (sa & 1) == 0;
  const T pa = phi0s[ra];
  const T pb = rb >= 0 ? phi0s[rb] : (T)0;
  const long long off_a = a.offsets[ra];
  const long long off_b = rb >= 0 ? a.offsets[rb] : 0;

  // slot s of thread q holds index k = q + s Q of the length-M sequences
  cplx_t<T> v[16];

  if (MODE == kSynth || MODE == kSandwich) {
    // synthesis layout (coef_syn_at): m stride 2 npairs elements
    const long long fs = 2 * a.npairs;
    const cplx_t<T> *Fa = Gin + coef_syn_at(tc, 0, sa >> 1, mmax, a.npairs) + (sa & 1);
    const cplx_t<T> *Fb = Gin + coef_syn_at(tc, 0, sb >> 1, mmax, a.npairs) + (sb & 1);
    // Inverse DFT of Z = Za + i Zb through conj: x = conj(DFT(conj Z)).
    const bool plain = N > 2LL * mmax && pa == (T)0 && pb == (T)0;
    if (on) {
#pragma unroll
      for (int s = 0; s < 16; ++s) {
        const long long k = q + s * Q;
        cplx_t<T> val = {0, 0};
        if (k < N) {
          cplx_t<T> za, zb = {0, 0};
          if (plain && joint) {
            plain_bin2<T>(Fa, fs, k, N, mmax, za, zb);
          } else if (plain) {
            za = plain_bin<T>(Fa, fs, k, N, mmax);
            if (rb >= 0) zb = plain_bin<T>(Fb, fs, k, N, mmax);
          } else {
            za = fold_bin<T>(Fa, fs, k, N, mmax, pa);
            if (rb >= 0) zb = fold_bin<T>(Fb, fs, k, N, mmax, pb);
          }
          const cplx_t<T> Z = {za.x - zb.y, za.y + zb.x};
          val = cmul<T>(cconj<T>(Z), w[k]);
        }
        v[s] = val;
      }
    }
    bluestein_conv<T, GB, LOGM>(v, buf, logm, q, tws, a.log_mmax_fft, bhat);
    // DFT(conj Z)_j = w_j c_j and x = conj of it: ring a gets Re(w_j c_j),
    // ring b gets -Im(w_j c_j).
    if (on) {
      T *row = map_out + t * map_ts + (long long)c * a.npix;
      const T *drow = diag + t * diag_ts + (long long)c * a.npix;
#pragma unroll
      for (int s = 0; s < 16; ++s) {
        const long long j = q + s * Q;
        if (j < N) {
          const cplx_t<T> o = cmul<T>(w[j], v[s]);
          T va = o.x;
          T vb = -o.y;
          if (MODE == kSynth) {
            if (a.weighted) {
              va *= ringw[ra];
              if (rb >= 0) vb *= ringw[rb];
            }
            row[off_a + j] = va;
            if (rb >= 0) row[off_b + j] = vb;
          } else {
            // pixel multiply, then the forward-DFT input z'_j w_j
            va *= drow[off_a + j];
            vb = rb >= 0 ? vb * drow[off_b + j] : (T)0;
            v[s] = cmul<T>(cplx_t<T>{va, vb}, w[j]);
          }
        } else if (MODE == kSandwich) {
          v[s] = {0, 0};
        }
      }
    }
  }

  if (MODE == kAdjoint) {
    if (on) {
      const T *row = map_in + t * map_ts + (long long)c * a.npix;
#pragma unroll
      for (int s = 0; s < 16; ++s) {
        const long long j = q + s * Q;
        cplx_t<T> val = {0, 0};
        if (j < N) {
          T va = row[off_a + j];
          T vb = rb >= 0 ? row[off_b + j] : (T)0;
          if (a.weighted) {
            va *= ringw[ra];
            if (rb >= 0) vb *= ringw[rb];
          }
          val = cmul<T>(cplx_t<T>{va, vb}, w[j]);
        }
        v[s] = val;
      }
    }
  }

  if (MODE == kAdjoint || MODE == kSandwich) {
    bluestein_conv<T, GB, LOGM>(v, buf, logm, q, tws, a.log_mmax_fft, bhat);
    // X_k = w_k c_k, gathered in the buffer (natural order) for the Hermitian split
    block_sync<GB>();
    if (on) {
#pragma unroll
      for (int s = 0; s < 16; ++s) {
        const long long k = q + s * Q;
        if (k < N) buf[pad16((int)k)] = cmul<T>(w[k], v[s]);
      }
    }
    block_sync<GB>();
    // m-blocked pair-major output (coef_pair_at); the two rings of a mirror
    // pair: one 16-byte (float) or 32-byte (double) store per m
    cplx_t<T> *Ga = Gout + coef_pair_at<T>(tc, 0, sa >> 1, mmax, a.npairs) + (sa & 1);
    cplx_t<T> *Gb = Gout + coef_pair_at<T>(tc, 0, sb >> 1, mmax, a.npairs) + (sb & 1);
    constexpr int MB = coef_mb<T>();
    const long long gp_blk = 2 * MB * a.npairs;  // elements per m block
    typedef T T4 __attribute__((ext_vector_type(4)));
    for (int m = threadIdx.x; m <= mmax; m += blockDim.x) {
      const long long k = m % N;
      const long long kk = (N - k) % N;
      const cplx_t<T> X = buf[pad16((int)k)];
      const cplx_t<T> Xc = cconj<T>(buf[pad16((int)kk)]);
      cplx_t<T> xa = {(T)0.5 * (X.x + Xc.x), (T)0.5 * (X.y + Xc.y)};
      const cplx_t<T> d = {(T)0.5 * (X.x - Xc.x), (T)0.5 * (X.y - Xc.y)};
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
      if (rb >= 0 && m > 0 && pb != (T)0) {
        T sn, cs;
        sincos_t(-(T)m * pb, &sn, &cs);
        xb = cmul<T>(xb, cplx_t<T>{cs, sn});
      }
      const long long om = (m / MB) * gp_blk + 2 * (m % MB);
      if (joint) {
        *reinterpret_cast<T4 *>(Ga + om) = T4{xa.x, xa.y, xb.x, xb.y};
      } else {
        Ga[om] = xa;
        if (rb >= 0) Gb[om] = xb;
      }
    }
  }
}

// ------------------------------------------------------ spectral sandwich
//
// The pixel sandwich of a fixed pixel diagonal as a convolution over m.
// Ring a with spectrum C_m (Hermitian-extended to |m| <= mmax) and pixel
// diagonal D_j gives H_m = sum_j D_j x_j exp(-i m phi_j)
//                        = sum_m' K(m - m') C_m',
// K(d) = sum_j D_j exp(-i d phi_j), |d| <= 2 mmax: a linear convolution,
// exact as a cyclic one of length M >= 4 mmax + 1 with C placed at index
// m mod M. Two rings a, b packed as Z = C_a + i C_b: the M-point DFT of a
// Hermitian sequence is real, so DFT(Z) = DFT(C_a) + i DFT(C_b) splits into
// its real and imaginary parts, and the sandwich is
//   F = IDFT_M(Re DFT_M(Z) * S_a + i Im DFT_M(Z) * S_b),
//   S_r[k] = DFT_M(K_r)[k] / M (real),
// H_a,m = (F_m + conj F_-m) / 2, H_b,m = (F_m - conj F_-m) / (2 i): two FFTs
// of length M per ring pair instead of four Bluestein FFTs, no chirps, any
// ring length and phi0. The multipliers S_r depend on the diagonal only and
// are computed once (conv_mult_kernel).

// One workgroup per (ring pair, component, group of TPB transforms): the
// multipliers are read once per group (TPB 8 for workgroups of up to 256
// threads; 1 beyond, where the loop's registers would spill). LOGM > 0: the
// FFT length fixed at compile time (logm_rt ignored).
template <typename T, int MAXT, bool GB, int TPB, int LOGM = 0>
__global__ void __launch_bounds__(MAXT)
ring_conv_kernel(RingArgs a, int logm_rt, const cplx_t<T> *twiddle, const T *mult,
                 long long mult_cs, const cplx_t<T> *Gin, cplx_t<T> *Gout, long long ntrans,
                 cplx_t<T> *gbuf, long long region) {
  extern __shared__ unsigned char smem_raw[];
  cplx_t<T> *buf = GB ? gbuf + ((long long)blockIdx.y * gridDim.x + blockIdx.x) * region
                      : reinterpret_cast<cplx_t<T> *>(smem_raw);
  __shared__ cplx_t<T> tw_lo[kTwTab], tw_hi[kTwTab];
  const Twiddles<T> tws = load_twiddles<T>(twiddle, a.log_mmax_fft, tw_lo, tw_hi);
  // float32 at M = 4096: direct twiddle tables (34 KiB of LDS; the
  // workgroups per CU are bound by registers, not LDS, at this size)
  constexpr bool DIRECT = LOGM == 12 && sizeof(T) == 4;
  __shared__ __align__(16) cplx_t<T> tdir[DIRECT ? 4096 + 256 : 1];
  if constexpr (DIRECT) {
    const int sh = a.log_mmax_fft - 12;
    for (int i = threadIdx.x; i < 4096; i += blockDim.x) tdir[i] = twiddle[i << sh];
    for (int i = threadIdx.x; i < 256; i += blockDim.x) tdir[4096 + i] = twiddle[i << (sh + 4)];
  }
  const Twiddles4096<T> twd = {tdir + 4096, tdir};
  // workgroups of one XCD take a contiguous range of ring pairs (they share
  // the cache lines of the pair-major coefficients)
  const long long gx = gridDim.x;
  const long long job = (gx & 7) == 0 ? (long long)(blockIdx.x & 7) * (gx >> 3) + (blockIdx.x >> 3)
                                      : (long long)blockIdx.x;
  const int c = (int)(blockIdx.y % a.ncomp);
  const long long t_lo = (long long)(blockIdx.y / a.ncomp) * TPB;
  const long long t_hi = min(ntrans, t_lo + TPB);
  const int ra = a.fft_a[job];
  const int rb = a.fft_b[job];
  const int logm = LOGM > 0 ? LOGM : logm_rt;
  const int M = 1 << logm;
  const int Q = M >> 4;
  const int q = threadIdx.x;
  const bool on = q < Q;
  const int mmax = a.mmax;
  const int sa = a.ring_slot[ra], sb = rb >= 0 ? a.ring_slot[rb] : 0;
  const bool joint = rb >= 0 && sb == sa + 1 && (sa & 1) == 0;
  // The multipliers stay in registers across the group where the register
  // budget allows (float, 256 threads), else they are read (from L2 after
  // the first transform) at the multiply.
  constexpr bool KEEP = sizeof(T) == 4 && TPB > 1;
  const T *pma = mult + c * mult_cs + (long long)ra * M;
  const T *pmb = mult + c * mult_cs + (long long)(rb >= 0 ? rb : ra) * M;
  T ma[16], mb[16];
  auto load_mult = [&]() {
#pragma unroll
    for (int s = 0; s < 16; ++s) {
      ma[s] = pma[q + s * Q];
      mb[s] = rb >= 0 ? pmb[q + s * Q] : (T)0;
    }
  };
  if (KEEP && on) load_mult();
  const long long fs = 2 * a.npairs;  // m stride of the synthesis layout
  constexpr int MB = coef_mb<T>();
  const long long gp_blk = 2 * MB * a.npairs;  // elements per m block of the adjoint layout
  typedef T T4 __attribute__((ext_vector_type(4)));
  // slot k holds m = k (k <= mmax) or m = k - M (k >= M - mmax), zero
  // between; the loads of all 16 slots are issued together (from a clamped,
  // valid m) and the slot's value selected afterwards. float32 with the
  // length fixed (LOGM): the next transform's loads are issued before this
  // one's FFTs, which then hide their latency.
  T4 w[16];
  auto issue = [&](long long t) {
    const long long tc = t * a.ncomp + c;
    const cplx_t<T> *Fa = Gin + coef_syn_at(tc, 0, sa >> 1, mmax, a.npairs) + (sa & 1);
    const cplx_t<T> *Fb = Gin + coef_syn_at(tc, 0, sb >> 1, mmax, a.npairs) + (sb & 1);
    if (!on) return;
    if (joint) {
#pragma unroll
      for (int s = 0; s < 16; ++s) {
        const int k = q + s * Q;
        const int am = k <= mmax ? k : (k >= M - mmax ? M - k : 0);
        w[s] = *reinterpret_cast<const T4 *>(Fa + (long long)am * fs);
      }
    } else {
#pragma unroll
      for (int s = 0; s < 16; ++s) {
        const int k = q + s * Q;
        const int am = k <= mmax ? k : (k >= M - mmax ? M - k : 0);
        const cplx_t<T> fa = Fa[(long long)am * fs];
        const cplx_t<T> fb = rb >= 0 ? Fb[(long long)am * fs] : cplx_t<T>{0, 0};
        w[s] = T4{fa.x, fa.y, fb.x, fb.y};
      }
    }
  };
  constexpr bool AHEAD = LOGM > 0 && sizeof(T) == 4;
  if (AHEAD) issue(t_lo);
  for (long long t = t_lo; t < t_hi; ++t) {
    const long long tc = t * a.ncomp + c;
    cplx_t<T> v[16];
    if (!AHEAD) issue(t);
    if (on) {
#pragma unroll
      for (int s = 0; s < 16; ++s) {
        const int k = q + s * Q;
        const bool in = k <= mmax || k >= M - mmax;
        // C_0 real; C_-m = conj C_m
        const T sg = k == 0 ? (T)0 : (k > mmax ? (T)-1 : (T)1);
        const cplx_t<T> za = {in ? w[s].x : (T)0, in ? sg * w[s].y : (T)0};
        const cplx_t<T> zb = {in ? w[s].z : (T)0, in ? sg * w[s].w : (T)0};
        v[s] = {za.x - zb.y, za.y + zb.x};
      }
    }
    if (AHEAD && t + 1 < t_hi) issue(t + 1);
    if constexpr (DIRECT)
      fft16<T, -1, GB, LOGM>(v, buf, logm, q, twd, a.log_mmax_fft);
    else
      fft16<T, -1, GB, LOGM>(v, buf, logm, q, tws, a.log_mmax_fft);
    if (on) {
      if (!KEEP) load_mult();
#pragma unroll
      for (int s = 0; s < 16; ++s) v[s] = {v[s].x * ma[s], v[s].y * mb[s]};
    }
    if constexpr (DIRECT)
      fft16<T, 1, GB, LOGM>(v, buf, logm, q, twd, a.log_mmax_fft);
    else
      fft16<T, 1, GB, LOGM>(v, buf, logm, q, tws, a.log_mmax_fft);
    block_sync<GB>();
    if (on) {
      // pad16(q + s Q) = pad16(q) + s (Q + Q / 16) for Q >= 16
      cplx_t<T> *dst = buf + pad16(q);
#pragma unroll
      for (int s = 0; s < 16; ++s) {
        const int k = q + s * Q;
        if (k <= mmax || k >= M - mmax) {
          if (Q >= 16)
            dst[s * (Q + (Q >> 4))] = v[s];
          else
            buf[pad16(k)] = v[s];
        }
      }
    }
    block_sync<GB>();
    cplx_t<T> *Ga = Gout + coef_pair_at<T>(tc, 0, sa >> 1, mmax, a.npairs) + (sa & 1);
    cplx_t<T> *Gb = Gout + coef_pair_at<T>(tc, 0, sb >> 1, mmax, a.npairs) + (sb & 1);
    for (int m = threadIdx.x; m <= mmax; m += blockDim.x) {
      const cplx_t<T> X = buf[pad16(m)];
      const cplx_t<T> Xc = cconj<T>(buf[pad16(m == 0 ? 0 : M - m)]);
      cplx_t<T> xa = {(T)0.5 * (X.x + Xc.x), (T)0.5 * (X.y + Xc.y)};
      const cplx_t<T> d = {(T)0.5 * (X.x - Xc.x), (T)0.5 * (X.y - Xc.y)};
      cplx_t<T> xb = {d.y, -d.x};  // d / i
      if (m == 0) {
        xa.y = 0;
        xb.y = 0;
      }
      const long long om = (m / MB) * gp_blk + 2 * (m % MB);
      if (joint) {
        *reinterpret_cast<T4 *>(Ga + om) = T4{xa.x, xa.y, xb.x, xb.y};
      } else {
        Ga[om] = xa;
        if (rb >= 0) Gb[om] = xb;
      }
    }
  }
}

// Multipliers of the spectral sandwich, in float64, per FFT job (rings a,
// b) and component c: the Bluestein DFT of z = D_a + i D_b gives X_k,
// k < N, and
//   K_a(d) = exp(-i d phi_a) (X_d + conj X_-d) / 2,
//   K_b(d) = exp(-i d phi_b) (X_d - conj X_-d) / (2 i)   (indices mod N),
// for d = 0 .. 2 mmax; Z = K_a + i K_b at index d mod M (|d| <= 2 mmax,
// K(-d) = conj K(d)), and S_a = Re DFT_M(Z) / M, S_b = Im DFT_M(Z) / M
// (the DFTs of the two Hermitian sequences are real). Two Bluestein FFTs and
// one FFT of length M per job and component.
template <typename T, int MAXT, bool GB>
__global__ void __launch_bounds__(MAXT)
conv_mult_kernel(RingArgs a, int logmc, const double2 *twiddle, const double2 *chirp,
                 const double2 *bhat_all, const double *phi0s, const T *diag, long long diag_cs,
                 T *mult, long long mult_cs, double2 *gbuf, long long region) {
  extern __shared__ unsigned char smem_raw[];
  double2 *buf = GB ? gbuf + ((long long)blockIdx.y * gridDim.x + blockIdx.x) * region
                    : reinterpret_cast<double2 *>(smem_raw);
  __shared__ double2 tw_lo[kTwTab], tw_hi[kTwTab];
  const Twiddles<double> tws = load_twiddles<double>(twiddle, a.log_mmax_fft, tw_lo, tw_hi);
  const long long job = blockIdx.x;
  const int c = blockIdx.y;
  const int ra = a.fft_a[job];
  const int rb = a.fft_b[job];
  const long long N = a.nphi[ra];
  const int logm = a.ring_logm[ra];
  const int Q = 1 << (logm - 4);
  const int q = threadIdx.x;
  const bool on = q < Q;
  const double2 *w = chirp + a.chirp_off[ra];
  const double2 *bhat = bhat_all + a.bhat_off[ra];
  const double pa = phi0s[ra], pb = rb >= 0 ? phi0s[rb] : 0.0;
  const T *Da = diag + c * diag_cs + a.offsets[ra];
  const T *Db = diag + c * diag_cs + (rb >= 0 ? a.offsets[rb] : a.offsets[ra]);
  double2 v[16];
  if (on) {
#pragma unroll
    for (int s = 0; s < 16; ++s) {
      const long long j = q + s * Q;
      double2 val = {0, 0};
      if (j < N) {
        const double da = (double)Da[j];
        const double db = rb >= 0 ? (double)Db[j] : 0.0;
        val = cmul<double>(double2{da, db}, w[j]);
      }
      v[s] = val;
    }
  }
  bluestein_conv<double, GB>(v, buf, logm, q, tws, a.log_mmax_fft, bhat);
  block_sync<GB>();
  if (on) {
#pragma unroll
    for (int s = 0; s < 16; ++s) {
      const long long k = q + s * Q;
      if (k < N) buf[pad16((int)k)] = cmul<double>(w[k], v[s]);
    }
  }
  block_sync<GB>();
  const int Mc = 1 << logmc;
  const int Qc = Mc >> 4;
  const bool onc = q < Qc;
  const int dmax = 2 * a.mmax;
  if (onc) {
#pragma unroll
    for (int s = 0; s < 16; ++s) {
      const int k = q + s * Qc;
      const bool in = k <= dmax || k >= Mc - dmax;
      const int ad = k <= dmax ? k : Mc - k;  // |d|
      double2 z = {0, 0};
      if (in) {
        const long long i1 = ad % N, i2 = (N - i1) % N;
        const double2 X = buf[pad16((int)i1)];
        const double2 Xc = cconj<double>(buf[pad16((int)i2)]);
        double2 ka = {0.5 * (X.x + Xc.x), 0.5 * (X.y + Xc.y)};
        double2 kb = {0.5 * (X.y - Xc.y), -0.5 * (X.x - Xc.x)};
        if (pa != 0.0) {
          double sn, cs;
          sincos(-(double)ad * pa, &sn, &cs);
          ka = cmul<double>(ka, double2{cs, sn});
        }
        if (pb != 0.0) {
          double sn, cs;
          sincos(-(double)ad * pb, &sn, &cs);
          kb = cmul<double>(kb, double2{cs, sn});
        }
        if (rb < 0) kb = {0, 0};
        if (k > dmax) {  // d < 0: K(d) = conj K(-d)
          ka.y = -ka.y;
          kb.y = -kb.y;
        }
        z = {ka.x - kb.y, ka.y + kb.x};
      }
      v[s] = z;
    }
  }
  fft16<double, -1, GB>(v, buf, logmc, q, tws, a.log_mmax_fft);
  if (onc) {
    const double inv = 1.0 / (double)Mc;
    T *oa = mult + c * mult_cs + (long long)ra * Mc;
    T *ob = mult + c * mult_cs + (long long)(rb >= 0 ? rb : ra) * Mc;
#pragma unroll
    for (int s = 0; s < 16; ++s) {
      const int k = q + s * Qc;
      oa[k] = (T)(v[s].x * inv);
      if (rb >= 0) ob[k] = (T)(v[s].y * inv);
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

// LDS of one workgroup (160 KiB on gfx950).
constexpr size_t kMaxLds = 160 * 1024;

// FFT buffer of one ring workgroup: M + M/16 (padding) complex elements.
template <typename T>
size_t ring_lds_bytes(int mmax_fft) {
  return (size_t)(mmax_fft + mmax_fft / 16) * sizeof(cplx_t<T>);
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
  a.ring_slot = p->ring_slot;
  a.nrings = p->nrings;
  a.npairs = p->npairs;
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
  const size_t lds = ring_lds_bytes<T>(p->mmax_fft);
  const int threads = std::max(64, p->mmax_fft / 16);
  OWMI_REQUIRE(threads <= 1024, "ring too long for one workgroup");
  PlanTables<T> &tb = tab<T>(p);
  dim3 grid((unsigned)p->nfft_jobs, (unsigned)nrows);
  OWMI_REQUIRE(p->log2_mmax_fft <= 14, "ring FFT length above 2^14");
  const bool gb = lds + 2 * kTwTab * sizeof(cplx_t<T>) > kMaxLds;
  cplx_t<T> *gbuf = nullptr;
  const long long region = p->mmax_fft + p->mmax_fft / 16;
  if (gb) {
    const size_t bytes = (size_t)p->nfft_jobs * nrows * region * sizeof(cplx_t<T>);
    OWMI_REQUIRE(bytes <= (size_t)32 << 30, "ring-FFT scratch above 32 GiB: split the batch");
    gbuf = static_cast<cplx_t<T> *>(p->ensure_work3(bytes));
  }
  // every ring of Bluestein length 4096 (nphi 1025 .. 2048, e.g. the
  // Gauss-Legendre rings of lmax 1023): length fixed at compile time
  auto kern = gb ? ring_kernel<T, MODE, 1024, true>
              : (threads == 256 && p->uniform_logm == 12) ? ring_kernel<T, MODE, 256, false, 12>
              : threads <= 256 ? ring_kernel<T, MODE, 256, false>
              : threads <= 512 ? ring_kernel<T, MODE, 512, false>
                               : ring_kernel<T, MODE, 1024, false>;
  const size_t dyn = gb ? 0 : lds;
  allow_lds(kern, dyn);
  hipLaunchKernelGGL(kern, grid, dim3(threads), dyn, st, ring_args<T>(p, ncomp, weighted),
                     reinterpret_cast<const cplx_t<T> *>(tb.twiddle),
                     reinterpret_cast<const cplx_t<T> *>(tb.chirp),
                     reinterpret_cast<const cplx_t<T> *>(tb.bhat), tb.ringw, tb.phi0, Gin, Gout,
                     map_in, map_out, map_ts, diag, diag_ts, gbuf, region);
}

template <typename T>
void launch_conv(owmi_plan_impl *p, const cplx_t<T> *G, cplx_t<T> *Gout, const T *mult,
                 long long mult_cs, int ncomp, long long ntrans, hipStream_t st) {
  const int logm = conv_logm(p);
  OWMI_REQUIRE(logm > 0, "the spectral sandwich does not apply to this plan");
  const int M = 1 << logm;
  // transforms per workgroup: the multipliers are read once per group
  // (OWMI_CONV_TPB = 1, 2, 4 or 8 overrides for measurements)
  const int threads = std::max(64, M / 16);
  static const int forced = [] {
    const char *e = std::getenv("OWMI_CONV_TPB");
    return e != nullptr ? std::atoi(e) : 0;
  }();
  const int tpb = threads > 256 ? 1 : (forced == 1 || forced == 2 || forced == 4) ? forced : 8;
  const long long ngroups = (ntrans + tpb - 1) / tpb;
  const long long nrows = ngroups * ncomp;
  OWMI_REQUIRE(nrows < 65536, "too many rows for one ring-FFT launch");
  const size_t lds = ring_lds_bytes<T>(M);
  PlanTables<T> &tb = tab<T>(p);
  dim3 grid((unsigned)p->nfft_jobs, (unsigned)nrows);
  const bool gb = lds + 2 * kTwTab * sizeof(cplx_t<T>) > kMaxLds;
  cplx_t<T> *gbuf = nullptr;
  const long long region = M + M / 16;
  if (gb) {
    const size_t bytes = (size_t)p->nfft_jobs * nrows * region * sizeof(cplx_t<T>);
    OWMI_REQUIRE(bytes <= (size_t)32 << 30, "ring-FFT scratch above 32 GiB: split the batch");
    gbuf = static_cast<cplx_t<T> *>(p->ensure_work3(bytes));
  }
  // M = 4096 (band limits 512 .. 1023) with its length fixed at compile time
  auto kern = gb ? ring_conv_kernel<T, 1024, true, 1>
              : threads > 512 ? ring_conv_kernel<T, 1024, false, 1>
              : threads > 256 ? ring_conv_kernel<T, 512, false, 1>
              : tpb == 1 ? ring_conv_kernel<T, 256, false, 1>
              : tpb == 2 ? ring_conv_kernel<T, 256, false, 2>
              : tpb == 4 ? ring_conv_kernel<T, 256, false, 4>
              : logm == 12 ? ring_conv_kernel<T, 256, false, 8, 12>
                           : ring_conv_kernel<T, 256, false, 8>;
  const size_t dyn = gb ? 0 : lds;
  allow_lds(kern, dyn);
  hipLaunchKernelGGL(kern, grid, dim3(threads), dyn, st, ring_args<T>(p, ncomp, 0), logm,
                     reinterpret_cast<const cplx_t<T> *>(tb.twiddle), mult, mult_cs, G, Gout,
                     ntrans, gbuf, region);
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

int conv_logm(const owmi_plan_impl *p) {
  int lm = 4;
  while ((1LL << lm) < 4LL * p->mmax + 1) ++lm;
  // the plan's twiddle table and FFT workgroup size cover M
  return (1 << lm) <= p->mmax_fft ? lm : 0;
}

template <typename T>
void launch_ring_conv(owmi_plan_impl *p, const cplx_t<T> *G, cplx_t<T> *Gout, const T *mult,
                      long long mult_cs, int ncomp, long long ntrans, hipStream_t st) {
  launch_conv<T>(p, G, Gout, mult, mult_cs, ncomp, ntrans, st);
}

template <typename T>
void launch_conv_mult(owmi_plan_impl *p, const T *diag, long long diag_cs, int ncomp, T *mult,
                      long long mult_cs, hipStream_t st) {
  const int logm = conv_logm(p);
  OWMI_REQUIRE(logm > 0, "the spectral sandwich does not apply to this plan");
  OWMI_REQUIRE(ncomp < 65536, "too many components");
  const size_t lds = ring_lds_bytes<double>(p->mmax_fft);
  const int threads = std::max(64, p->mmax_fft / 16);
  const bool gb = lds + 2 * kTwTab * sizeof(double2) > kMaxLds;
  double2 *gbuf = nullptr;
  const long long region = p->mmax_fft + p->mmax_fft / 16;
  if (gb) {
    const size_t bytes = (size_t)p->nfft_jobs * ncomp * region * sizeof(double2);
    OWMI_REQUIRE(bytes <= (size_t)32 << 30, "ring-FFT scratch above 32 GiB");
    gbuf = static_cast<double2 *>(p->ensure_work3(bytes));
  }
  auto kern = gb ? conv_mult_kernel<T, 1024, true>
              : threads <= 256 ? conv_mult_kernel<T, 256, false>
              : threads <= 512 ? conv_mult_kernel<T, 512, false>
                               : conv_mult_kernel<T, 1024, false>;
  const size_t dyn = gb ? 0 : lds;
  allow_lds(kern, dyn);
  hipLaunchKernelGGL(kern, dim3((unsigned)p->nfft_jobs, (unsigned)ncomp), dim3(threads), dyn, st,
                     ring_args<double>(p, ncomp, 0), logm,
                     reinterpret_cast<const double2 *>(p->t64.twiddle),
                     reinterpret_cast<const double2 *>(p->t64.chirp),
                     reinterpret_cast<const double2 *>(p->t64.bhat),
                     reinterpret_cast<const double *>(p->t64.phi0), diag, diag_cs, mult, mult_cs,
                     gbuf, region);
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
  int max_logm = 0;
  for (int lm : logms) max_logm = std::max(max_logm, lm);
  size_t lds = ((size_t)1 << max_logm) * sizeof(double2);
  OWMI_REQUIRE(lds <= kMaxLds, "Bluestein length exceeds LDS");
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
template void launch_ring_conv<double>(owmi_plan_impl *, const double2 *, double2 *,
                                       const double *, long long, int, long long, hipStream_t);
template void launch_ring_conv<float>(owmi_plan_impl *, const float2 *, float2 *, const float *,
                                      long long, int, long long, hipStream_t);
template void launch_conv_mult<double>(owmi_plan_impl *, const double *, long long, int, double *,
                                       long long, hipStream_t);
template void launch_conv_mult<float>(owmi_plan_impl *, const float *, long long, int, float *,
                                      long long, hipStream_t);

}  // namespace owmi
