// Streaming (HBM-bound) kernels: per-l multiply, truncation, wavelet
// accumulate, ring weights, pixel matvec, real-field alm dot product and the
// CG vector updates. One thread per complex element (or pixel), grid-stride
// over a fixed grid where a deterministic reduction order matters.
#include "common.hpp"

#include <map>
#include <mutex>
#include <tuple>
#include <vector>

namespace owmi {

static thread_local std::string g_last_error;
void set_error(const std::string &msg) { g_last_error = msg; }
void clear_error() { g_last_error.clear(); }

void *stream_scratch(int tag, size_t bytes, hipStream_t st) {
  static std::mutex mu;
  static std::map<std::tuple<int, hipStream_t, int>, std::pair<void *, size_t>> bufs;
  int dev = 0;
  OWMI_HIP(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lk(mu);
  const auto key = std::make_tuple(dev, st, tag);
  auto it = bufs.find(key);
  if (it != bufs.end()) {
    OWMI_REQUIRE(it->second.second == bytes, "stream_scratch: size changed for a tag");
    return it->second.first;
  }
  hipStreamCaptureMode mode = hipStreamCaptureModeRelaxed;
  OWMI_HIP(hipThreadExchangeStreamCaptureMode(&mode));
  void *p = nullptr;
  const hipError_t e = hipMalloc(&p, bytes);
  (void)hipThreadExchangeStreamCaptureMode(&mode);
  if (e != hipSuccess) throw Error(std::string("stream_scratch: hipMalloc: ") + hipGetErrorString(e));
  bufs[key] = {p, bytes};
  return p;
}

// (l, m) of flat index idx in a triangular lmax layout:
// idx = m (2 lmax + 1 - m) / 2 + l with m <= l <= lmax.
__device__ __forceinline__ void lm_of(long long idx, int lmax, int &l, int &m) {
  // m is the largest integer with m (2 lmax + 3 - m) / 2 <= idx.
  double b = 2.0 * lmax + 3.0;
  double disc = b * b - 8.0 * (double)idx;
  int mm = (int)((b - sqrt(disc)) * 0.5);
  if (mm < 0) mm = 0;
  if (mm > lmax) mm = lmax;
  while (mm > 0 && mstart_of(mm, lmax) + mm > idx) --mm;
  while (mm < lmax && mstart_of(mm + 1, lmax) + mm + 1 <= idx) ++mm;
  m = mm;
  l = (int)(idx - mstart_of(mm, lmax));
}

// ---------------------------------------------------------------- lmul

// One thread per (multipole, transform): the flat index i gives (l, m)
// once, and the thread applies the NC x NC matrix of its multipole to the
// NC components of its transform. A grid over the flat index (blockIdx.x)
// and the transform (blockIdx.y) keeps the resident waves on one narrow,
// advancing window of each array; a grid over m with several transforms
// per thread measured 1.4-1.8x slower at lmax 1023.
template <typename T, int NC>
__global__ void __launch_bounds__(256)
lmul_full_kernel(const T *__restrict__ lmat, const cplx_t<T> *in, cplx_t<T> *out, int lmax,
                 long long nelem, long long ntrans) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long t = blockIdx.y;
  if (i >= nelem || t >= ntrans) return;
  int l, m;
  lm_of(i, lmax, l, m);
  const int nell = lmax + 1;
  const cplx_t<T> *src = in + t * NC * nelem + i;
  cplx_t<T> v[NC];
#pragma unroll
  for (int j = 0; j < NC; ++j) v[j] = src[j * nelem];
  cplx_t<T> *dst = out + t * NC * nelem + i;
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    T re = 0, im = 0;
#pragma unroll
    for (int j = 0; j < NC; ++j) {
      const T a = lmat[(c * NC + j) * nell + l];
      re += a * v[j].x;
      im += a * v[j].y;
    }
    dst[c * nelem] = {re, im};
  }
}

// Loads of kLoadB elements are issued before their stores. in and out may
// be the same array (in place), so the compiler cannot move a load above an
// earlier store by itself; every thread reads and writes only its own
// elements, so the reordering leaves the result unchanged.
constexpr int kLoadB = 4;

template <typename T>
__global__ void lmul_full_generic_kernel(const T *__restrict__ lmat, const cplx_t<T> *in,
                                         cplx_t<T> *out, int lmax, int nc,
                                         long long nelem, long long ntrans) {
  // Used for ncomp > 4; in-place safe because each thread reads all
  // components of its element before writing any.
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  long long t = blockIdx.y;
  if (i >= nelem || t >= ntrans) return;
  int l, m;
  lm_of(i, lmax, l, m);
  const int nell = lmax + 1;
  const cplx_t<T> *src = in + t * nc * nelem + i;
  cplx_t<T> *dst = out + t * nc * nelem + i;
  cplx_t<T> v[16];
  for (int j = 0; j < nc; ++j) v[j] = src[j * nelem];
  for (int c = 0; c < nc; ++c) {
    T re = 0, im = 0;
    for (int j = 0; j < nc; ++j) {
      T a = lmat[(c * nc + j) * nell + l];
      re += a * v[j].x;
      im += a * v[j].y;
    }
    dst[c * nelem] = {re, im};
  }
}

// Diagonal lmat (nc, nell): one thread per (multipole, transform) applies
// the nc diagonal entries of its l to the nc components of its transform.
template <typename T>
__global__ void __launch_bounds__(256)
lmul_diag_kernel(const T *__restrict__ lmat, const cplx_t<T> *in, cplx_t<T> *out, int lmax,
                 int nc, long long nelem, long long ntrans) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long t = blockIdx.y;
  if (i >= nelem || t >= ntrans) return;
  int l, m;
  lm_of(i, lmax, l, m);
  const int nell = lmax + 1;
  const cplx_t<T> *src = in + t * nc * nelem + i;
  cplx_t<T> *dst = out + t * nc * nelem + i;
  int c = 0;
  for (; c + kLoadB <= nc; c += kLoadB) {
    cplx_t<T> v[kLoadB];
#pragma unroll
    for (int b = 0; b < kLoadB; ++b) v[b] = src[(c + b) * nelem];
#pragma unroll
    for (int b = 0; b < kLoadB; ++b) {
      const T a = lmat[(c + b) * nell + l];
      dst[(c + b) * nelem] = {a * v[b].x, a * v[b].y};
    }
  }
  if (nc - c == 3) {
    const cplx_t<T> v0 = src[c * nelem], v1 = src[(c + 1) * nelem], v2 = src[(c + 2) * nelem];
    const T a0 = lmat[c * nell + l], a1 = lmat[(c + 1) * nell + l], a2 = lmat[(c + 2) * nell + l];
    dst[c * nelem] = {a0 * v0.x, a0 * v0.y};
    dst[(c + 1) * nelem] = {a1 * v1.x, a1 * v1.y};
    dst[(c + 2) * nelem] = {a2 * v2.x, a2 * v2.y};
    return;
  }
  for (; c < nc; ++c) {
    const T a = lmat[c * nell + l];
    const cplx_t<T> v = src[c * nelem];
    dst[c * nelem] = {a * v.x, a * v.y};
  }
}

// out = La a + Lb b per (l, m), full (NC, NC, lmax + 1) matrices (La absent:
// the identity). a and b may alias out: every thread reads all components
// of both inputs at its element before writing.
template <typename T, int NC, bool HASA>
__global__ void __launch_bounds__(256)
lmul_add_kernel(const T *__restrict__ la, const cplx_t<T> *a, const T *__restrict__ lb,
                const cplx_t<T> *b, cplx_t<T> *out, int lmax, long long nelem,
                long long ntrans) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long t = blockIdx.y;
  if (i >= nelem || t >= ntrans) return;
  int l, m;
  lm_of(i, lmax, l, m);
  const int nell = lmax + 1;
  const long long base = t * NC * nelem + i;
  cplx_t<T> va[NC], vb[NC];
#pragma unroll
  for (int j = 0; j < NC; ++j) {
    va[j] = a[base + j * nelem];
    vb[j] = b[base + j * nelem];
  }
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    T ra = 0, ia = 0;
    if (HASA) {
#pragma unroll
      for (int j = 0; j < NC; ++j) {
        const T w = la[(c * NC + j) * nell + l];
        ra += w * va[j].x;
        ia += w * va[j].y;
      }
    } else {
      ra = va[c].x;
      ia = va[c].y;
    }
    T rb = 0, ib = 0;
#pragma unroll
    for (int j = 0; j < NC; ++j) {
      const T w = lb[(c * NC + j) * nell + l];
      rb += w * vb[j].x;
      ib += w * vb[j].y;
    }
    out[base + c * nelem] = {ra + rb, ia + ib};
  }
}

// out[t, c, p] = x[t, c, p] + dcomp[c] * mask[c, p] * (a[t, c, p] - b[t, c, p]),
// each of x, b, dcomp and mask optional (absent: 0, 0, 1, 1); rows (t, c) of
// npix pixels. The difference is rounded first, then each factor, then the
// sum (the order of the separate array operations it replaces). Any array
// may alias out: every thread reads its element of each before writing.
template <typename T>
__global__ void __launch_bounds__(256)
map_fused_kernel(T *out, const T *x, const T *a, const T *b, const T *__restrict__ dcomp,
                 const T *__restrict__ mask, int ncomp, long long npix) {
#pragma clang fp contract(off)
  const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= npix) return;
  const long long row = blockIdx.y;
  const int c = (int)(row % ncomp);
  const long long i = row * npix + p;
  T v = a[i];
  if (b) v = v - b[i];
  if (mask) v = v * mask[(long long)c * npix + p];
  if (dcomp) v = v * dcomp[c];
  out[i] = x ? x[i] + v : v;
}

// map_fused_kernel on 16-byte vectors of V elements (npix % V == 0, all
// rows 16-byte aligned); the same arithmetic per element.
template <typename T, int V>
__global__ void __launch_bounds__(256)
map_fused_vec_kernel(T *out, const T *x, const T *a, const T *b, const T *__restrict__ dcomp,
                     const T *__restrict__ mask, int ncomp, long long npix) {
#pragma clang fp contract(off)
  struct alignas(16) Vec { T v[V]; };
  const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long nq = npix / V;
  if (q >= nq) return;
  const long long row = blockIdx.y;
  const int c = (int)(row % ncomp);
  const long long i = row * nq + q;
  Vec va = reinterpret_cast<const Vec *>(a)[i];
  if (b) {
    const Vec vb = reinterpret_cast<const Vec *>(b)[i];
#pragma unroll
    for (int k = 0; k < V; ++k) va.v[k] = va.v[k] - vb.v[k];
  }
  if (mask) {
    const Vec vm = reinterpret_cast<const Vec *>(mask)[(long long)c * nq + q];
#pragma unroll
    for (int k = 0; k < V; ++k) va.v[k] = va.v[k] * vm.v[k];
  }
  if (dcomp) {
    const T d = dcomp[c];
#pragma unroll
    for (int k = 0; k < V; ++k) va.v[k] = va.v[k] * d;
  }
  if (x) {
    const Vec vx = reinterpret_cast<const Vec *>(x)[i];
#pragma unroll
    for (int k = 0; k < V; ++k) va.v[k] = vx.v[k] + va.v[k];
  }
  reinterpret_cast<Vec *>(out)[i] = va;
}

template <typename T>
static void launch_map_fused(T *out, const T *x, const T *a, const T *b, const T *dcomp,
                             const T *mask, int ncomp, long long npix, long long rows,
                             hipStream_t st) {
  constexpr int V = 16 / sizeof(T);
  auto al = [](const void *ptr) { return ptr == nullptr || ((uintptr_t)ptr % 16) == 0; };
  if (npix % V == 0 && al(out) && al(x) && al(a) && al(b) && al(mask)) {
    dim3 grid(ceil_div(npix / V, 256), (unsigned)rows);
    hipLaunchKernelGGL((map_fused_vec_kernel<T, V>), grid, dim3(256), 0, st, out, x, a, b,
                       dcomp, mask, ncomp, npix);
    return;
  }
  dim3 grid(ceil_div(npix, 256), (unsigned)rows);
  hipLaunchKernelGGL(map_fused_kernel<T>, grid, dim3(256), 0, st, out, x, a, b, dcomp, mask,
                     ncomp, npix);
}

template <typename T, int NC>
static void launch_lmul_add_nc(const T *la, const cplx_t<T> *a, const T *lb, const cplx_t<T> *b,
                               cplx_t<T> *out, int lmax, long long nelem, long long ntrans,
                               hipStream_t st) {
  dim3 grid(ceil_div(nelem, 256), (unsigned)ntrans);
  if (la)
    hipLaunchKernelGGL((lmul_add_kernel<T, NC, true>), grid, dim3(256), 0, st, la, a, lb, b, out,
                       lmax, nelem, ntrans);
  else
    hipLaunchKernelGGL((lmul_add_kernel<T, NC, false>), grid, dim3(256), 0, st, la, a, lb, b,
                       out, lmax, nelem, ntrans);
}

template <typename T>
static void lmul_add_impl(const void *lmat_a, const void *a, const void *lmat_b, const void *b,
                          void *out, int lmax, int ncomp, long long ntrans, hipStream_t st) {
  const long long nelem = nelem_of(lmax);
  OWMI_REQUIRE(ntrans < 65536, "too many transforms for one lmul_add launch");
  const T *La = static_cast<const T *>(lmat_a);
  const T *Lb = static_cast<const T *>(lmat_b);
  auto *A = static_cast<const cplx_t<T> *>(a);
  auto *B = static_cast<const cplx_t<T> *>(b);
  auto *O = static_cast<cplx_t<T> *>(out);
  switch (ncomp) {
    case 1: launch_lmul_add_nc<T, 1>(La, A, Lb, B, O, lmax, nelem, ntrans, st); break;
    case 2: launch_lmul_add_nc<T, 2>(La, A, Lb, B, O, lmax, nelem, ntrans, st); break;
    case 3: launch_lmul_add_nc<T, 3>(La, A, Lb, B, O, lmax, nelem, ntrans, st); break;
    case 4: launch_lmul_add_nc<T, 4>(La, A, Lb, B, O, lmax, nelem, ntrans, st); break;
    default: throw Error("lmul_add: ncomp must lie in [1, 4]");
  }
}

template <typename T>
static void lmul_impl(const void *lmat, int diag, const void *in, void *out, int lmax,
                      int ncomp, long long ntrans, hipStream_t st) {
  long long nelem = nelem_of(lmax);
  dim3 block(256);
  const T *L = static_cast<const T *>(lmat);
  auto *I = static_cast<const cplx_t<T> *>(in);
  auto *O = static_cast<cplx_t<T> *>(out);
  OWMI_REQUIRE(ntrans < 65536, "too many transforms for one lmul launch");
  dim3 grid(ceil_div(nelem, 256), (unsigned)ntrans);
  if (diag) {
    hipLaunchKernelGGL(lmul_diag_kernel<T>, grid, block, 0, st, L, I, O, lmax, ncomp, nelem,
                       ntrans);
    return;
  }
  switch (ncomp) {
    case 1: hipLaunchKernelGGL((lmul_full_kernel<T, 1>), grid, block, 0, st, L, I, O, lmax, nelem, ntrans); break;
    case 2: hipLaunchKernelGGL((lmul_full_kernel<T, 2>), grid, block, 0, st, L, I, O, lmax, nelem, ntrans); break;
    case 3: hipLaunchKernelGGL((lmul_full_kernel<T, 3>), grid, block, 0, st, L, I, O, lmax, nelem, ntrans); break;
    case 4: hipLaunchKernelGGL((lmul_full_kernel<T, 4>), grid, block, 0, st, L, I, O, lmax, nelem, ntrans); break;
    default:
      OWMI_REQUIRE(ncomp <= 16, "ncomp > 16 not supported");
      hipLaunchKernelGGL(lmul_full_generic_kernel<T>, grid, block, 0, st, L, I, O, lmax, ncomp, nelem, ntrans);
  }
}

// ------------------------------------------------------ trunc / wlm2alm

template <typename T>
__global__ void trunc_kernel(const cplx_t<T> *alm, cplx_t<T> *out, int lmax, int lmax_out,
                             long long nelem_in, long long nelem_out) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  long long row = blockIdx.y;
  if (i >= nelem_out) return;
  int l, m;
  lm_of(i, lmax_out, l, m);
  out[row * nelem_out + i] = alm[row * nelem_in + mstart_of(m, lmax) + l];
}

template <typename T>
__global__ void wlm2alm_kernel(const T *__restrict__ w, const cplx_t<T> *wlm, cplx_t<T> *alm,
                               int lmax_w, int lmax_a, long long nelem_w, long long nelem_a) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  long long row = blockIdx.y;
  if (i >= nelem_w) return;
  int l, m;
  lm_of(i, lmax_w, l, m);
  cplx_t<T> v = wlm[row * nelem_w + i];
  T a = w[l];
  cplx_t<T> *dst = alm + row * nelem_a + mstart_of(m, lmax_a) + l;
  cplx_t<T> o = *dst;
  *dst = {o.x + a * v.x, o.y + a * v.y};
}

// ---------------------------------------------------------- ring weight

template <typename T>
__global__ void ringweight_kernel(T *imap, const double *__restrict__ weight,
                                  const long long *__restrict__ nphi,
                                  const long long *__restrict__ offsets,
                                  const long long *__restrict__ stride, long long map_stride) {
  long long r = blockIdx.x;
  long long k = blockIdx.y;
  T w = (T)weight[r];
  long long n = nphi[r], off = offsets[r], s = stride[r];
  T *base = imap + k * map_stride + off;
  for (long long j = threadIdx.x; j < n; j += blockDim.x) base[j * s] *= w;
}

// ---------------------------------------------------------- pix matvec

// Pixel-diagonal multiply of ntrans maps of n elements by one diagonal: each
// thread holds one 16-byte vector of the diagonal and applies it to up to
// kPixTG maps (blockIdx.y groups them). Requires 16-byte aligned buffers and
// n a multiple of the vector width; pix_diag_scalar_kernel otherwise.
constexpr int kPixTG = 4;

template <typename V, typename T>
__global__ void __launch_bounds__(256)
pix_diag_kernel(const T *__restrict__ mat, const T *in, T *out, long long n, long long ntrans) {
  constexpr int W = sizeof(V) / sizeof(T);
  const long long v = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= n / W) return;
  const V d = reinterpret_cast<const V *>(mat)[v];
  const long long t1 = min(ntrans, (long long)(blockIdx.y + 1) * kPixTG);
  long long t = (long long)blockIdx.y * kPixTG;
  // kLoadB maps' loads before their stores (in place is allowed; each
  // thread touches only its own vectors)
  for (; t + kLoadB <= t1; t += kLoadB) {
    V x[kLoadB];
#pragma unroll
    for (int b = 0; b < kLoadB; ++b) x[b] = reinterpret_cast<const V *>(in + (t + b) * n)[v];
#pragma unroll
    for (int b = 0; b < kLoadB; ++b) reinterpret_cast<V *>(out + (t + b) * n)[v] = d * x[b];
  }
  for (; t < t1; ++t)
    reinterpret_cast<V *>(out + t * n)[v] = d * reinterpret_cast<const V *>(in + t * n)[v];
}

template <typename T>
__global__ void __launch_bounds__(256)
pix_diag_scalar_kernel(const T *__restrict__ mat, const T *in, T *out, long long n,
                       long long ntrans) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const T d = mat[i];
  const long long t1 = min(ntrans, (long long)(blockIdx.y + 1) * kPixTG);
  for (long long t = (long long)blockIdx.y * kPixTG; t < t1; ++t) out[t * n + i] = d * in[t * n + i];
}

template <typename T>
static void launch_pix_diag(const T *mat, const T *in, T *out, long long n, long long ntrans,
                            hipStream_t st) {
  typedef T V __attribute__((ext_vector_type(16 / sizeof(T))));
  constexpr int W = 16 / sizeof(T);
  auto aligned = [](const void *p) { return (reinterpret_cast<uintptr_t>(p) & 15) == 0; };
  const unsigned tg = (unsigned)ceil_div(ntrans, kPixTG);
  if (n % W == 0 && aligned(mat) && aligned(in) && aligned(out)) {
    dim3 g(ceil_div(n / W, 256), tg);
    hipLaunchKernelGGL((pix_diag_kernel<V, T>), g, dim3(256), 0, st, mat, in, out, n, ntrans);
  } else {
    dim3 g(ceil_div(n, 256), tg);
    hipLaunchKernelGGL(pix_diag_scalar_kernel<T>, g, dim3(256), 0, st, mat, in, out, n, ntrans);
  }
}

template <typename T>
__global__ void pix_full_kernel(const T *__restrict__ mat, const T *in, T *out, int nc,
                                long long npix) {
  long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  long long t = blockIdx.y;
  if (p >= npix) return;
  const T *src = in + t * nc * npix;
  T *dst = out + t * nc * npix;
  T v[16];
  for (int j = 0; j < nc; ++j) v[j] = src[j * npix + p];
  for (int i = 0; i < nc; ++i) {
    T acc = 0;
    for (int j = 0; j < nc; ++j) acc += mat[(i * nc + j) * npix + p] * v[j];
    dst[i * npix + p] = acc;
  }
}

// --------------------------------------------------------------- dot

constexpr int kDotBlocks = 1024;
constexpr int kDotThreads = 256;

__global__ void __launch_bounds__(kDotThreads)
dot_final_kernel(const double *partial, int n, double *out) {
  __shared__ double red[kDotThreads];
  double acc = 0.0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) acc += partial[i];
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int s = kDotThreads / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = red[0];
}

// out[b] = sum of the n partials of product b (one block per product).
__global__ void __launch_bounds__(kDotThreads)
dot_batch_final_kernel(const double *partial, int n, double *out) {
  __shared__ double red[kDotThreads];
  const double *src = partial + (long long)blockIdx.x * n;
  double acc = 0.0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) acc += src[i];
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int s = kDotThreads / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[blockIdx.x] = red[0];
}

// Partial sums of the two-pass reductions (kDotBlocks doubles), one buffer
// per device and stream.
constexpr int kTagDot = 1;
static double *dot_scratch(hipStream_t st) {
  return (double *)stream_scratch(kTagDot, kDotBlocks * sizeof(double), st);
}

// The partial sums of owmi_alm_dot_batch_noalloc: a fixed device array
// (one copy per GPU, 8 KiB), shared by every call on that GPU. Host code
// never passes its name to a kernel: the address each device holds it at
// comes from hipGetSymbolAddress (noalloc_scratch_ptr).
static __device__ __align__(256) double scratch[kDotBlocks];

// Device address of `scratch` on the current device, looked up once per
// device (no allocation, host or device).
constexpr int kMaxDevices = 64;
static double *noalloc_scratch_ptr() {
  static std::mutex mu;
  static double *ptrs[kMaxDevices] = {};
  int dev = 0;
  OWMI_HIP(hipGetDevice(&dev));
  OWMI_REQUIRE(dev >= 0 && dev < kMaxDevices, "alm_dot_batch_noalloc: device index above 63");
  std::lock_guard<std::mutex> lk(mu);
  if (ptrs[dev] != nullptr) return ptrs[dev];
  double *p = nullptr;
  const hipError_t e = hipGetSymbolAddress((void **)&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || p == nullptr)
    throw Error(std::string("alm_dot_batch_noalloc: hipGetSymbolAddress(scratch): ") +
                hipGetErrorString(e));
  ptrs[dev] = p;
  return p;
}

// Batched dot: nbatch independent products, each over nrows rows of nelem
// complex elements; block (x, b) reduces a strided slice of every row of
// product b. The first lmax + 1 elements of a row (m = 0) count once, the
// rest twice (real field, m < 0 implied). kLoadB elements per thread and
// step are loaded before they are summed.
template <typename T>
__global__ void __launch_bounds__(kDotThreads)
dot_batch_partial_kernel(const cplx_t<T> *a, const cplx_t<T> *b, int lmax, long long nelem,
                         long long nrows, double *partial) {
  __shared__ double red[kDotThreads];
  const long long stride = (long long)gridDim.x * blockDim.x;
  const long long j0 = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  double acc = 0.0;
  for (long long r = 0; r < nrows; ++r) {
    const long long base = ((long long)blockIdx.y * nrows + r) * nelem;
    const cplx_t<T> *ar = a + base, *br = b + base;
    long long j = j0;
    for (; j + (kLoadB - 1) * stride < nelem; j += kLoadB * stride) {
      cplx_t<T> x[kLoadB], y[kLoadB];
#pragma unroll
      for (int k = 0; k < kLoadB; ++k) {
        x[k] = ar[j + k * stride];
        y[k] = br[j + k * stride];
      }
#pragma unroll
      for (int k = 0; k < kLoadB; ++k) {
        const double v = (double)x[k].x * (double)y[k].x + (double)x[k].y * (double)y[k].y;
        acc += (j + k * stride <= lmax) ? v : 2.0 * v;
      }
    }
    for (; j < nelem; j += stride) {
      const cplx_t<T> x = ar[j], y = br[j];
      const double v = (double)x.x * (double)y.x + (double)x.y * (double)y.y;
      acc += (j <= lmax) ? v : 2.0 * v;
    }
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int s = kDotThreads / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[(long long)blockIdx.y * gridDim.x + blockIdx.x] = red[0];
}

// ------------------------------------------------------------ CG updates

template <typename T>
__global__ void cg_xr_kernel(cplx_t<T> *x, cplx_t<T> *r, const cplx_t<T> *p,
                             const cplx_t<T> *q, const double *num, const double *den,
                             long long n) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  T alpha = (T)(*num / *den);
  cplx_t<T> pv = p[i], qv = q[i], xv = x[i], rv = r[i];
  x[i] = {xv.x + alpha * pv.x, xv.y + alpha * pv.y};
  r[i] = {rv.x - alpha * qv.x, rv.y - alpha * qv.y};
}

// x (precision X) += alpha p; r (precision R) -= alpha q; alpha in double.
template <typename X, typename R>
__global__ void cg_xr_mixed_kernel(cplx_t<X> *x, cplx_t<R> *r, const cplx_t<R> *p,
                                   const cplx_t<R> *q, const double *num, const double *den,
                                   long long n) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double alpha = *num / *den;
  cplx_t<R> pv = p[i], qv = q[i], rv = r[i];
  cplx_t<X> xv = x[i];
  x[i] = {(X)(xv.x + alpha * (double)pv.x), (X)(xv.y + alpha * (double)pv.y)};
  const R a = (R)alpha;
  r[i] = {rv.x - a * qv.x, rv.y - a * qv.y};
}

template <typename T>
__global__ void cg_p_kernel(cplx_t<T> *p, const cplx_t<T> *z, const double *num,
                            const double *den, long long n) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  T beta = (T)(*num / *den);
  cplx_t<T> pv = p[i], zv = z[i];
  p[i] = {zv.x + beta * pv.x, zv.y + beta * pv.y};
}

// Batched updates: transform b (blockIdx.y) uses its own scalars
// num[b] / den[b]; a zero denominator gives a zero coefficient, which
// freezes a right-hand side whose residual and direction were cleared.
__device__ __forceinline__ double safe_ratio(const double *num, const double *den, long long b) {
  const double d = den[b];
  return d != 0.0 ? num[b] / d : 0.0;
}

template <typename X, typename R>
__global__ void cg_xr_batch_kernel(cplx_t<X> *x, cplx_t<R> *r, const cplx_t<R> *p,
                                   const cplx_t<R> *q, const double *num, const double *den,
                                   long long n) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const long long b = blockIdx.y, o = b * n + i;
  const double alpha = safe_ratio(num, den, b);
  const cplx_t<R> pv = p[o], qv = q[o], rv = r[o];
  const cplx_t<X> xv = x[o];
  x[o] = {(X)(xv.x + alpha * (double)pv.x), (X)(xv.y + alpha * (double)pv.y)};
  const R a = (R)alpha;
  r[o] = {rv.x - a * qv.x, rv.y - a * qv.y};
}

// lmul_add_kernel with the real-field <b, out> of every transform folded in
// (the CG coefficient <p, A p> when b = p and out = A p): block (x, t)
// takes a strided slice of the elements of transform t (all components of
// an element in one thread, m = 0 elements counted once, the rest twice)
// and writes its partial sum; dot_batch_final_kernel adds the partials of
// each transform in fixed order. a may alias out (every input of an element
// is read before it is written).
template <typename T, int NC, bool HASA>
__global__ void __launch_bounds__(kDotThreads)
lmul_add_dot_kernel(const T *__restrict__ la, const cplx_t<T> *a, const T *__restrict__ lb,
                    const cplx_t<T> *b, cplx_t<T> *out, int lmax, long long nelem,
                    double *partial) {
  __shared__ double red[kDotThreads];
  const long long t = blockIdx.y;
  const long long stride = (long long)gridDim.x * blockDim.x;
  const int nell = lmax + 1;
  double acc = 0.0;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < nelem; i += stride) {
    int l, m;
    lm_of(i, lmax, l, m);
    const long long base = t * NC * nelem + i;
    cplx_t<T> va[NC], vb[NC];
#pragma unroll
    for (int j = 0; j < NC; ++j) {
      va[j] = a[base + j * nelem];
      vb[j] = b[base + j * nelem];
    }
    double v = 0.0;
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      T ra = 0, ia = 0;
      if (HASA) {
#pragma unroll
        for (int j = 0; j < NC; ++j) {
          const T w = la[(c * NC + j) * nell + l];
          ra += w * va[j].x;
          ia += w * va[j].y;
        }
      } else {
        ra = va[c].x;
        ia = va[c].y;
      }
      T rb = 0, ib = 0;
#pragma unroll
      for (int j = 0; j < NC; ++j) {
        const T w = lb[(c * NC + j) * nell + l];
        rb += w * vb[j].x;
        ib += w * vb[j].y;
      }
      const cplx_t<T> o = {ra + rb, ia + ib};
      out[base + c * nelem] = o;
      v += (double)vb[c].x * (double)o.x + (double)vb[c].y * (double)o.y;
    }
    acc += (i <= lmax) ? v : 2.0 * v;
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int s = kDotThreads / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[t * gridDim.x + blockIdx.x] = red[0];
}

// The update of cg_xr_batch_kernel with the real-field <r, r> of the new
// residual folded in: block (x, b) updates a strided slice of every row of
// right-hand side b (the slices of dot_batch_partial_kernel, m = 0 elements
// counted once, the rest twice) and writes its partial sum of |r|^2 (of r as
// stored, in precision R); dot_batch_final_kernel adds the partials of each
// right-hand side in fixed order.
template <typename X, typename R>
__global__ void __launch_bounds__(kDotThreads)
cg_xr_rr_batch_kernel(cplx_t<X> *x, cplx_t<R> *r, const cplx_t<R> *p, const cplx_t<R> *q,
                      const double *num, const double *den, int lmax, long long nelem,
                      long long nrows, double *partial) {
  __shared__ double red[kDotThreads];
  const long long b = blockIdx.y;
  const double alpha = safe_ratio(num, den, b);
  const R a = (R)alpha;
  const long long stride = (long long)gridDim.x * blockDim.x;
  const long long j0 = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  double acc = 0.0;
  for (long long row = 0; row < nrows; ++row) {
    const long long base = (b * nrows + row) * nelem;
    for (long long j = j0; j < nelem; j += stride) {
      const long long o = base + j;
      const cplx_t<R> pv = p[o], qv = q[o], rv = r[o];
      const cplx_t<X> xv = x[o];
      x[o] = {(X)(xv.x + alpha * (double)pv.x), (X)(xv.y + alpha * (double)pv.y)};
      const cplx_t<R> rn = {rv.x - a * qv.x, rv.y - a * qv.y};
      r[o] = rn;
      const double v = (double)rn.x * (double)rn.x + (double)rn.y * (double)rn.y;
      acc += (j <= lmax) ? v : 2.0 * v;
    }
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int s = kDotThreads / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[b * gridDim.x + blockIdx.x] = red[0];
}

template <typename T>
__global__ void cg_p_batch_kernel(cplx_t<T> *p, const cplx_t<T> *z, const double *num,
                                  const double *den, long long n) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const long long b = blockIdx.y, o = b * n + i;
  const T beta = (T)safe_ratio(num, den, b);
  const cplx_t<T> pv = p[o], zv = z[o];
  p[o] = {zv.x + beta * pv.x, zv.y + beta * pv.y};
}

template <typename T>
__global__ void axpby_kernel(T a, const cplx_t<T> *x, T b, cplx_t<T> *y, long long n) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  cplx_t<T> xv = x[i], yv = y[i];
  y[i] = {a * xv.x + b * yv.x, a * xv.y + b * yv.y};
}

// One complex element per thread (a 16- or 8-byte load, an 8- or 16-byte
// store).
template <typename S, typename D>
__global__ void __launch_bounds__(256)
convert_kernel(const cplx_t<S> *src, cplx_t<D> *dst, long long n) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const cplx_t<S> v = src[i];
  dst[i] = {(D)v.x, (D)v.y};
}

}  // namespace owmi

using namespace owmi;

template <typename T, int NC>
static void lmul_add_dot_nc(const T *la, const cplx_t<T> *a, const T *lb, const cplx_t<T> *b,
                            cplx_t<T> *out, int lmax, long long ntrans, double *dev_dot,
                            hipStream_t st) {
  const long long nelem = nelem_of(lmax);
  const int bpb = kDotBlocks / (int)ntrans;
  double *part = dot_scratch(st);
  dim3 g((unsigned)bpb, (unsigned)ntrans);
  if (la)
    hipLaunchKernelGGL((lmul_add_dot_kernel<T, NC, true>), g, dim3(kDotThreads), 0, st, la, a,
                       lb, b, out, lmax, nelem, part);
  else
    hipLaunchKernelGGL((lmul_add_dot_kernel<T, NC, false>), g, dim3(kDotThreads), 0, st, la, a,
                       lb, b, out, lmax, nelem, part);
  hipLaunchKernelGGL(dot_batch_final_kernel, dim3((unsigned)ntrans), dim3(kDotThreads), 0, st,
                     part, bpb, dev_dot);
}

template <typename T>
static void lmul_add_dot_impl(const void *la, const void *a, const void *lb, const void *b,
                              void *out, int lmax, int ncomp, long long ntrans, double *dev_dot,
                              hipStream_t st) {
  using C = cplx_t<T>;
  const T *pla = (const T *)la, *plb = (const T *)lb;
  const C *pa = (const C *)a, *pb = (const C *)b;
  C *po = (C *)out;
  switch (ncomp) {
    case 1: return lmul_add_dot_nc<T, 1>(pla, pa, plb, pb, po, lmax, ntrans, dev_dot, st);
    case 2: return lmul_add_dot_nc<T, 2>(pla, pa, plb, pb, po, lmax, ntrans, dev_dot, st);
    case 3: return lmul_add_dot_nc<T, 3>(pla, pa, plb, pb, po, lmax, ntrans, dev_dot, st);
    case 4: return lmul_add_dot_nc<T, 4>(pla, pa, plb, pb, po, lmax, ntrans, dev_dot, st);
    default: throw Error("lmul_add_dot: ncomp must be in 1..4");
  }
}

extern "C" {

const char *owmi_last_error(void) { return owmi::g_last_error.c_str(); }

int owmi_abi_version(void) { return 1; }

int owmi_lmul(int dtype, const void *lmat, int diag, const void *in, void *out, int lmax,
              int ncomp, long long ntrans, void *stream) {
  return guard([&] {
    check_dtype(dtype);
    OWMI_REQUIRE(lmax >= 0 && ncomp >= 1 && ntrans >= 0, "invalid lmul sizes");
    if (ntrans == 0) return;
    if (dtype == OWMI_F64)
      lmul_impl<double>(lmat, diag, in, out, lmax, ncomp, ntrans, as_stream(stream));
    else
      lmul_impl<float>(lmat, diag, in, out, lmax, ncomp, ntrans, as_stream(stream));
  });
}

int owmi_map_fused(int dtype, void *out, const void *x, const void *a, const void *b,
                   const void *dcomp, const void *mask, int ncomp, long long npix,
                   long long ntrans, void *stream) {
  return guard([&] {
    check_dtype(dtype);
    OWMI_REQUIRE(out != nullptr && a != nullptr, "map_fused: out and a are required");
    OWMI_REQUIRE(ncomp >= 1 && npix >= 0 && ntrans >= 0, "invalid map_fused sizes");
    const long long rows = (long long)ncomp * ntrans;
    OWMI_REQUIRE(rows < 65536, "map_fused: too many rows for one launch");
    if (rows == 0 || npix == 0) return;
    if (dtype == OWMI_F64)
      launch_map_fused<double>((double *)out, (const double *)x, (const double *)a,
                               (const double *)b, (const double *)dcomp, (const double *)mask,
                               ncomp, npix, rows, as_stream(stream));
    else
      launch_map_fused<float>((float *)out, (const float *)x, (const float *)a,
                              (const float *)b, (const float *)dcomp, (const float *)mask, ncomp,
                              npix, rows, as_stream(stream));
  });
}

int owmi_lmul_add(int dtype, const void *lmat_a, const void *a, const void *lmat_b,
                  const void *b, void *out, int lmax, int ncomp, long long ntrans,
                  void *stream) {
  return guard([&] {
    check_dtype(dtype);
    OWMI_REQUIRE(lmax >= 0 && ncomp >= 1 && ntrans >= 0, "invalid lmul_add sizes");
    OWMI_REQUIRE(lmat_b != nullptr && a != nullptr && b != nullptr && out != nullptr,
                 "lmul_add: lmat_b, a, b and out are required");
    if (ntrans == 0) return;
    if (dtype == OWMI_F64)
      lmul_add_impl<double>(lmat_a, a, lmat_b, b, out, lmax, ncomp, ntrans, as_stream(stream));
    else
      lmul_add_impl<float>(lmat_a, a, lmat_b, b, out, lmax, ncomp, ntrans, as_stream(stream));
  });
}

int owmi_lmul_add_dot(int dtype, const void *lmat_a, const void *a, const void *lmat_b,
                      const void *b, void *out, int lmax, int ncomp, long long ntrans,
                      double *dev_dot, void *stream) {
  return guard([&] {
    check_dtype(dtype);
    OWMI_REQUIRE(lmax >= 0 && ncomp >= 1 && ntrans >= 1 && ntrans <= kDotBlocks,
                 "invalid lmul_add_dot sizes (ntrans must be in [1, 1024])");
    OWMI_REQUIRE(lmat_b != nullptr && a != nullptr && b != nullptr && out != nullptr &&
                     dev_dot != nullptr,
                 "lmul_add_dot: lmat_b, a, b, out and dev_dot are required");
    OWMI_REQUIRE(out != b, "lmul_add_dot: b (the dot's first factor) may not alias out");
    if (dtype == OWMI_F64)
      lmul_add_dot_impl<double>(lmat_a, a, lmat_b, b, out, lmax, ncomp, ntrans, dev_dot,
                                as_stream(stream));
    else
      lmul_add_dot_impl<float>(lmat_a, a, lmat_b, b, out, lmax, ncomp, ntrans, dev_dot,
                               as_stream(stream));
  });
}

int owmi_trunc_alm(int dtype, const void *alm, void *alm_out, int lmax, int lmax_out,
                   long long nrows, void *stream) {
  return guard([&] {
    check_dtype(dtype);
    OWMI_REQUIRE(lmax_out <= lmax && lmax_out >= 0, "lmax_out must be in [0, lmax]");
    OWMI_REQUIRE(nrows < 65536, "too many rows");
    if (nrows == 0) return;
    long long ni = nelem_of(lmax), no = nelem_of(lmax_out);
    dim3 grid(ceil_div(no, 256), (unsigned)nrows);
    if (dtype == OWMI_F64)
      hipLaunchKernelGGL(trunc_kernel<double>, grid, dim3(256), 0, as_stream(stream),
                         (const double2 *)alm, (double2 *)alm_out, lmax, lmax_out, ni, no);
    else
      hipLaunchKernelGGL(trunc_kernel<float>, grid, dim3(256), 0, as_stream(stream),
                         (const float2 *)alm, (float2 *)alm_out, lmax, lmax_out, ni, no);
  });
}

int owmi_wlm2alm(int dtype, const void *w_ell, const void *wlm, void *alm, int lmax_w,
                 int lmax_a, long long nrows, void *stream) {
  return guard([&] {
    check_dtype(dtype);
    OWMI_REQUIRE(lmax_w <= lmax_a && lmax_w >= 0, "lmax_w must be in [0, lmax_a]");
    OWMI_REQUIRE(nrows < 65536, "too many rows");
    if (nrows == 0) return;
    long long nw = nelem_of(lmax_w), na = nelem_of(lmax_a);
    dim3 grid(ceil_div(nw, 256), (unsigned)nrows);
    if (dtype == OWMI_F64)
      hipLaunchKernelGGL(wlm2alm_kernel<double>, grid, dim3(256), 0, as_stream(stream),
                         (const double *)w_ell, (const double2 *)wlm, (double2 *)alm, lmax_w,
                         lmax_a, nw, na);
    else
      hipLaunchKernelGGL(wlm2alm_kernel<float>, grid, dim3(256), 0, as_stream(stream),
                         (const float *)w_ell, (const float2 *)wlm, (float2 *)alm, lmax_w,
                         lmax_a, nw, na);
  });
}

int owmi_alm_dot(int dtype, const void *a, const void *b, int lmax, long long nrows,
                 double *dev_out, void *stream) {
  return guard([&] {
    check_dtype(dtype);
    long long nelem = nelem_of(lmax);
    hipStream_t st = as_stream(stream);
    double *part = dot_scratch(st);
    if (dtype == OWMI_F64)
      hipLaunchKernelGGL(dot_batch_partial_kernel<double>, dim3(kDotBlocks), dim3(kDotThreads), 0, st,
                         (const double2 *)a, (const double2 *)b, lmax, nelem, nrows, part);
    else
      hipLaunchKernelGGL(dot_batch_partial_kernel<float>, dim3(kDotBlocks), dim3(kDotThreads), 0, st,
                         (const float2 *)a, (const float2 *)b, lmax, nelem, nrows, part);
    hipLaunchKernelGGL(dot_final_kernel, dim3(1), dim3(kDotThreads), 0, st, part, kDotBlocks,
                       dev_out);
  });
}

// The batched dot through a partial-sum buffer `part` of kDotBlocks doubles.
static void alm_dot_batch_impl(int dtype, const void *a, const void *b, int lmax,
                               long long nrows, long long nbatch, double *dev_out,
                               double *part, hipStream_t st) {
  const long long nelem = nelem_of(lmax);
  const int bpb = kDotBlocks / (int)nbatch;
  dim3 g(bpb, (unsigned)nbatch);
  if (dtype == OWMI_F64)
    hipLaunchKernelGGL(dot_batch_partial_kernel<double>, g, dim3(kDotThreads), 0, st,
                       (const double2 *)a, (const double2 *)b, lmax, nelem, nrows, part);
  else
    hipLaunchKernelGGL(dot_batch_partial_kernel<float>, g, dim3(kDotThreads), 0, st,
                       (const float2 *)a, (const float2 *)b, lmax, nelem, nrows, part);
  hipLaunchKernelGGL(dot_batch_final_kernel, dim3((unsigned)nbatch), dim3(kDotThreads), 0, st,
                     part, bpb, dev_out);
}

int owmi_alm_dot_batch(int dtype, const void *a, const void *b, int lmax, long long nrows,
                       long long nbatch, double *dev_out, void *stream) {
  return guard([&] {
    check_dtype(dtype);
    OWMI_REQUIRE(nbatch >= 1 && nbatch <= kDotBlocks, "nbatch must be in [1, 1024]");
    OWMI_REQUIRE(lmax >= 0 && nrows >= 0, "invalid alm_dot_batch sizes");
    hipStream_t st = as_stream(stream);
    alm_dot_batch_impl(dtype, a, b, lmax, nrows, nbatch, dev_out, dot_scratch(st), st);
  });
}

int owmi_alm_dot_batch_noalloc(int dtype, const void *a, const void *b, int lmax,
                               long long nrows, long long nbatch, double *dev_out,
                               void *stream) {
  return guard([&] {
    check_dtype(dtype);
    OWMI_REQUIRE(nbatch >= 1 && nbatch <= kDotBlocks,
                 "alm_dot_batch_noalloc: nbatch must be in [1, 1024] (the fixed scratch holds "
                 "1024 partial sums)");
    OWMI_REQUIRE(lmax >= 0 && nrows >= 0, "invalid alm_dot_batch_noalloc sizes");
    hipStream_t st = as_stream(stream);
    // every partial is written by the first pass before the second reads
    // it: nothing needs clearing between calls
    alm_dot_batch_impl(dtype, a, b, lmax, nrows, nbatch, dev_out, noalloc_scratch_ptr(), st);
  });
}

int owmi_cg_xr_batch(int dtype_x, void *x, int dtype_r, void *r, const void *p, const void *q,
                     const double *dev_num, const double *dev_den, long long n,
                     long long nbatch, void *stream) {
  return guard([&] {
    check_dtype(dtype_x);
    check_dtype(dtype_r);
    OWMI_REQUIRE(nbatch >= 1 && nbatch < 65536, "nbatch must be in [1, 65535]");
    if (n == 0) return;
    dim3 g(ceil_div(n, 256), (unsigned)nbatch);
    hipStream_t st = as_stream(stream);
    if (dtype_x == OWMI_F64 && dtype_r == OWMI_F32)
      hipLaunchKernelGGL((cg_xr_batch_kernel<double, float>), g, dim3(256), 0, st, (double2 *)x,
                         (float2 *)r, (const float2 *)p, (const float2 *)q, dev_num, dev_den, n);
    else if (dtype_x == OWMI_F64 && dtype_r == OWMI_F64)
      hipLaunchKernelGGL((cg_xr_batch_kernel<double, double>), g, dim3(256), 0, st,
                         (double2 *)x, (double2 *)r, (const double2 *)p, (const double2 *)q,
                         dev_num, dev_den, n);
    else if (dtype_x == OWMI_F32 && dtype_r == OWMI_F32)
      hipLaunchKernelGGL((cg_xr_batch_kernel<float, float>), g, dim3(256), 0, st, (float2 *)x,
                         (float2 *)r, (const float2 *)p, (const float2 *)q, dev_num, dev_den, n);
    else
      throw Error("x must be at least as precise as r");
  });
}

int owmi_cg_xr_rr_batch(int dtype_x, void *x, int dtype_r, void *r, const void *p,
                        const void *q, const double *dev_num, const double *dev_den, int lmax,
                        long long nrows, long long nbatch, double *dev_rr, void *stream) {
  return guard([&] {
    check_dtype(dtype_x);
    check_dtype(dtype_r);
    OWMI_REQUIRE(lmax >= 0 && nrows >= 1, "invalid cg_xr_rr_batch sizes");
    OWMI_REQUIRE(nbatch >= 1 && nbatch <= kDotBlocks, "nbatch must be in [1, 1024]");
    OWMI_REQUIRE(dev_rr != nullptr, "cg_xr_rr_batch: dev_rr is required");
    hipStream_t st = as_stream(stream);
    const long long nelem = nelem_of(lmax);
    const int bpb = kDotBlocks / (int)nbatch;
    double *part = dot_scratch(st);
    dim3 g((unsigned)bpb, (unsigned)nbatch);
    if (dtype_x == OWMI_F64 && dtype_r == OWMI_F32)
      hipLaunchKernelGGL((cg_xr_rr_batch_kernel<double, float>), g, dim3(kDotThreads), 0, st,
                         (double2 *)x, (float2 *)r, (const float2 *)p, (const float2 *)q,
                         dev_num, dev_den, lmax, nelem, nrows, part);
    else if (dtype_x == OWMI_F64 && dtype_r == OWMI_F64)
      hipLaunchKernelGGL((cg_xr_rr_batch_kernel<double, double>), g, dim3(kDotThreads), 0, st,
                         (double2 *)x, (double2 *)r, (const double2 *)p, (const double2 *)q,
                         dev_num, dev_den, lmax, nelem, nrows, part);
    else if (dtype_x == OWMI_F32 && dtype_r == OWMI_F32)
      hipLaunchKernelGGL((cg_xr_rr_batch_kernel<float, float>), g, dim3(kDotThreads), 0, st,
                         (float2 *)x, (float2 *)r, (const float2 *)p, (const float2 *)q,
                         dev_num, dev_den, lmax, nelem, nrows, part);
    else
      throw Error("x must be at least as precise as r");
    hipLaunchKernelGGL(dot_batch_final_kernel, dim3((unsigned)nbatch), dim3(kDotThreads), 0, st,
                       part, bpb, dev_rr);
  });
}

int owmi_cg_p_batch(int dtype, void *p, const void *z, const double *dev_num,
                    const double *dev_den, long long n, long long nbatch, void *stream) {
  return guard([&] {
    check_dtype(dtype);
    OWMI_REQUIRE(nbatch >= 1 && nbatch < 65536, "nbatch must be in [1, 65535]");
    if (n == 0) return;
    dim3 g(ceil_div(n, 256), (unsigned)nbatch);
    if (dtype == OWMI_F64)
      hipLaunchKernelGGL(cg_p_batch_kernel<double>, g, dim3(256), 0, as_stream(stream),
                         (double2 *)p, (const double2 *)z, dev_num, dev_den, n);
    else
      hipLaunchKernelGGL(cg_p_batch_kernel<float>, g, dim3(256), 0, as_stream(stream),
                         (float2 *)p, (const float2 *)z, dev_num, dev_den, n);
  });
}

int owmi_apply_ringweight(int dtype, void *imap, const double *weight, const long long *nphi,
                          const long long *offsets, const long long *stride, long long nrow,
                          long long nmaps, long long map_stride, void *stream) {
  return guard([&] {
    check_dtype(dtype);
    OWMI_REQUIRE(nmaps < 65536, "too many maps");
    if (nrow == 0 || nmaps == 0) return;
    dim3 grid((unsigned)nrow, (unsigned)nmaps);
    if (dtype == OWMI_F64)
      hipLaunchKernelGGL(ringweight_kernel<double>, grid, dim3(256), 0, as_stream(stream),
                         (double *)imap, weight, nphi, offsets, stride, map_stride);
    else
      hipLaunchKernelGGL(ringweight_kernel<float>, grid, dim3(256), 0, as_stream(stream),
                         (float *)imap, weight, nphi, offsets, stride, map_stride);
  });
}

int owmi_pix_matvec(int dtype, const void *mat, int kind, const void *in, void *out, int ncomp,
                    long long npix, long long ntrans, void *stream) {
  return guard([&] {
    check_dtype(dtype);
    OWMI_REQUIRE(ncomp >= 1 && ncomp <= 16, "ncomp must be in [1, 16]");
    OWMI_REQUIRE(kind == 0 || kind == 1, "kind must be 0 (diagonal) or 1 (full)");
    OWMI_REQUIRE(ntrans >= 0 && ntrans < 65536, "ntrans must be in [0, 65535]");
    if (ntrans == 0 || npix == 0) return;
    hipStream_t st = as_stream(stream);
    if (kind == 0) {
      long long n = npix * ncomp;
      if (dtype == OWMI_F64)
        launch_pix_diag<double>((const double *)mat, (const double *)in, (double *)out, n, ntrans,
                                st);
      else
        launch_pix_diag<float>((const float *)mat, (const float *)in, (float *)out, n, ntrans, st);
    } else {
      dim3 g(ceil_div(npix, 256), (unsigned)ntrans);
      if (dtype == OWMI_F64)
        hipLaunchKernelGGL(pix_full_kernel<double>, g, dim3(256), 0, st,
                           (const double *)mat, (const double *)in, (double *)out, ncomp, npix);
      else
        hipLaunchKernelGGL(pix_full_kernel<float>, g, dim3(256), 0, st,
                           (const float *)mat, (const float *)in, (float *)out, ncomp, npix);
    }
  });
}

int owmi_cg_xr_mixed(int dtype_x, void *x, int dtype_r, void *r, const void *p, const void *q,
                     const double *dev_num, const double *dev_den, long long n, void *stream) {
  return guard([&] {
    check_dtype(dtype_x);
    check_dtype(dtype_r);
    if (n == 0) return;
    dim3 g(ceil_div(n, 256));
    hipStream_t st = as_stream(stream);
    if (dtype_x == OWMI_F64 && dtype_r == OWMI_F32)
      hipLaunchKernelGGL((cg_xr_mixed_kernel<double, float>), g, dim3(256), 0, st, (double2 *)x,
                         (float2 *)r, (const float2 *)p, (const float2 *)q, dev_num, dev_den, n);
    else if (dtype_x == OWMI_F64 && dtype_r == OWMI_F64)
      hipLaunchKernelGGL((cg_xr_mixed_kernel<double, double>), g, dim3(256), 0, st,
                         (double2 *)x, (double2 *)r, (const double2 *)p, (const double2 *)q,
                         dev_num, dev_den, n);
    else if (dtype_x == OWMI_F32 && dtype_r == OWMI_F32)
      hipLaunchKernelGGL((cg_xr_mixed_kernel<float, float>), g, dim3(256), 0, st, (float2 *)x,
                         (float2 *)r, (const float2 *)p, (const float2 *)q, dev_num, dev_den, n);
    else
      throw Error("x must be at least as precise as r");
  });
}

int owmi_cg_xr(int dtype, void *x, void *r, const void *p, const void *q, const double *dev_num,
               const double *dev_den, long long n, void *stream) {
  return guard([&] {
    check_dtype(dtype);
    if (n == 0) return;
    dim3 g(ceil_div(n, 256));
    if (dtype == OWMI_F64)
      hipLaunchKernelGGL(cg_xr_kernel<double>, g, dim3(256), 0, as_stream(stream), (double2 *)x,
                         (double2 *)r, (const double2 *)p, (const double2 *)q, dev_num, dev_den, n);
    else
      hipLaunchKernelGGL(cg_xr_kernel<float>, g, dim3(256), 0, as_stream(stream), (float2 *)x,
                         (float2 *)r, (const float2 *)p, (const float2 *)q, dev_num, dev_den, n);
  });
}

int owmi_cg_p(int dtype, void *p, const void *z, const double *dev_num, const double *dev_den,
              long long n, void *stream) {
  return guard([&] {
    check_dtype(dtype);
    if (n == 0) return;
    dim3 g(ceil_div(n, 256));
    if (dtype == OWMI_F64)
      hipLaunchKernelGGL(cg_p_kernel<double>, g, dim3(256), 0, as_stream(stream), (double2 *)p,
                         (const double2 *)z, dev_num, dev_den, n);
    else
      hipLaunchKernelGGL(cg_p_kernel<float>, g, dim3(256), 0, as_stream(stream), (float2 *)p,
                         (const float2 *)z, dev_num, dev_den, n);
  });
}

int owmi_axpby(int dtype, double a, const void *x, double b, void *y, long long n, void *stream) {
  return guard([&] {
    check_dtype(dtype);
    if (n == 0) return;
    dim3 g(ceil_div(n, 256));
    if (dtype == OWMI_F64)
      hipLaunchKernelGGL(axpby_kernel<double>, g, dim3(256), 0, as_stream(stream), a,
                         (const double2 *)x, b, (double2 *)y, n);
    else
      hipLaunchKernelGGL(axpby_kernel<float>, g, dim3(256), 0, as_stream(stream), (float)a,
                         (const float2 *)x, (float)b, (float2 *)y, n);
  });
}

int owmi_convert(int dtype_src, const void *src, int dtype_dst, void *dst, long long n,
                 void *stream) {
  return guard([&] {
    check_dtype(dtype_src);
    check_dtype(dtype_dst);
    if (n == 0) return;
    dim3 g(ceil_div(n, 256));
    hipStream_t st = as_stream(stream);
    if (dtype_src == OWMI_F64 && dtype_dst == OWMI_F32)
      hipLaunchKernelGGL((convert_kernel<double, float>), g, dim3(256), 0, st,
                         (const double2 *)src, (float2 *)dst, n);
    else if (dtype_src == OWMI_F32 && dtype_dst == OWMI_F64)
      hipLaunchKernelGGL((convert_kernel<float, double>), g, dim3(256), 0, st,
                         (const float2 *)src, (double2 *)dst, n);
    else if (dtype_src == OWMI_F64)
      hipLaunchKernelGGL((convert_kernel<double, double>), g, dim3(256), 0, st,
                         (const double2 *)src, (double2 *)dst, n);
    else
      hipLaunchKernelGGL((convert_kernel<float, float>), g, dim3(256), 0, st,
                         (const float2 *)src, (float2 *)dst, n);
  });
}

}  // extern "C"
