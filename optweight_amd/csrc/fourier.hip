// 2-D Fourier (flat-sky) kernels for the Fourier-wavelet noise model:
// per-pixel matrices on 2-D Fourier coefficients, the gather of kernel
// slices (2-D coefficients -> per-scale coefficient blocks) and the
// scatter-add back, the real-field inner product of half-plane
// coefficients, and the fix-ups that give the library c2c/r2c/c2r passes
// the exact real-field conventions of a 2-D real FFT.
//
// Layouts: a coefficient array is (nlead, ny, nlx) complex, rows of the
// half plane lx = 0 .. nlx - 1 contiguous. Scale j of a kernel set is a
// block (nlead, ny_j, nx_j) of the packed buffer, its row r taken from
// full row rows[j][r] and its columns from x0_j .. x0_j + nx_j - 1.
// Descriptor table (int64, kDescLen per scale): block offset in complex
// elements, ny_j, nx_j, x0_j, offset into the row table, offset into the
// packed kernel values (ny_j * nx_j reals per scale), first thread block of
// the scale in the gather's block list (blocks of kThreads elements).

#include <map>
#include <mutex>

#include "common.hpp"

namespace owmi {
namespace {

constexpr int kThreads = 256;
constexpr int kDescLen = 8;
constexpr int kDotBlocks = 1024;

// --------------------------------------------------------------- fmul

// out[i, p] = sum_j M[i, j, p] in[j, p] on complex in/out, real M
// (ncomp, ncomp, npix); in may alias out (all inputs are read first).
template <typename T, int NC>
__global__ void __launch_bounds__(kThreads)
fmul_full_kernel(const T *__restrict__ mat, const cplx_t<T> *in, cplx_t<T> *out,
                 long long npix, long long ntrans) {
  const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long t = blockIdx.y;
  if (p >= npix || t >= ntrans) return;
  const cplx_t<T> *src = in + t * NC * npix;
  cplx_t<T> *dst = out + t * NC * npix;
  cplx_t<T> v[NC];
#pragma unroll
  for (int j = 0; j < NC; ++j) v[j] = src[j * npix + p];
  T m[NC * NC];
#pragma unroll
  for (int k = 0; k < NC * NC; ++k) m[k] = mat[k * npix + p];
#pragma unroll
  for (int i = 0; i < NC; ++i) {
    T re = 0, im = 0;
#pragma unroll
    for (int j = 0; j < NC; ++j) {
      re += m[i * NC + j] * v[j].x;
      im += m[i * NC + j] * v[j].y;
    }
    dst[i * npix + p] = cplx_t<T>{re, im};
  }
}

// Generic component count (ncomp > 4): same product, components in a loop.
template <typename T>
__global__ void __launch_bounds__(kThreads)
fmul_full_generic_kernel(const T *__restrict__ mat, const cplx_t<T> *in, cplx_t<T> *out,
                         int nc, long long npix, long long ntrans, cplx_t<T> *tmp) {
  const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long t = blockIdx.y;
  if (p >= npix || t >= ntrans) return;
  const cplx_t<T> *src = in + t * nc * npix;
  cplx_t<T> *dst = out + t * nc * npix;
  cplx_t<T> *buf = tmp + t * nc * npix;
  for (int i = 0; i < nc; ++i) {
    T re = 0, im = 0;
    for (int j = 0; j < nc; ++j) {
      const T mm = mat[((long long)i * nc + j) * npix + p];
      const cplx_t<T> v = src[j * npix + p];
      re += mm * v.x;
      im += mm * v.y;
    }
    buf[i * npix + p] = cplx_t<T>{re, im};
  }
  for (int i = 0; i < nc; ++i) dst[i * npix + p] = buf[i * npix + p];
}

// out[t, c, p] = d[c * dstride + p] * in[t, c, p] over rows (t, c) of
// ncomp components (dstride 0: one diagonal shared by all components).
template <typename T>
__global__ void __launch_bounds__(kThreads)
fmul_diag_kernel(const T *__restrict__ d, const cplx_t<T> *in, cplx_t<T> *out, long long npix,
                 long long dstride, long long nrows, int ncomp) {
  const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long row = blockIdx.y;
  if (p >= npix || row >= nrows) return;
  const T m = d[(row % ncomp) * dstride + p];
  const cplx_t<T> v = in[row * npix + p];
  out[row * npix + p] = cplx_t<T>{m * v.x, m * v.y};
}

// ------------------------------------------------------- gather / scatter

// small_j[b, r, c] = full[b, rows_j[r], x0_j + c] * kern_j[r, c], block j
// starting at nlead * off_j in small; grid (sum_j ceil(ny_j nx_j / 256),
// 1, nlead), each thread block inside one scale (its first block in desc).
template <typename T>
__global__ void __launch_bounds__(kThreads)
fwav_gather_kernel(const cplx_t<T> *__restrict__ full, long long ny, long long nlx,
                   int nscale, const long long *__restrict__ desc,
                   const int *__restrict__ rows, const T *__restrict__ kern,
                   cplx_t<T> *__restrict__ small) {
  int j = 0;
  while (j + 1 < nscale && desc[(long long)(j + 1) * kDescLen + 6] <= (long long)blockIdx.x) ++j;
  const long long *dj = desc + (long long)j * kDescLen;
  const long long off = dj[0], x0 = dj[3], roff = dj[4], koff = dj[5];
  const unsigned nxj = (unsigned)dj[2];
  const long long cnt = dj[1] * dj[2];
  const long long e = ((long long)blockIdx.x - dj[6]) * blockDim.x + threadIdx.x;
  if (e >= cnt) return;
  const long long b = blockIdx.z;
  const unsigned r = (unsigned)e / nxj;
  const long long c = e - (long long)r * nxj;
  const long long y = rows[roff + r];
  const cplx_t<T> v = full[(b * ny + y) * nlx + x0 + c];
  const T k = kern[koff + e];
  small[(long long)gridDim.z * off + b * cnt + e] = cplx_t<T>{v.x * k, v.y * k};
}

// full[b, y, x] = sum_j [y, x in slice j] small_j[b, r_j(y), x - x0_j] *
// kern_j[r_j(y), x - x0_j], scales added in table order from zero;
// invrows[j * ny + y] = r_j(y) or -1. Every element of full is written.
// grid (ceil(nlx / 256), ny, nlead): the row and the scales' row indices
// are uniform per thread block.
template <typename T>
__global__ void __launch_bounds__(kThreads)
fwav_scatter_kernel(const cplx_t<T> *__restrict__ small, long long ny, long long nlx,
                    long long nlead, int nscale, const long long *__restrict__ desc,
                    const int *__restrict__ invrows, const T *__restrict__ kern,
                    cplx_t<T> *__restrict__ full) {
  const long long x = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= nlx) return;
  const long long y = blockIdx.y;
  const long long b = blockIdx.z;
  const long long i = (b * ny + y) * nlx + x;
  T re = 0, im = 0;
  for (int j = 0; j < nscale; ++j) {
    const long long *dj = desc + (long long)j * kDescLen;
    const long long nxj = dj[2], x0 = dj[3];
    if (x < x0 || x >= x0 + nxj) continue;
    const int r = invrows[(long long)j * ny + y];
    if (r < 0) continue;
    const long long nyj = dj[1];
    const long long e = (long long)r * nxj + (x - x0);
    const cplx_t<T> v = small[nlead * dj[0] + b * nyj * nxj + e];
    const T k = kern[dj[5] + e];
    re += v.x * k;
    im += v.y * k;
  }
  full[i] = cplx_t<T>{re, im};
}

// ------------------------------------------------------------- fix-ups

// mode 0 (after a real-to-complex transform of (nlead, ny, nx) real data):
// the self-conjugate coefficients (ly in {0, ny/2}, lx in {0, nx/2}) are
// real; their rounding-level imaginary parts are set to zero.
// mode 1 (before the complex-to-real pass along x, after the inverse c2c
// along y): the real part alone of column 0 (and column nlx - 1 for even
// nx) enters a real-field synthesis; the imaginary parts are zeroed.
template <typename T>
__global__ void __launch_bounds__(kThreads)
fourier_fix_kernel(cplx_t<T> *data, long long nlead, long long ny, long long nlx, int nx_even,
                   int mode) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (mode == 1) {
    const long long nrows = nlead * ny;
    if (i >= nrows) return;
    data[i * nlx].y = T(0);
    if (nx_even && nlx > 1) data[i * nlx + nlx - 1].y = T(0);
    return;
  }
  if (i >= nlead) return;
  cplx_t<T> *base = data + i * ny * nlx;
  base[0].y = T(0);
  if (nx_even && nlx > 1) base[nlx - 1].y = T(0);
  if (ny % 2 == 0 && ny > 1) {
    base[(ny / 2) * nlx].y = T(0);
    if (nx_even && nlx > 1) base[(ny / 2) * nlx + nlx - 1].y = T(0);
  }
}

// ------------------------------------------------------------------ dot

// Partials of the real-field inner product of half-plane coefficients:
// s_main = sum_(rows, x) (2 - [x == 0]) Re(f conj g), s_last = sum over
// column nlx - 1 of Re(f conj g), nonreal = number of coefficients of f at
// (ly = 0 and ly = ny/2 for even ny, lx = nlx - 1) with a nonzero
// imaginary part. Three partials per block.
template <typename T>
__global__ void __launch_bounds__(kThreads)
fdot_partial_kernel(const cplx_t<T> *__restrict__ f, const cplx_t<T> *__restrict__ g,
                    long long nlead, long long ny, long long nlx, double *partial) {
  __shared__ double red[3][kThreads];
  const long long n = nlead * ny * nlx;
  const long long stride = (long long)gridDim.x * blockDim.x;
  double s_main = 0.0, s_last = 0.0, nonreal = 0.0;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const cplx_t<T> a = f[i];
    const cplx_t<T> b = g[i];
    const double p = (double)a.x * (double)b.x + (double)a.y * (double)b.y;
    const long long x = i % nlx;
    s_main += (x == 0 ? 1.0 : 2.0) * p;
    if (x == nlx - 1) {
      s_last += p;
      const long long y = (i / nlx) % ny;
      if ((y == 0 || (ny % 2 == 0 && y == ny / 2)) && a.y != T(0)) nonreal += 1.0;
    }
  }
  red[0][threadIdx.x] = s_main;
  red[1][threadIdx.x] = s_last;
  red[2][threadIdx.x] = nonreal;
  __syncthreads();
  for (int s = kThreads / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      red[0][threadIdx.x] += red[0][threadIdx.x + s];
      red[1][threadIdx.x] += red[1][threadIdx.x + s];
      red[2][threadIdx.x] += red[2][threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    partial[blockIdx.x] = red[0][0];
    partial[gridDim.x + blockIdx.x] = red[1][0];
    partial[2 * gridDim.x + blockIdx.x] = red[2][0];
  }
}

// out = s_main - [nx even] s_last; nx_parity 1 / 0 gives the parity, -1
// takes it from the data (even when no inspected coefficient of f has an
// imaginary part).
__global__ void __launch_bounds__(kThreads)
fdot_final_kernel(const double *partial, int nblk, int nx_parity, double *out) {
  __shared__ double red[3][kThreads];
  double a = 0.0, b = 0.0, c = 0.0;
  for (int i = threadIdx.x; i < nblk; i += blockDim.x) {
    a += partial[i];
    b += partial[nblk + i];
    c += partial[2 * nblk + i];
  }
  red[0][threadIdx.x] = a;
  red[1][threadIdx.x] = b;
  red[2][threadIdx.x] = c;
  __syncthreads();
  for (int s = kThreads / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      red[0][threadIdx.x] += red[0][threadIdx.x + s];
      red[1][threadIdx.x] += red[1][threadIdx.x + s];
      red[2][threadIdx.x] += red[2][threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const bool even = nx_parity >= 0 ? nx_parity == 1 : red[2][0] == 0.0;
    *out = red[0][0] - (even ? red[1][0] : 0.0);
  }
}

// Partial sums of owmi_fourier_dot (3 kDotBlocks doubles), one buffer per
// device and stream.
constexpr int kTagFourierDot = 2;
double *fdot_scratch(hipStream_t st) {
  return (double *)stream_scratch(kTagFourierDot, 3 * kDotBlocks * sizeof(double), st);
}

template <typename T>
void launch_fmul(const void *mat, int kind, const void *in, void *out, int ncomp,
                 long long npix, long long ntrans, long long dstride, hipStream_t st) {
  const cplx_t<T> *src = (const cplx_t<T> *)in;
  cplx_t<T> *dst = (cplx_t<T> *)out;
  const T *m = (const T *)mat;
  if (kind == 0) {
    const long long nrows = (long long)ncomp * ntrans;
    OWMI_REQUIRE(nrows <= 65535, "fmul: too many rows for one launch");
    hipLaunchKernelGGL(fmul_diag_kernel<T>, dim3(ceil_div(npix, kThreads), (unsigned)nrows),
                       dim3(kThreads), 0, st, m, src, dst, npix, dstride, nrows, ncomp);
    return;
  }
  OWMI_REQUIRE(ntrans <= 65535, "fmul: too many transforms for one launch");
  dim3 grid(ceil_div(npix, kThreads), (unsigned)ntrans);
  switch (ncomp) {
    case 1: hipLaunchKernelGGL((fmul_full_kernel<T, 1>), grid, dim3(kThreads), 0, st, m, src, dst, npix, ntrans); break;
    case 2: hipLaunchKernelGGL((fmul_full_kernel<T, 2>), grid, dim3(kThreads), 0, st, m, src, dst, npix, ntrans); break;
    case 3: hipLaunchKernelGGL((fmul_full_kernel<T, 3>), grid, dim3(kThreads), 0, st, m, src, dst, npix, ntrans); break;
    case 4: hipLaunchKernelGGL((fmul_full_kernel<T, 4>), grid, dim3(kThreads), 0, st, m, src, dst, npix, ntrans); break;
    default:
      throw Error("fmul: full matrices need ncomp <= 4 (use the generic entry)");
  }
}

}  // namespace
}  // namespace owmi

using namespace owmi;

extern "C" {

int owmi_fmul(int dtype, const void *mat, int kind, const void *in, void *out, int ncomp,
              long long npix, long long ntrans, long long diag_stride, void *stream) {
  return guard([&] {
    check_dtype(dtype);
    OWMI_REQUIRE(kind == 0 || kind == 1, "kind must be 0 (diagonal) or 1 (full)");
    OWMI_REQUIRE(ncomp >= 1 && npix >= 0 && ntrans >= 0, "bad sizes");
    if (npix == 0 || ntrans == 0) return;
    if (dtype == OWMI_F64)
      launch_fmul<double>(mat, kind, in, out, ncomp, npix, ntrans, diag_stride, as_stream(stream));
    else
      launch_fmul<float>(mat, kind, in, out, ncomp, npix, ntrans, diag_stride, as_stream(stream));
  });
}

int owmi_fmul_generic(int dtype, const void *mat, const void *in, void *out, void *tmp,
                      int ncomp, long long npix, long long ntrans, void *stream) {
  return guard([&] {
    check_dtype(dtype);
    OWMI_REQUIRE(ncomp >= 1 && npix >= 0 && ntrans >= 0 && ntrans <= 65535, "bad sizes");
    if (npix == 0 || ntrans == 0) return;
    dim3 grid(ceil_div(npix, kThreads), (unsigned)ntrans);
    if (dtype == OWMI_F64)
      hipLaunchKernelGGL(fmul_full_generic_kernel<double>, grid, dim3(kThreads), 0,
                         as_stream(stream), (const double *)mat, (const double2 *)in,
                         (double2 *)out, ncomp, npix, ntrans, (double2 *)tmp);
    else
      hipLaunchKernelGGL(fmul_full_generic_kernel<float>, grid, dim3(kThreads), 0,
                         as_stream(stream), (const float *)mat, (const float2 *)in,
                         (float2 *)out, ncomp, npix, ntrans, (float2 *)tmp);
  });
}

int owmi_fwav_gather(int dtype, const void *full, long long nlead, long long ny, long long nlx,
                     int nscale, const long long *desc, const int *rows, const void *kern,
                     long long nblocks, void *small, void *stream) {
  return guard([&] {
    check_dtype(dtype);
    OWMI_REQUIRE(nscale >= 0 && nlead >= 0 && nlead <= 65535,
                 "fwav_gather: nlead must lie in [0, 65535]");
    OWMI_REQUIRE(nblocks >= 0 && nblocks < (1LL << 31), "fwav_gather: bad block count");
    if (nscale == 0 || nlead == 0 || nblocks == 0) return;
    dim3 grid((unsigned)nblocks, 1, (unsigned)nlead);
    if (dtype == OWMI_F64)
      hipLaunchKernelGGL(fwav_gather_kernel<double>, grid, dim3(kThreads), 0, as_stream(stream),
                         (const double2 *)full, ny, nlx, nscale, desc, rows,
                         (const double *)kern, (double2 *)small);
    else
      hipLaunchKernelGGL(fwav_gather_kernel<float>, grid, dim3(kThreads), 0, as_stream(stream),
                         (const float2 *)full, ny, nlx, nscale, desc, rows,
                         (const float *)kern, (float2 *)small);
  });
}

int owmi_fwav_scatter(int dtype, const void *small, long long nlead, long long ny, long long nlx,
                      int nscale, const long long *desc, const int *invrows, const void *kern,
                      void *full, void *stream) {
  return guard([&] {
    check_dtype(dtype);
    OWMI_REQUIRE(nscale >= 0 && nlead >= 0 && nlead <= 65535 && ny >= 0 && ny <= 65535,
                 "fwav_scatter: ny and nlead must lie in [0, 65535]");
    if (nlead * ny * nlx == 0) return;
    dim3 grid(ceil_div(nlx, kThreads), (unsigned)ny, (unsigned)nlead);
    if (dtype == OWMI_F64)
      hipLaunchKernelGGL(fwav_scatter_kernel<double>, grid, dim3(kThreads), 0, as_stream(stream),
                         (const double2 *)small, ny, nlx, nlead, nscale, desc, invrows,
                         (const double *)kern, (double2 *)full);
    else
      hipLaunchKernelGGL(fwav_scatter_kernel<float>, grid, dim3(kThreads), 0, as_stream(stream),
                         (const float2 *)small, ny, nlx, nlead, nscale, desc, invrows,
                         (const float *)kern, (float2 *)full);
  });
}

int owmi_fourier_fix(int dtype, void *data, long long nlead, long long ny, long long nlx,
                     int nx_even, int mode, void *stream) {
  return guard([&] {
    check_dtype(dtype);
    OWMI_REQUIRE(mode == 0 || mode == 1, "mode must be 0 (after r2c) or 1 (before c2r)");
    const long long n = mode == 1 ? nlead * ny : nlead;
    if (n == 0 || nlx == 0) return;
    if (dtype == OWMI_F64)
      hipLaunchKernelGGL(fourier_fix_kernel<double>, dim3(ceil_div(n, kThreads)), dim3(kThreads),
                         0, as_stream(stream), (double2 *)data, nlead, ny, nlx, nx_even, mode);
    else
      hipLaunchKernelGGL(fourier_fix_kernel<float>, dim3(ceil_div(n, kThreads)), dim3(kThreads),
                         0, as_stream(stream), (float2 *)data, nlead, ny, nlx, nx_even, mode);
  });
}

int owmi_fourier_dot(int dtype, const void *f, const void *g, long long nlead, long long ny,
                     long long nlx, int nx_parity, double *dev_out, void *stream) {
  return guard([&] {
    check_dtype(dtype);
    OWMI_REQUIRE(nx_parity >= -1 && nx_parity <= 1, "nx_parity must be -1, 0 or 1");
    hipStream_t st = as_stream(stream);
    double *part = fdot_scratch(st);
    if (dtype == OWMI_F64)
      hipLaunchKernelGGL(fdot_partial_kernel<double>, dim3(kDotBlocks), dim3(kThreads), 0, st,
                         (const double2 *)f, (const double2 *)g, nlead, ny, nlx, part);
    else
      hipLaunchKernelGGL(fdot_partial_kernel<float>, dim3(kDotBlocks), dim3(kThreads), 0, st,
                         (const float2 *)f, (const float2 *)g, nlead, ny, nlx, part);
    hipLaunchKernelGGL(fdot_final_kernel, dim3(1), dim3(kThreads), 0, st, part, kDotBlocks,
                       nx_parity, dev_out);
  });
}

}  // extern "C"
