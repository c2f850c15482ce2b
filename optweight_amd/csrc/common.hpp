// Shared helpers for the owmi HIP library (gfx950).
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstdio>
#include <string>
#include <stdexcept>

#include "owmi.h"

namespace owmi {

// Thread-local last-error string behind owmi_last_error().
void set_error(const std::string &msg);
void clear_error();

struct Error : std::runtime_error {
  using std::runtime_error::runtime_error;
};

#define OWMI_HIP(expr)                                                        \
  do {                                                                        \
    hipError_t e_ = (expr);                                                   \
    if (e_ != hipSuccess)                                                     \
      throw ::owmi::Error(std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

#define OWMI_REQUIRE(cond, msg)                   \
  do {                                            \
    if (!(cond)) throw ::owmi::Error(msg);        \
  } while (0)

// Wrap an ABI body: translate exceptions and launch errors into a status.
template <typename F>
int guard(F &&f) {
  try {
    clear_error();
    f();
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
      set_error(std::string("kernel launch: ") + hipGetErrorString(e));
      return 2;
    }
    return 0;
  } catch (const std::exception &ex) {
    set_error(ex.what());
    return 1;
  } catch (...) {
    set_error("unknown error");
    return 1;
  }
}

inline hipStream_t as_stream(void *s) { return reinterpret_cast<hipStream_t>(s); }

// Device buffer of `bytes` bytes private to (current device, stream, tag):
// the partial sums of a two-pass reduction, so reductions queued on
// different streams never share one. Allocated on the first call for the
// key (with the thread's stream-capture mode relaxed, so a first call
// inside a HIP-graph recording allocates outside the graph) and kept until
// the process ends; `bytes` must be the same on every call with one tag.
void *stream_scratch(int tag, size_t bytes, hipStream_t st);

inline void check_dtype(int dtype) {
  OWMI_REQUIRE(dtype == OWMI_F32 || dtype == OWMI_F64, "dtype must be OWMI_F32 or OWMI_F64");
}

// Complex type per real type.
template <typename T> struct cplx;
template <> struct cplx<float> { using type = float2; };
template <> struct cplx<double> { using type = double2; };
template <typename T> using cplx_t = typename cplx<T>::type;

__host__ __device__ inline long long nelem_of(int lmax) {
  return ((long long)lmax * lmax + 3LL * lmax) / 2 + 1;
}
__host__ __device__ inline long long mstart_of(int m, int lmax) {
  return (long long)m * (2LL * lmax + 1 - m) / 2;
}

inline unsigned int ceil_div(long long a, long long b) {
  return (unsigned int)((a + b - 1) / b);
}

}  // namespace owmi
