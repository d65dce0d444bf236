// Shared helpers for the gfx950 DS hot-path kernels: error reporting, launch
// checks, wave-level reductions and complex helpers.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <string>

#include "../../include/moraine_hip.h"

namespace mr {

void set_error(const char* fmt, ...);

#define MR_CHECK_ARG(cond, ...)          \
  do {                                   \
    if (!(cond)) {                       \
      ::mr::set_error(__VA_ARGS__);      \
      return MR_ERR_ARG;                 \
    }                                    \
  } while (0)

#define MR_CHECK_HIP(expr)                                                      \
  do {                                                                          \
    hipError_t e__ = (expr);                                                    \
    if (e__ != hipSuccess) {                                                    \
      ::mr::set_error("%s failed: %s", #expr, hipGetErrorString(e__));          \
      return MR_ERR_HIP;                                                        \
    }                                                                           \
  } while (0)

// Launch status: kernel launch errors are reported from hipGetLastError.
inline int launch_status(const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    set_error("%s launch failed: %s", what, hipGetErrorString(e));
    return MR_ERR_HIP;
  }
  return MR_OK;
}

inline hipStream_t as_stream(mr_stream_t s) { return reinterpret_cast<hipStream_t>(s); }

inline unsigned div_up(unsigned long long a, unsigned long long b) { return (unsigned)((a + b - 1) / b); }

// ---- wave (64-lane) reductions -------------------------------------------
__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
__device__ __forceinline__ int wave_sum(int v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// ---- complex helpers (float2 / double2 as complex) ------------------------
struct c64 {
  float re, im;
};
struct c128 {
  double re, im;
};
__device__ __forceinline__ c128 cmul(c128 a, c128 b) {
  return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
__device__ __forceinline__ c128 cmulc(c128 a, c128 b) {  // a * conj(b)
  return {a.re * b.re + a.im * b.im, a.im * b.re - a.re * b.im};
}

}  // namespace mr
