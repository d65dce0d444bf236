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

// ---- DPP / readlane cross-lane helpers (no LDS round trip) -----------------
// DPP controls (gfx9 family): quad_perm [1,0,3,2] and [2,3,0,1] are lane xor 1
// and xor 2; row_half_mirror / row_mirror pair the halves of 8- and 16-lane
// groups; wave_shr:1 / wave_shl:1 shift by one lane across the whole wave.
constexpr int DPP_XOR1 = 0xB1;
constexpr int DPP_XOR2 = 0x4E;
constexpr int DPP_HALF_MIRROR = 0x141;
constexpr int DPP_MIRROR = 0x140;
constexpr int DPP_WAVE_SHR1 = 0x138;  // lane i <- lane i-1 (lane 0 <- 0)
constexpr int DPP_WAVE_SHL1 = 0x130;  // lane i <- lane i+1 (lane 63 <- 0)
constexpr int DPP_ROW_BCAST15 = 0x142;  // row r <- lane 15 of row r-1 (row 0 reads 0)
constexpr int DPP_ROW_BCAST31 = 0x143;  // rows 2, 3 <- lane 31 (rows 0, 1 read 0)

template <int CTRL>
__device__ __forceinline__ int dpp_i(int v) {
  return __builtin_amdgcn_update_dpp(0, v, CTRL, 0xF, 0xF, false);
}
template <int CTRL>
__device__ __forceinline__ float dpp_f(float v) {
  return __int_as_float(dpp_i<CTRL>(__float_as_int(v)));
}
template <int CTRL>
__device__ __forceinline__ double dpp_d(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = dpp_i<CTRL>((int)(b & 0xffffffffll));
  const int hi = dpp_i<CTRL>((int)(b >> 32));
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
__device__ __forceinline__ int rl_i(int v, int l) { return __builtin_amdgcn_readlane(v, l); }
__device__ __forceinline__ float rl_f(float v, int l) { return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l)); }
__device__ __forceinline__ double rl_d(double v, int l) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffll), l);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// value of lane l ^ 32 (exchange between the two half-waves)
__device__ __forceinline__ float swap_half_f(float v) { return __shfl_xor(v, 32, 64); }

// Sum over all 64 lanes, result uniform (every lane must be active).  Fixed
// association order: after the in-row steps every lane holds its row sum r_i;
// row_bcast:15 then gives row 3 r3 + r2 and row 1 r1 + r0, and row_bcast:31
// adds row 1 into row 3, so lane 63 holds (r2 + r3) + (r0 + r1).
__device__ __forceinline__ float wsum_f(float v) {
  v += dpp_f<DPP_XOR1>(v);
  v += dpp_f<DPP_XOR2>(v);
  v += dpp_f<DPP_HALF_MIRROR>(v);
  v += dpp_f<DPP_MIRROR>(v);
  v += dpp_f<DPP_ROW_BCAST15>(v);
  v += dpp_f<DPP_ROW_BCAST31>(v);
  return rl_f(v, 63);
}
__device__ __forceinline__ double wsum_d(double v) {
  v += dpp_d<DPP_XOR1>(v);
  v += dpp_d<DPP_XOR2>(v);
  v += dpp_d<DPP_HALF_MIRROR>(v);
  v += dpp_d<DPP_MIRROR>(v);
  v += dpp_d<DPP_ROW_BCAST15>(v);
  v += dpp_d<DPP_ROW_BCAST31>(v);
  return rl_d(v, 63);
}
__device__ __forceinline__ double wmin_d(double v) {
  v = fmin(v, dpp_d<DPP_XOR1>(v));
  v = fmin(v, dpp_d<DPP_XOR2>(v));
  v = fmin(v, dpp_d<DPP_HALF_MIRROR>(v));
  v = fmin(v, dpp_d<DPP_MIRROR>(v));
  return fmin(fmin(rl_d(v, 0), rl_d(v, 16)), fmin(rl_d(v, 32), rl_d(v, 48)));
}
__device__ __forceinline__ double wmax_d(double v) {
  v = fmax(v, dpp_d<DPP_XOR1>(v));
  v = fmax(v, dpp_d<DPP_XOR2>(v));
  v = fmax(v, dpp_d<DPP_HALF_MIRROR>(v));
  v = fmax(v, dpp_d<DPP_MIRROR>(v));
  return fmax(fmax(rl_d(v, 0), rl_d(v, 16)), fmax(rl_d(v, 32), rl_d(v, 48)));
}
// Lane holding the largest key (ties -> smallest lane), uniform result.
__device__ __forceinline__ int wargmax_d(double key, int lane) {
  int idx = lane;
#define MR_AM_STEP(CTRL)                                  \
  {                                                       \
    const double ok = dpp_d<CTRL>(key);                   \
    const int oi = dpp_i<CTRL>(idx);                      \
    if (ok > key || (ok == key && oi < idx)) {            \
      key = ok;                                           \
      idx = oi;                                           \
    }                                                     \
  }
  MR_AM_STEP(DPP_XOR1)
  MR_AM_STEP(DPP_XOR2)
  MR_AM_STEP(DPP_HALF_MIRROR)
  MR_AM_STEP(DPP_MIRROR)
#undef MR_AM_STEP
  double bk = rl_d(key, 0);
  int bi = rl_i(idx, 0);
#pragma unroll
  for (int r = 1; r < 4; ++r) {
    const double k2 = rl_d(key, 16 * r);
    const int i2 = rl_i(idx, 16 * r);
    if (k2 > bk || (k2 == bk && i2 < bi)) {
      bk = k2;
      bi = i2;
    }
  }
  return __builtin_amdgcn_readfirstlane(bi);
}

// Wave max of an int, uniform result.
__device__ __forceinline__ int wmax_i(int v) {
  v = max(v, dpp_i<DPP_XOR1>(v));
  v = max(v, dpp_i<DPP_XOR2>(v));
  v = max(v, dpp_i<DPP_HALF_MIRROR>(v));
  v = max(v, dpp_i<DPP_MIRROR>(v));
  v = max(v, dpp_i<DPP_ROW_BCAST15>(v));  // lane 63 ends with the max of all four rows
  v = max(v, dpp_i<DPP_ROW_BCAST31>(v));
  return rl_i(v, 63);
}
// Pivot key for an integer max: the high word of a non-negative double (whose
// bit patterns order as the values do) with its low BITS bits replaced by
// (2^BITS - 1 - id), so the max picks the largest value to within a relative
// 2^-(20 - BITS) and, among equal keys, the smallest id; ineligible entries
// get -1.  id = (2^BITS - 1) - (max & (2^BITS - 1)).
template <int BITS>
__device__ __forceinline__ int pivot_key(double mag, bool eligible, int id) {
  constexpr int m = (1 << BITS) - 1;
  return eligible ? ((__double2hiint(mag) & ~m) | (m - id)) : -1;
}

// FP64 reciprocal: hardware approximation refined by two Newton steps (~1 ulp).
__device__ __forceinline__ double rcp_d(double x) {
  double r = __builtin_amdgcn_rcp(x);
  r = fma(r, fma(-x, r, 1.0), r);
  r = fma(r, fma(-x, r, 1.0), r);
  return r;
}

// ---- complex elements of complex64 (float2) / complex128 (double2) arrays ---
template <typename Z>
struct ZT;
template <>
struct ZT<float2> {
  using S = float;
};
template <>
struct ZT<double2> {
  using S = double;
};
template <typename Z>
__host__ __device__ __forceinline__ Z mkz(typename ZT<Z>::S x, typename ZT<Z>::S y) {
  Z v;
  v.x = x;
  v.y = y;
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
