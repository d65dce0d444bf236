// Helpers shared by the EMI kernels: uniform runtime indexing of per-lane
// register arrays and the triu pair index.
#pragma once
#include "common.hpp"

namespace mr {

// c(n) of the FP64 re-solve test of the EMI kernels: n / (60 sqrt 60)
__host__ __device__ inline double emi_flag_scale(int n) { return (double)n / (60.0 * 7.745966692414834); }

constexpr int EMI_WIDE_MAX_N = 200;  // largest n served by the register-resident kernels

// Per-lane register array read at a wave-uniform runtime index through
// M0-relative moves (s_set_gpr_idx_on + v_mov) instead of a branch tree: the
// entries live in vectors of 32 registers (32 floats or 16 doubles), the
// largest the backend indexes without going through scratch memory, plus one
// tail vector of the next power of two above the remainder, so at most a few
// registers are padding.  get/set take compile-time indices (inside unrolled
// loops), dyn a uniform runtime one.
constexpr int pow2_at_least(int x) { return x <= 2 ? 2 : 2 * pow2_at_least((x + 1) / 2); }

template <typename T, int NM>
struct RegArr {
  static constexpr int CH = 128 / sizeof(T);
  static constexpr int NF = NM / CH;  // full chunks
  static constexpr int REM = NM - NF * CH;
  static constexpr int RV = pow2_at_least(REM);
  typedef T vec __attribute__((ext_vector_type(CH)));
  typedef T tail_vec __attribute__((ext_vector_type(RV)));
  vec v[NF > 0 ? NF : 1];
  tail_vec tl;
  __device__ __forceinline__ T get(int j) const { return j < NF * CH ? v[j / CH][j % CH] : tl[j - NF * CH]; }
  __device__ __forceinline__ void set(int j, T x) {
    if (j < NF * CH)
      v[j / CH][j % CH] = x;
    else
      tl[j - NF * CH] = x;
  }
  __device__ __forceinline__ T dyn(int k) const {
    T r = T(0);
#pragma unroll
    for (int c = 0; c < NF; ++c)
      if (k / CH == c) r = v[c][k % CH];
    if (REM > 0 && k >= NF * CH) r = tl[k - NF * CH];
    return r;
  }
};

// index of pair (a, b), a < b, in triu order (moraine/tnet.py:15-20)
__device__ __forceinline__ int pair_index(int a, int b, int n) { return a * (2 * n - a - 1) / 2 + (b - a - 1); }

}  // namespace mr
