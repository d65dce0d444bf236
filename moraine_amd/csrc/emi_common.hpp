// Helpers shared by the EMI kernels: uniform runtime indexing of per-lane
// register arrays and the triu pair index.
#pragma once
#include "common.hpp"

namespace mr {

constexpr int EMI_WIDE_MAX_N = 200;  // largest n served by the register-resident kernels

#define MR_C8(B) C(B + 0) C(B + 1) C(B + 2) C(B + 3) C(B + 4) C(B + 5) C(B + 6) C(B + 7)
#define MR_C64 MR_C8(0) MR_C8(8) MR_C8(16) MR_C8(24) MR_C8(32) MR_C8(40) MR_C8(48) MR_C8(56)
#define MR_C128 MR_C64 MR_C8(64) MR_C8(72) MR_C8(80) MR_C8(88) MR_C8(96) MR_C8(104) MR_C8(112) MR_C8(120)

// a[k] for a wave-uniform runtime k: a switch (scalar branch tree + v_mov), so
// the array stays in VGPRs instead of going to scratch memory.
template <typename V, int NM>
__device__ __forceinline__ V reg_get(const V (&a)[NM], int k) {
  static_assert(NM <= 128, "reg_get covers at most 128 entries");
  V v = V(0);
  switch (k) {
#define C(i) \
  case i:    \
    if constexpr ((i) < NM) v = a[i]; \
    break;
    MR_C128
#undef C
    default:
      break;
  }
  return v;
}

// index of pair (a, b), a < b, in triu order (moraine/tnet.py:15-20)
__device__ __forceinline__ int pair_index(int a, int b, int n) { return a * (2 * n - a - 1) / 2 + (b - a - 1); }

}  // namespace mr
