// Temporal coherence of one point, one wave (moraine/pl.py:80-123 for the DS
// form, moraine/pqm.py:46-121 for the PS form):
//   DS: t = |sum_p c_p/|c_p| conj(v_a) v_b| / n_pairs
//   PS: t = |sum_p c_p conj(v_a) v_b / (|c_p||v_a||v_b|)| / n_pairs
// Lanes take pairs p = lane, lane + 64, ...; the partial sums reduce in a fixed
// order, so every caller gets the same value for the same inputs.
#pragma once
#include "common.hpp"

namespace mr {

// coh_at(p, a, b): coherence of pair p = (a, b); ph_at(i): phase of image i
template <bool DS, class CohFn, class PhFn>
__device__ __forceinline__ float tcoh_point(const CohFn& coh_at, const PhFn& ph_at, const int32_t* __restrict__ pairs,
                                            int npairs, int lane) {
  float sr = 0.f, si = 0.f;
  for (int p = lane; p < npairs; p += 64) {
    const int a = pairs[2 * p], b = pairs[2 * p + 1];
    const float2 x = coh_at(p, a, b), va = ph_at(a), vb = ph_at(b);
    // conj(va) * vb
    float tr = va.x * vb.x + va.y * vb.y;
    float ti = va.x * vb.y - va.y * vb.x;
    float sc;
    if (DS) {
      sc = 1.f / sqrtf(x.x * x.x + x.y * x.y);
    } else {
      sc = 1.f / sqrtf((x.x * x.x + x.y * x.y) * (va.x * va.x + va.y * va.y) * (vb.x * vb.x + vb.y * vb.y));
    }
    sr += (x.x * tr - x.y * ti) * sc;
    si += (x.x * ti + x.y * tr) * sc;
  }
  sr = wave_sum(sr);
  si = wave_sum(si);
  return sqrtf(sr * sr + si * si) / (float)npairs;
}

}  // namespace mr
