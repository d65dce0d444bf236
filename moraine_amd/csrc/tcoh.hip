// Temporal coherence for gfx950.
//
// mr_ds_temp_coh replaces moraine.ds_temp_coh (moraine/pl.py:126-145; numba
// :80-100; kernel :103-123):  t = |sum_p coh_p/|coh_p| conj(ph_a) ph_b| / n_pairs.
// mr_temp_coh replaces moraine.temp_coh (moraine/pqm.py:124-159; numba :46-71;
// kernel :99-121):  t = |sum_p intf_p conj(r_a) r_b / (|intf_p||r_a||r_b|)| / n_pairs.
// One wave per point streams the point's pair row (coalesced, the HBM-bound
// part), gathers the phase terms from the point's N-vector staged in LDS (the
// gathers are data-dependent, so from global memory each would be a 64-address
// load), and reduces with cross-lane shuffles.  Inputs are read only (the numba
// temp_coh normalises its inputs in place; this one does not).
#include "tcoh.hpp"

namespace mr {

template <bool DS>
__global__ __launch_bounds__(256) void tcoh_kernel(const float2* __restrict__ c, const float2* __restrict__ v,
                                                   long long npts, int n, const int32_t* __restrict__ pairs,
                                                   int npairs, float* __restrict__ out) {
  extern __shared__ float2 tcoh_ph[];  // n phases per wave
  const int lane = threadIdx.x & 63;
  float2* ph = tcoh_ph + (threadIdx.x >> 6) * n;
  const long long w0 = ((long long)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const long long nw = ((long long)gridDim.x * blockDim.x) >> 6;
  for (long long pt = w0; pt < npts; pt += nw) {
    const float2* cp = c + pt * npairs;
    const float2* vp = v + pt * n;
    for (int i = lane; i < n; i += 64) ph[i] = vp[i];
    __builtin_amdgcn_wave_barrier();
    const float t = tcoh_point<DS>([&](int p, int, int) { return cp[p]; }, [&](int i) { return ph[i]; }, pairs,
                                   npairs, lane);
    if (lane == 0) out[pt] = t;
    __builtin_amdgcn_wave_barrier();  // every lane is past its reads before the next point's phases land
  }
}

// complex128 inputs: the same sums in FP64 (the reference's accumulator has
// the input's complex type), the result rounded to float32 as its output is.
template <bool DS>
__global__ __launch_bounds__(256) void tcoh64_kernel(const double2* __restrict__ c, const double2* __restrict__ v,
                                                     long long npts, int n, const int32_t* __restrict__ pairs,
                                                     int npairs, float* __restrict__ out) {
  extern __shared__ double2 tcoh_ph64[];
  const int lane = threadIdx.x & 63;
  double2* ph = tcoh_ph64 + (threadIdx.x >> 6) * n;
  const long long w0 = ((long long)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const long long nw = ((long long)gridDim.x * blockDim.x) >> 6;
  for (long long pt = w0; pt < npts; pt += nw) {
    const double2* cp = c + pt * npairs;
    const double2* vp = v + pt * n;
    for (int i = lane; i < n; i += 64) ph[i] = vp[i];
    __builtin_amdgcn_wave_barrier();
    double sr = 0.0, si = 0.0;
    for (int p = lane; p < npairs; p += 64) {
      const int a = pairs[2 * p], b = pairs[2 * p + 1];
      const double2 x = cp[p], va = ph[a], vb = ph[b];
      const double tr = va.x * vb.x + va.y * vb.y, ti = va.x * vb.y - va.y * vb.x;  // conj(va) vb
      const double den = DS ? sqrt(x.x * x.x + x.y * x.y)
                            : sqrt((x.x * x.x + x.y * x.y) * (va.x * va.x + va.y * va.y) * (vb.x * vb.x + vb.y * vb.y));
      sr += (x.x * tr - x.y * ti) / den;
      si += (x.x * ti + x.y * tr) / den;
    }
    sr = wave_sum(sr);
    si = wave_sum(si);
    if (lane == 0) out[pt] = (float)(sqrt(sr * sr + si * si) / (double)npairs);
    __builtin_amdgcn_wave_barrier();
  }
}

template <bool DS>
static int tcoh64_launch(const double* c, const double* v, long long n_pts, int n, const int32_t* pairs, int n_pairs,
                         float* out, mr_stream_t stream, const char* name) {
  MR_CHECK_ARG(n_pts >= 0 && n >= 1 && n_pairs >= 0, "%s: bad sizes", name);
  MR_CHECK_ARG(n <= 2048, "%s: n=%d above 2048 (the phases of a point are staged in LDS)", name, n);
  if (n_pts == 0) return MR_OK;
  MR_CHECK_ARG(c && v && pairs && out, "%s: null pointer", name);
  const int nb = (int)std::min<long long>(div_up((unsigned long long)n_pts, 4), 1 << 16);
  hipLaunchKernelGGL((tcoh64_kernel<DS>), dim3(nb), dim3(256), (size_t)4 * n * sizeof(double2), as_stream(stream),
                     reinterpret_cast<const double2*>(c), reinterpret_cast<const double2*>(v), n_pts, n, pairs,
                     n_pairs, out);
  return launch_status(name);
}

template <bool DS>
static int tcoh_launch(const float* c, const float* v, long long n_pts, int n, const int32_t* pairs, int n_pairs,
                       float* out, mr_stream_t stream, const char* name) {
  MR_CHECK_ARG(n_pts >= 0 && n >= 1 && n_pairs >= 0, "%s: bad sizes", name);
  MR_CHECK_ARG(n <= 4096, "%s: n=%d above 4096 (the phases of a point are staged in LDS)", name, n);
  if (n_pts == 0) return MR_OK;
  MR_CHECK_ARG(c && v && pairs && out, "%s: null pointer", name);
  const int nb = (int)std::min<long long>(div_up((unsigned long long)n_pts, 4), 1 << 16);
  hipLaunchKernelGGL((tcoh_kernel<DS>), dim3(nb), dim3(256), (size_t)4 * n * sizeof(float2), as_stream(stream),
                     reinterpret_cast<const float2*>(c), reinterpret_cast<const float2*>(v), n_pts, n, pairs,
                     n_pairs, out);
  return launch_status(name);
}

}  // namespace mr

extern "C" {
int mr_ds_temp_coh(const float* coh, const float* ph, long long n_pts, int n, const int32_t* pairs, int n_pairs,
                   float* t_coh, mr_stream_t stream) {
  return mr::tcoh_launch<true>(coh, ph, n_pts, n, pairs, n_pairs, t_coh, stream, "ds_temp_coh");
}
int mr_ds_temp_coh_c128(const double* coh, const double* ph, long long n_pts, int n, const int32_t* pairs,
                        int n_pairs, float* t_coh, mr_stream_t stream) {
  return mr::tcoh64_launch<true>(coh, ph, n_pts, n, pairs, n_pairs, t_coh, stream, "ds_temp_coh_c128");
}
int mr_temp_coh_c128(const double* intf, const double* rslc, long long n_pts, int n, const int32_t* pairs,
                     int n_pairs, float* t_coh, mr_stream_t stream) {
  return mr::tcoh64_launch<false>(intf, rslc, n_pts, n, pairs, n_pairs, t_coh, stream, "temp_coh_c128");
}
int mr_temp_coh(const float* intf, const float* rslc, long long n_pts, int n, const int32_t* pairs, int n_pairs,
                float* t_coh, mr_stream_t stream) {
  return mr::tcoh_launch<false>(intf, rslc, n_pts, n, pairs, n_pairs, t_coh, stream, "temp_coh");
}
}
