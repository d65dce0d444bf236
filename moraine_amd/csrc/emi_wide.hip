// EMI phase linking for n <= 32 and 64 < n <= 200 on gfx950 (moraine/pl.py:45-65):
// the drop-in kernel over coherence rows in global memory, one workgroup per
// point, running the per-point body of emi_wide.hpp.
#include "emi_wide.hpp"

namespace mr {

template <int R, int E, int GS, int GH, bool CHK = false>
__global__ __launch_bounds__(R == 2 ? 512 : 64 * EW_MAX_WAVES)
__attribute__((amdgpu_waves_per_eu(E <= 16 ? 4 : E <= 64 ? 3 : 1))) void emi_wide_kernel(
    const float2* __restrict__ coh, long long npts, int n, int ref, float2* __restrict__ ph,
    float* __restrict__ quality, double tau = 0.0, unsigned* __restrict__ flag_count = nullptr,
    int* __restrict__ flag_list = nullptr) {
  extern __shared__ __align__(16) unsigned char smem[];
  ew_lds_init<R * E>(smem);
  const int P = n * (n - 1) / 2;
  for (long long pt = blockIdx.x; pt < npts; pt += gridDim.x)
    ew_point<R, E, GS, GH, CHK>(EwPairRow{coh + pt * P, n}, n, ref, smem, EwDenseOut{ph, quality, pt, n}, tau,
                                flag_count, flag_list);
}

int emi_wide_launch(const float2* coh, long long npts, int n, int ref, float2* ph, float* quality, hipStream_t st,
                    double tau, unsigned* flag_count, int* flag_list) {
  const int nb = (int)std::min<long long>(npts, 1 << 20);
#define MR_EW(R, E, GS, GH)                                                                                       \
  if (flag_list)                                                                                                  \
    hipLaunchKernelGGL((emi_wide_kernel<R, E, GS, GH, true>), dim3(nb), dim3(((R) * n + 63) / 64 * 64),          \
                       ew_lds_bytes((R) * (E)), st, coh, npts, n, ref, ph, quality, tau, flag_count, flag_list);  \
  else                                                                                                            \
    hipLaunchKernelGGL((emi_wide_kernel<R, E, GS, GH>), dim3(nb), dim3(((R) * n + 63) / 64 * 64),                \
                       ew_lds_bytes((R) * (E)), st, coh, npts, n, ref, ph, quality)
  if (n <= 32) {
    MR_EW(2, 16, 8, 4);
  } else if (n <= 96) {
    MR_EW(2, 48, 4, 2);
  } else if (n <= 128) {
    MR_EW(2, 64, 4, 2);
  } else if (n <= 160) {
    MR_EW(2, 80, 8, 4);
  } else {
    MR_EW(2, 100, 4, 2);
  }
#undef MR_EW
  return launch_status("emi_wide");
}

}  // namespace mr
