// Smallest eigenpair of a real symmetric tridiagonal matrix held one entry per
// lane, shared by the one-wave EMI kernels (moraine/pl.py:45-65 restated: the
// eigenvector of the smallest eigenvalue, here of the tridiagonal form).
#pragma once
#include "emi_common.hpp"

namespace mr {

// Lane i holds d_i = T[i][i] and e_i = T[i+1][i] (e_{n-1} = 0; lanes >= n
// hold zeros).  Returns z_i, lane i's entry of the unit eigenvector of the
// smallest eigenvalue, and rq = z^T T z.  tde (64 x double2) and tee (64 x
// double) are LDS the wave owns; T is staged there for the serial recurrences.
// rounds: multisection rounds at most (each narrows the bracket 65-fold; they
// stop once it is at FP64 resolution).
__device__ __forceinline__ double emi_tri_smallest(int n, int lane, double my_d, double my_e, double2* tde,
                                                   double* tee, double& rq, int rounds = 5) {
  // ---- step 4a: smallest eigenvalue of T by multisection ------------------
  const double e_prev = dpp_d<DPP_WAVE_SHR1>(my_e);  // e_{i-1} (0 for lane 0)
  // T to LDS: the serial recurrences below read d_i, e_{i-1}^2 and e_i as one
  // uniform-address LDS load per step instead of two readlanes per double
  __builtin_amdgcn_wave_barrier();  // every lane is past its last read of these bytes
  tde[lane] = make_double2(my_d, e_prev * e_prev);
  tee[lane] = my_e;
  __builtin_amdgcn_wave_barrier();
  double lo, hi, scale;
  {
    const double off = e_prev + my_e;
    lo = wmin_d(lane < n ? my_d - off : 1e308);
    hi = wmax_d(lane < n ? my_d + off : -1e308);
    const double sc = wmax_d(lane < n ? fabs(my_d) + off : 0.0);
    scale = sc > 0.0 ? sc : 1.0;
    lo -= 1e-14 * scale;
    hi += 1e-14 * scale;
  }
  const double pivmin = 1e-300 + scale * scale * 1e-290;
  // five rounds narrow the Gershgorin bracket by 65^5 ~ 1.2e9, finer than the
  // FP32 reduction resolves the eigenvalue; the reported quality is the
  // Rayleigh quotient of the eigenvector, accurate to second order
  for (int round = 0; round < rounds; ++round) {
    if (hi - lo <= 2e-15 * scale) break;
    const double x = lo + (hi - lo) * (double)(lane + 1) * (1.0 / 65.0);
    // Sturm count of eigenvalues below x
    double q = tde[0].x - x;
    int cnt = q < 0.0;
    for (int i = 1; i < n; ++i) {
      if (fabs(q) < pivmin) q = -pivmin;
      double r = __builtin_amdgcn_rcp(q);
      r = fma(r, fma(-q, r, 1.0), r);  // one Newton step on the reciprocal
      const double2 t = tde[i];
      q = (t.x - x) - t.y * r;
      cnt += q < 0.0;
    }
    const unsigned long long m = __ballot(cnt >= 1);
    if (m == 0ull) {
      lo = rl_d(x, 63);
    } else {
      const int L = __ffsll((long long)m) - 1;
      const double xl = rl_d(x, L);
      const double xp = rl_d(x, L > 0 ? L - 1 : 0);
      hi = xl;
      if (L > 0) lo = xp;
    }
  }
  const double lam = 0.5 * (lo + hi);

  // ---- step 4b: eigenvector of T by a twisted factorisation ---------------
  // D+_i = (d_i - lam) - e_{i-1}^2 / D+_{i-1},  D-_i = (d_i - lam) - e_i^2 / D-_{i+1},
  // gamma_i = D+_i + D-_i - (d_i - lam); twist r = argmin |gamma|;
  // z_r = 1, z_i = -(e_i / D+_i) z_{i+1} (i < r), z_{i+1} = -(e_i / D-_{i+1}) z_i (i >= r).
  const double tiny = 2.2e-16 * scale + 1e-300;
  double my_dp = 0.0, my_dm = 0.0, my_rhi = 0.0;
  {
    double dp = tde[0].x - lam;
    if (fabs(dp) < tiny) dp = tiny;
    if (lane == 0) my_dp = dp;
    for (int i = 1; i < n; ++i) {
      const double2 t = tde[i];
      dp = (t.x - lam) - t.y * rcp_d(dp);
      if (fabs(dp) < tiny) dp = tiny;
      if (lane == i) my_dp = dp;
    }
    double dm = tde[n - 1].x - lam;
    if (fabs(dm) < tiny) dm = tiny;
    if (lane == n - 1) my_dm = dm;
    for (int i = n - 2; i >= 0; --i) {
      const double ei = tee[i];
      const double rd = rcp_d(dm);
      if (lane == i) my_rhi = -ei * rd;  // dm = D-_{i+1}
      dm = (tde[i].x - lam) - ei * ei * rd;
      if (fabs(dm) < tiny) dm = tiny;
      if (lane == i) my_dm = dm;
    }
  }
  const double gam = lane < n ? fabs(my_dp + my_dm - (my_d - lam)) : 1e308;
  const int rtw = wargmax_d(-gam, lane);
  const double my_rlo = -my_e / my_dp;
  double my_z = (lane == rtw) ? 1.0 : 0.0;
  {
    double z = 1.0;
    for (int i = rtw - 1; i >= 0; --i) {
      z *= rl_d(my_rlo, i);
      if (lane == i) my_z = z;
    }
    z = 1.0;
    for (int i = rtw + 1; i < n; ++i) {
      z *= rl_d(my_rhi, i - 1);
      if (lane == i) my_z = z;
    }
  }
  my_z *= 1.0 / sqrt(wsum_d(my_z * my_z));
  // Rayleigh quotient of T for the reported eigenvalue
  const double z_prev = dpp_d<DPP_WAVE_SHR1>(my_z);
  const double z_next = dpp_d<DPP_WAVE_SHL1>(my_z);
  const double tz = lane < n ? my_d * my_z + e_prev * z_prev + my_e * z_next : 0.0;
  rq = wsum_d(my_z * tz);
  return my_z;
}

}  // namespace mr
