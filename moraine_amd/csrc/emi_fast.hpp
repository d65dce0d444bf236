// One-wave EMI for n <= 64 (moraine/pl.py:45-65 restated; algorithm notes at
// the top of emi.hip): shared by the EMI kernel and the fused batched kernel.
#pragma once
#include "emi_tri.hpp"

namespace mr {

// One wave, n <= NM <= 64: the smallest eigenpair of M = inv(|Gamma|) o Gamma
// for one point.  coh_at(a, b), a < b, returns the coherence of pair (a, b).
// On return lane i holds (vr, vi) = v_i, the eigenvector before the reference
// rotation, and rq the Rayleigh quotient (smallest eigenvalue).  lds:
// EMI_FAST_LDS_BYTES of 16-byte aligned LDS owned by this wave.
constexpr int EMI_FAST_LDS_BYTES = 1536;

// CHECK: also decide whether the point needs the FP64 re-solve (*flag, from
// tau; see the end of the function).
template <int NM, bool CHECK = false, class Src>
__device__ __forceinline__ void emi_fast_point(const Src& coh_at, int n, int lane, unsigned char* lds, float& vr,
                                               float& vi, double& rq, double tau = 0.0, bool* flag = nullptr) {
  double* colb = reinterpret_cast<double*>(lds);        // 64: sweep column broadcast
  float2* ub = reinterpret_cast<float2*>(lds + 512);    // 64: reflector u
  float2* qb = ub + 64;                                 // 64: q
  // after the reduction the same bytes serve the uniform reads of T
  double2* tde = reinterpret_cast<double2*>(lds);       // 64: (d_i, e_{i-1}^2)
  double* tee = reinterpret_cast<double*>(lds + 1024);  // 64: e_i
  // ---- step 1: A = |Gamma| (row per lane, FP64), sweep inverse ----------
  RegArr<double, NM> a;
#pragma unroll
  for (int j = 0; j < NM; ++j) {
    double v = 0.0;
    if (lane < n && j < n) {
      if (j == lane) {
        v = 1.0;
      } else {
        const float2 g = coh_at(j > lane ? lane : j, j > lane ? j : lane);
        v = sqrt((double)g.x * g.x + (double)g.y * g.y);
      }
    }
    a.set(j, v);
    if ((j & 7) == 7) __builtin_amdgcn_sched_barrier(0);
  }
  double dg = lane < n ? 1.0 : 0.0;  // own diagonal entry, tracked through the sweep
  bool swept = lane >= n;
  for (int s = 0; s < n; ++s) {
    int k = 63 - (wmax_i(pivot_key<6>(fabs(dg), !swept, lane)) & 63);  // unswept row, largest |diagonal|
    __asm__ volatile("" : "+s"(k));  // an opaque index keeps the row in VGPRs (M0-relative read, no scratch copy)
    const double c = a.dyn(k);
    // Broadcasting d-1 in place of the pivot d lets the rank-1 update itself
    // write column k:  c - (c/d)(d-1) = c/d.  The pivot lane's own diagonal
    // comes out as 2 - 1/d instead of -1/d; that entry is never read again
    // by the sweep, so the +2 offset is removed once, when M is formed.
    colb[lane] = (lane == k) ? c - 1.0 : c;
    __builtin_amdgcn_wave_barrier();
    const double dk = rl_d(c, k);
    const double rd = rcp_d(dk);
    const double f = (lane == k) ? (1.0 - rd) : c * rd;
    {
      // software-pipelined broadcast of colb: group g+1 is loaded before the
      // FMAs of group g, so LDS latency overlaps the FP64 FMAs
      double cur[8], nxt[8];
#pragma unroll
      for (int t = 0; t < 8; ++t) cur[t] = colb[t];
#pragma unroll
      for (int g = 0; g < NM / 8; ++g) {
        if (g + 1 < NM / 8) {
#pragma unroll
          for (int t = 0; t < 8; ++t) nxt[t] = colb[(g + 1) * 8 + t];
        }
#pragma unroll
        for (int t = 0; t < 8; ++t) a.set(g * 8 + t, fma(-f, cur[t], a.get(g * 8 + t)));
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int t = 0; t < 8; ++t) cur[t] = nxt[t];
      }
    }
    dg = fma(-f, c, dg);
    swept = swept || (lane == k);
    __builtin_amdgcn_wave_barrier();
  }
  // a now holds -inv(A) row `lane` (diagonal entry offset by +2)

  // ---- step 2: M = inv(A) o Gamma (FP32 complex row per lane) ------------
  RegArr<float, NM> mr_, mi_;
#pragma unroll
  for (int j = 0; j < NM; ++j) {
    float re = 0.f, im = 0.f;
    if (lane < n && j < n) {
      const double ai = (j == lane) ? 2.0 - a.get(j) : -a.get(j);  // undo the sweep's +2 on the diagonal
      if (j == lane) {
        re = (float)ai;
      } else if (j > lane) {
        const float2 g = coh_at(lane, j);
        re = (float)(ai * g.x);
        im = (float)(ai * g.y);
      } else {
        const float2 g = coh_at(j, lane);
        re = (float)(ai * g.x);
        im = (float)(-ai * g.y);
      }
    }
    mr_.set(j, re);
    mi_.set(j, im);
    if ((j & 7) == 7) __builtin_amdgcn_sched_barrier(0);
  }

  // ---- step 3: Householder tridiagonalisation -----------------------------
  // T is kept distributed: lane i holds d_i = T[i][i], e_i = |T[i+1][i]|, the
  // unitary scaling D_i, and (alpha_i, 1/den_i) of reflector i.
  double my_d = 0.0, my_e = 0.0;
  float my_ar = 0.f, my_ai = 0.f, my_hv = 0.f;
  float my_dr = 1.f, my_di = 0.f;
  float dcr = 1.f, dci = 0.f;  // running phase product D_k (uniform)
  // Steps 0..30 run one lane per row.  From step 31 on the active block lies in
  // rows/columns 32..63 and moves to two lanes per row (late_steps below), so
  // the lanes of reduced rows stop issuing half of the remaining work.
  constexpr bool kSplit = NM == 64;
  const bool late = kSplit && n > 33;
  const int k1 = late ? 31 : n - 2;
  for (int k = 0; k < k1; ++k) {
    const float xr_raw = mr_.dyn(k);
    const float xi_raw = mi_.dyn(k);
    const bool below = lane > k && lane < n;
    const float xr = below ? xr_raw : 0.f, xi = below ? xi_raw : 0.f;
    const float nrm2 = wsum_f(xr * xr + xi * xi);
    const float x0r = rl_f(xr, k + 1), x0i = rl_f(xi, k + 1);
    const float dkk = rl_f(xr_raw, k);
    const float nx = __builtin_amdgcn_sqrtf(nrm2);
    const float ax0 = __builtin_amdgcn_sqrtf(x0r * x0r + x0i * x0i);
    const float rax0 = ax0 > 0.f ? __builtin_amdgcn_rcpf(ax0) : 0.f;
    const float phr = ax0 > 0.f ? x0r * rax0 : 1.f, phi = ax0 > 0.f ? x0i * rax0 : 0.f;
    const float alr = -phr * nx, ali = -phi * nx;  // alpha = -phase * ||x||
    const float den2 = 2.f * nx * (nx + ax0);
    const float inv = den2 > 0.f ? __builtin_amdgcn_rsqf(den2) : 0.f;
    float ur = (lane == k + 1) ? (x0r - alr) : xr;
    float ui = (lane == k + 1) ? (x0i - ali) : xi;
    ur *= inv;
    ui *= inv;
    // column k keeps x (rows > k are never updated again); the reflector is
    // rebuilt from x, alpha and 1/den during the back-transformation
    if (lane == k) {
      my_d = dkk;
      my_e = nx;
      my_ar = alr;
      my_ai = ali;
      my_hv = inv;
    }
    {
      const float sr = nx > 0.f ? -phr : 1.f, si = nx > 0.f ? -phi : 0.f;  // alpha/|alpha|
      const float r = dcr * sr - dci * si, i = dcr * si + dci * sr;
      dcr = r;
      dci = i;
    }
    if (lane == k + 1) {
      my_dr = dcr;
      my_di = dci;
    }
    ub[lane] = make_float2(ur, ui);
    __builtin_amdgcn_wave_barrier();
    // p = M22 u  (4-column groups left of the active block or beyond n are skipped)
    float pr = 0.f, pi = 0.f;
    {
      float2 cu[4], nu[4];
#pragma unroll
      for (int t = 0; t < 4; ++t) cu[t] = ub[t];
#pragma unroll
      for (int g = 0; g < NM / 4; ++g) {
        const bool act = (g * 4 + 3 > k) && (g * 4 < n);
        if (g + 1 < NM / 4) {
#pragma unroll
          for (int t = 0; t < 4; ++t) nu[t] = ub[(g + 1) * 4 + t];
        }
        if (act) {
#pragma unroll
          for (int t = 0; t < 4; ++t) {
            const int j = g * 4 + t;
            pr = fmaf(mr_.get(j), cu[t].x, pr);
            pr = fmaf(-mi_.get(j), cu[t].y, pr);
            pi = fmaf(mr_.get(j), cu[t].y, pi);
            pi = fmaf(mi_.get(j), cu[t].x, pi);
          }
        }
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int t = 0; t < 4; ++t) cu[t] = nu[t];
      }
    }
    const float cval = wsum_f(ur * pr + ui * pi);  // real(u^H p)
    const float qr = below ? 2.f * (pr - cval * ur) : 0.f;
    const float qi = below ? 2.f * (pi - cval * ui) : 0.f;
    qb[lane] = make_float2(qr, qi);
    __builtin_amdgcn_wave_barrier();
    // M22 -= u q^H + q u^H
    {
      // opaque copies of k and n: the group tests are recomputed here as scalar
      // compares instead of being carried over from the product loop as lane
      // masks (a VGPR select and compare per group)
      int ku = k, nu_ = n;
      __asm__ volatile("" : "+s"(ku), "+s"(nu_));
      float2 cu[4], cq[4], nu[4], nq[4];
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        cu[t] = ub[t];
        cq[t] = qb[t];
      }
#pragma unroll
      for (int g = 0; g < NM / 4; ++g) {
        const bool act = (g * 4 + 3 > ku) && (g * 4 < nu_);
        if (g + 1 < NM / 4) {
#pragma unroll
          for (int t = 0; t < 4; ++t) {
            nu[t] = ub[(g + 1) * 4 + t];
            nq[t] = qb[(g + 1) * 4 + t];
          }
        }
        if (act) {
#pragma unroll
          for (int t = 0; t < 4; ++t) {
            const int j = g * 4 + t;
            float r = mr_.get(j), i = mi_.get(j);
            r = fmaf(-ur, cq[t].x, r);
            r = fmaf(-ui, cq[t].y, r);
            i = fmaf(ur, cq[t].y, i);
            i = fmaf(-ui, cq[t].x, i);
            r = fmaf(-qr, cu[t].x, r);
            r = fmaf(-qi, cu[t].y, r);
            i = fmaf(qr, cu[t].y, i);
            i = fmaf(-qi, cu[t].x, i);
            mr_.set(j, r);
            mi_.set(j, i);
          }
        }
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int t = 0; t < 4; ++t) {
          cu[t] = nu[t];
          cq[t] = nq[t];
        }
      }
    }
    __builtin_amdgcn_wave_barrier();
  }
  const int hh = lane >> 5, bb = lane & 31;  // late steps: block row 32 + bb, columns 32 + 2s + hh
  if (late) {
    // Step 31's column (rows 32..63) stays in register 31; then lane (bb, hh)
    // takes M[32 + bb][32 + 2s + hh] into register 32 + s (s = 0..15): odd
    // columns are the lane's own, even ones come from lane bb + 32.  Registers
    // 32..47 of rows 0..31 are dead by now (only columns k of rows > k are
    // read again, by the back-transformation).
#pragma unroll
    for (int s2 = 0; s2 < 16; ++s2) {
      const float er = swap_half_f(mr_.get(32 + 2 * s2)), ei = swap_half_f(mi_.get(32 + 2 * s2));
      const float nr = hh ? mr_.get(33 + 2 * s2) : er, ni = hh ? mi_.get(33 + 2 * s2) : ei;
      mr_.set(32 + s2, nr);
      mi_.set(32 + s2, ni);
    }
    float x31r = mr_.get(31), x31i = mi_.get(31);
    {
      const float pr = swap_half_f(x31r), pi = swap_half_f(x31i);  // lane bb takes row 32 + bb's
      x31r = hh ? x31r : pr;
      x31i = hh ? x31i : pi;
    }
    const float d31 = rl_f(mr_.get(31), 31);  // M[31][31] (row 31 never moved)
    for (int k = 31; k + 2 < n; ++k) {
      // column k, this lane's row R = 32 + bb (register 32 + (k-32)/2 on half (k-32)&1)
      float xr_raw, xi_raw, dkk;
      if (k == 31) {
        xr_raw = x31r;
        xi_raw = x31i;
        dkk = d31;
      } else {
        const int s2 = (k - 32) >> 1, hk = (k - 32) & 1;
        const float o_r = mr_.dyn(32 + s2), o_i = mi_.dyn(32 + s2);
        const float p_r = swap_half_f(o_r), p_i = swap_half_f(o_i);
        xr_raw = hh == hk ? o_r : p_r;
        xi_raw = hh == hk ? o_i : p_i;
        dkk = rl_f(xr_raw, k - 32);
      }
      const int R = 32 + bb;
      const bool below = R > k && R < n;
      const float xr = below ? xr_raw : 0.f, xi = below ? xi_raw : 0.f;
      const float nrm2 = wsum_f(hh == 0 ? xr * xr + xi * xi : 0.f);
      const float x0r = rl_f(xr, k + 1 - 32), x0i = rl_f(xi, k + 1 - 32);
      const float nx = __builtin_amdgcn_sqrtf(nrm2);
      const float ax0 = __builtin_amdgcn_sqrtf(x0r * x0r + x0i * x0i);
      const float rax0 = ax0 > 0.f ? __builtin_amdgcn_rcpf(ax0) : 0.f;
      const float phr = ax0 > 0.f ? x0r * rax0 : 1.f, phi = ax0 > 0.f ? x0i * rax0 : 0.f;
      const float alr = -phr * nx, ali = -phi * nx;
      const float den2 = 2.f * nx * (nx + ax0);
      const float inv = den2 > 0.f ? __builtin_amdgcn_rsqf(den2) : 0.f;
      float ur = (R == k + 1) ? (x0r - alr) : xr;
      float ui = (R == k + 1) ? (x0i - ali) : xi;
      ur *= inv;
      ui *= inv;
      if (lane == k) {
        my_d = dkk;
        my_e = nx;
        my_ar = alr;
        my_ai = ali;
        my_hv = inv;
      }
      {
        const float sr = nx > 0.f ? -phr : 1.f, si = nx > 0.f ? -phi : 0.f;
        const float r = dcr * sr - dci * si, i = dcr * si + dci * sr;
        dcr = r;
        dci = i;
      }
      if (lane == k + 1) {
        my_dr = dcr;
        my_di = dci;
      }
      if (hh == 0) ub[R] = make_float2(ur, ui);
      __builtin_amdgcn_wave_barrier();
      // p = M22 u: each lane sums its eight interleaved columns per group, the
      // two lanes of a row then add their partial sums
      float pr = 0.f, pi = 0.f;
#pragma unroll
      for (int g = 0; g < 4; ++g) {
        const bool act = 32 + 2 * (4 * g + 3) + 1 > k;
        if (act) {
#pragma unroll
          for (int t = 0; t < 4; ++t) {
            const int s2 = 4 * g + t;
            const float2 cu = ub[32 + 2 * s2 + hh];
            pr = fmaf(mr_.get(32 + s2), cu.x, pr);
            pr = fmaf(-mi_.get(32 + s2), cu.y, pr);
            pi = fmaf(mr_.get(32 + s2), cu.y, pi);
            pi = fmaf(mi_.get(32 + s2), cu.x, pi);
          }
        }
        __builtin_amdgcn_sched_barrier(0);
      }
      pr += swap_half_f(pr);
      pi += swap_half_f(pi);
      const float cval = wsum_f(hh == 0 ? ur * pr + ui * pi : 0.f);  // real(u^H p)
      const float qr = below ? 2.f * (pr - cval * ur) : 0.f;
      const float qi = below ? 2.f * (pi - cval * ui) : 0.f;
      if (hh == 0) qb[R] = make_float2(qr, qi);
      __builtin_amdgcn_wave_barrier();
      // M22 -= u q^H + q u^H over this lane's columns (group tests on an
      // opaque k, as in the steps before the split)
      int ku = k;
      __asm__ volatile("" : "+s"(ku));
#pragma unroll
      for (int g = 0; g < 4; ++g) {
        const bool act = 32 + 2 * (4 * g + 3) + 1 > ku;
        if (act) {
#pragma unroll
          for (int t = 0; t < 4; ++t) {
            const int s2 = 4 * g + t;
            const float2 cu = ub[32 + 2 * s2 + hh], cq = qb[32 + 2 * s2 + hh];
            float r = mr_.get(32 + s2), i = mi_.get(32 + s2);
            r = fmaf(-ur, cq.x, r);
            r = fmaf(-ui, cq.y, r);
            i = fmaf(ur, cq.y, i);
            i = fmaf(-ui, cq.x, i);
            r = fmaf(-qr, cu.x, r);
            r = fmaf(-qi, cu.y, r);
            i = fmaf(qr, cu.y, i);
            i = fmaf(-qi, cu.x, i);
            mr_.set(32 + s2, r);
            mi_.set(32 + s2, i);
          }
        }
        __builtin_amdgcn_sched_barrier(0);
      }
      __builtin_amdgcn_wave_barrier();
    }
  }
  {
    // trailing 2x2 block
    float r0, i0, r1;
    int l0, l1;  // lanes holding rows n-2 / n-1 of column n-2 (r0, i0) and of column n-1 (r1)
    if (late) {
      const int c0 = n - 2 - 32, c1 = n - 1 - 32;
      r0 = mr_.dyn(32 + (c0 >> 1));
      i0 = mi_.dyn(32 + (c0 >> 1));
      r1 = mr_.dyn(32 + (c1 >> 1));
      l0 = 32 * (c0 & 1) - 32;  // + row
      l1 = 32 * (c1 & 1) - 32;
    } else {
      r0 = mr_.dyn(n - 2);
      i0 = mi_.dyn(n - 2);
      r1 = mr_.dyn(n - 1);
      l0 = 0;
      l1 = 0;
    }
    const float dm2 = rl_f(r0, n - 2 + l0);
    const float dm1 = rl_f(r1, n - 1 + l1);
    const float er = rl_f(r0, n - 1 + l0), ei = rl_f(i0, n - 1 + l0);
    const float ae = sqrtf(er * er + ei * ei);
    if (lane == n - 2) {
      my_d = dm2;
      my_e = ae;
    }
    if (lane == n - 1) {
      my_d = dm1;
      my_e = 0.0;
      const float sr = ae > 0.f ? er / ae : 1.f, si = ae > 0.f ? ei / ae : 0.f;
      my_dr = dcr * sr - dci * si;
      my_di = dcr * si + dci * sr;
    }
  }

  // ---- step 4: smallest eigenpair of T ------------------------------------
  const double my_z = emi_tri_smallest(n, lane, my_d, my_e, tde, tee, rq);

  // ---- back-transformation v = H_0 ... H_{n-3} D z ------------------------
  vr = 0.f;
  vi = 0.f;
  if (lane < n) {
    const float zi = (float)my_z;
    vr = my_dr * zi;
    vi = my_di * zi;
  }
  for (int k = n - 3; k >= 0; --k) {
    float xr, xi;
    if (late && k >= 32) {
      // column k of the late block: lane i >= 32 (row i) owns it on half 1,
      // else it sits on lane i - 32
      const int s2 = (k - 32) >> 1;
      const float o_r = mr_.dyn(32 + s2), o_i = mi_.dyn(32 + s2);
      const float p_r = swap_half_f(o_r), p_i = swap_half_f(o_i);
      xr = (k - 32) & 1 ? o_r : p_r;
      xi = (k - 32) & 1 ? o_i : p_i;
    } else {
      xr = mr_.dyn(k);
      xi = mi_.dyn(k);
    }
    const bool below = lane > k && lane < n;
    const float alr = rl_f(my_ar, k), ali = rl_f(my_ai, k), hv = rl_f(my_hv, k);
    const float ar = below ? ((lane == k + 1) ? xr - alr : xr) * hv : 0.f;
    const float ai = below ? ((lane == k + 1) ? xi - ali : xi) * hv : 0.f;
    // s = u^H v
    const float sr = wsum_f(ar * vr + ai * vi);
    const float si = wsum_f(ar * vi - ai * vr);
    vr -= 2.f * (ar * sr - ai * si);
    vi -= 2.f * (ar * si + ai * sr);
  }
  if constexpr (CHECK) {
    // The FP32 stages perturb M by about eps32 ||T||, which moves the unit
    // eigenvector by eps32 ||T|| / gap and the phase of entry i by that over
    // |v_i|.  Flag the point when eps32 ||T|| c(n) / (gap min_i |v_i|) > tau,
    // c(n) = n / (60 sqrt 60) (measured against FP64: the worst errors sit
    // near 1/sqrt(60) of the bare estimate at n = 60 and grow with n), i.e.
    // when T (still in LDS) has a second eigenvalue below lambda + delta,
    // delta = eps32 ||T|| c(n) / (tau min_i |v_i|): one Sturm count.
    const double av = lane < n ? (double)sqrtf(vr * vr + vi * vi) : 1e300;
    const double vmin = fmax(wmin_d(av), 1e-300);
    const double tn = wmax_d(lane < n ? fabs(my_d) : 0.0) + 2.0 * wmax_d(lane < n ? fabs(my_e) : 0.0);
    const double x = rq + 5.9604644775390625e-8 * tn * emi_flag_scale(n) / (tau * vmin);
    const double pivmin = 1e-300 + tn * tn * 1e-290;
    double q = tde[0].x - x;
    int cnt = q < 0.0;
    for (int i = 1; i < n; ++i) {
      if (fabs(q) < pivmin) q = -pivmin;
      double r = __builtin_amdgcn_rcp(q);
      r = fma(r, fma(-q, r, 1.0), r);
      const double2 t = tde[i];
      q = (t.x - x) - t.y * r;
      cnt += q < 0.0;
    }
    *flag = cnt >= 2 && isfinite(x);
  }
}

// Steps 1-2 of emi_fast_point for the shift-invert path: M = inv(|Gamma|) o
// Gamma into mr_, mi_ (row `lane`).  Returns false, as soon as a sweep pivot is
// not positive, when |Gamma| is not positive definite: the sweep pivots are
// those of a symmetric elimination with symmetric pivoting, so their signs are
// |Gamma|'s inertia (Sylvester); all positive <=> PD.
constexpr int EMI_SI_GD = 8;  // FP64 sweep: columns per software-pipelined LDS group
template <int NM, class Src>
__device__ __forceinline__ bool emi_fast_form_m(const Src& coh_at, int n, int lane, unsigned char* lds,
                                                RegArr<float, NM>& mr_, RegArr<float, NM>& mi_,
                                                double* __restrict__ xs = nullptr) {
  double* colb = reinterpret_cast<double*>(lds);        // 64: sweep column broadcast
  // ---- step 1: A = |Gamma| (row per lane, FP64), sweep inverse ----------
  RegArr<double, NM> a;
#pragma unroll
  for (int j = 0; j < NM; ++j) {
    double v = 0.0;
    if (lane < n && j < n) {
      if (j == lane) {
        v = 1.0;
      } else {
        const float2 g = coh_at(j > lane ? lane : j, j > lane ? j : lane);
        v = sqrt((double)g.x * g.x + (double)g.y * g.y);
      }
    }
    a.set(j, v);
    if ((j & 7) == 7) __builtin_amdgcn_sched_barrier(0);
  }
  // Pivots in natural order: this pass only takes a PD |Gamma|, for which the
  // symmetric elimination needs no pivoting, and the first non-positive pivot
  // (in any order) shows that |Gamma| is not PD (Sylvester: the pivots of the
  // natural order are ratios of leading principal minors).
  for (int s = 0; s < n; ++s) {
    int k = s;
    __asm__ volatile("" : "+s"(k));  // an opaque index keeps the row in VGPRs (M0-relative read, no scratch copy)
    const double c = a.dyn(k);
    // Broadcasting d-1 in place of the pivot d lets the rank-1 update itself
    // write column k:  c - (c/d)(d-1) = c/d.  The pivot lane's own diagonal
    // comes out as 2 - 1/d instead of -1/d; that entry is never read again
    // by the sweep, so the +2 offset is removed once, when M is formed.
    colb[lane] = (lane == k) ? c - 1.0 : c;
    __builtin_amdgcn_wave_barrier();
    const double dk = rl_d(c, k);
    if (!(dk > 0.0)) return false;  // |Gamma| not PD: the point goes to the Householder pass
    const double rd = rcp_d(dk);
    const double f = (lane == k) ? (1.0 - rd) : c * rd;
    {
      // software-pipelined broadcast of colb: group g+1 is loaded before the
      // FMAs of group g, so LDS latency overlaps the FP64 FMAs
      // (groups of GD columns; a last partial group when NM is not a multiple)
      constexpr int GD = EMI_SI_GD, NG = (NM + GD - 1) / GD;
      double cur[GD], nxt[GD];
#pragma unroll
      for (int t = 0; t < GD; ++t) cur[t] = colb[t];
#pragma unroll
      for (int g = 0; g < NG; ++g) {
        if (g + 1 < NG) {
#pragma unroll
          for (int t = 0; t < GD; ++t)
            if ((g + 1) * GD + t < NM) nxt[t] = colb[(g + 1) * GD + t];
        }
#pragma unroll
        for (int t = 0; t < GD; ++t)
          if (g * GD + t < NM) a.set(g * GD + t, fma(-f, cur[t], a.get(g * GD + t)));
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int t = 0; t < GD; ++t) cur[t] = nxt[t];
      }
    }
    __builtin_amdgcn_wave_barrier();
  }
  // a now holds -inv(A) row `lane` (diagonal entry offset by +2)
  if (xs != nullptr) {
    // X = inv(A) in FP64 by packed columns for the eigen-residual check: entry
    // (i, j), i <= j, at j(j+1)/2 + i, so each store covers consecutive lanes
#pragma unroll
    for (int j = 0; j < NM; ++j)
      if (lane <= j && j < n) xs[j * (j + 1) / 2 + lane] = (j == lane) ? 2.0 - a.get(j) : -a.get(j);
  }

  // ---- step 2: M = inv(A) o Gamma (FP32 complex row per lane) ------------
#pragma unroll
  for (int j = 0; j < NM; ++j) {
    float re = 0.f, im = 0.f;
    if (lane < n && j < n) {
      const double ai = (j == lane) ? 2.0 - a.get(j) : -a.get(j);  // undo the sweep's +2 on the diagonal
      if (j == lane) {
        re = (float)ai;
      } else if (j > lane) {
        const float2 g = coh_at(lane, j);
        re = (float)(ai * g.x);
        im = (float)(ai * g.y);
      } else {
        const float2 g = coh_at(j, lane);
        re = (float)(ai * g.x);
        im = (float)(-ai * g.y);
      }
    }
    mr_.set(j, re);
    mi_.set(j, im);
    if ((j & 7) == 7) __builtin_amdgcn_sched_barrier(0);
  }
  return true;
}

// Shift-invert formulation (|Gamma| PD): B = M - sigma I with sigma = 0.98
// below the smallest eigenvalue (M is near a closure-free matrix, whose
// smallest eigenvalue is 1), X = inv(B) by the Hermitian sweep in FP32 (the
// pivots certify B PD: one that is not positive means sigma >= lambda_min and
// the point goes back to the Householder path), then inverse iteration
// x <- X x / |X x| from the closure-free guess x_j = conj(Gamma_0j)/|Gamma_0j|
// until successive unit iterates differ by less than EMI_INV_TOL in every
// entry (at least EMI_INV_MIN, at most EMI_INV_MAX sweeps; not converged ->
// Householder path).  lambda_min = sigma + 1 / (x^H X x).  Returns false when
// the point must take the Householder path (M in mr_, mi_ is destroyed).
constexpr float EMI_INV_SIGMA = 0.98f;
constexpr float EMI_INV_TOL = 4e-7f;
constexpr int EMI_INV_MIN = 3;
constexpr int EMI_INV_MAX = 16;

__device__ __forceinline__ float wmax_f(float v) {
  v = fmaxf(v, dpp_f<DPP_XOR1>(v));
  v = fmaxf(v, dpp_f<DPP_XOR2>(v));
  v = fmaxf(v, dpp_f<DPP_HALF_MIRROR>(v));
  v = fmaxf(v, dpp_f<DPP_MIRROR>(v));
  return fmaxf(fmaxf(rl_f(v, 0), rl_f(v, 16)), fmaxf(rl_f(v, 32), rl_f(v, 48)));
}

// y = (X o Gamma) x in FP64, lane i: y_i.  xs: X = inv(|Gamma|) in FP64 by
// packed columns (emi_fast_form_m); xb: x broadcast (FP64); (x_r, x_i): this
// lane's entry.  Column j of the packed X covers rows i <= j; entry (i, j),
// i < j, also enters row j as X_ij conj(Gamma_ij) x_i (one wave sum per
// column).  One column at a time, the next column's load in flight (the
// caller holds the FP32 inverse in registers meanwhile).
template <class Src>
__device__ __forceinline__ void emi_m64_product(const Src& coh_at, int n, int lane, const double* __restrict__ xs,
                                                const double2* xb, double x_r, double x_i, double& yr, double& yi) {
  yr = 0.0;
  yi = 0.0;
  double a_next = lane == 0 ? xs[0] : 0.0;
#pragma unroll 1
  for (int j = 0; j < n; ++j) {
    const double a = a_next;
    const int jn = j + 1;
    a_next = (jn < n && lane <= jn) ? xs[jn * (jn + 1) / 2 + lane] : 0.0;
    const double2 xj = xb[j];
    double gr = 0.0, gi = 0.0;
    if (lane == j) {
      gr = 1.0;
    } else if (lane < j) {
      const float2 g = coh_at(lane, j);
      gr = g.x;
      gi = g.y;
    }
    yr = fma(a, gr * xj.x - gi * xj.y, yr);
    yi = fma(a, gr * xj.y + gi * xj.x, yi);
    const bool lo = lane < j;
    const double sr = wsum_d(lo ? a * (gr * x_r + gi * x_i) : 0.0);
    const double si = wsum_d(lo ? a * (gr * x_i - gi * x_r) : 0.0);
    if (lane == j) {
      yr += sr;
      yi += si;
    }
  }
}

// REF (the precision form): after the iteration converges, FP64 refinement
// steps through the same product.  Each forms the FP64 residual
// r = M x - lambda x (M = X o Gamma with X = inv(|Gamma|) in FP64 from xs,
// emi_fast_form_m; lambda = x^H M x) and the correction d = P X' r, X' the
// FP32 inverse of B and P the projector off x.  To first order the error of x
// is inv(M - lambda_1 I) r on the complement of x; X' shrinks each of its
// components by at most 1 - rho (rho: the iteration's contraction per sweep,
// estimated from its last changes), so the linked phase error is estimated as
// 2 max_i |d_i| / |x_i| / (1 - rho).  At or below tau on the first step the
// point keeps the plain outputs; above, x <- x - d until the estimate falls
// below tau / 16 (at most EMI_REF_STEPS corrections; lambda is then the FP64
// Rayleigh quotient).  *flag: no contraction estimate or no convergence, the
// point goes to the FP64 re-solve.
constexpr int EMI_REF_STEPS = 3;

template <int NM, bool REF = false, class Src>
__device__ __forceinline__ bool emi_shift_invert(const Src& coh_at, int n, int lane, unsigned char* lds,
                                                 RegArr<float, NM>& br, RegArr<float, NM>& bi, float& vr, float& vi,
                                                 double& rq, const double* __restrict__ xs = nullptr,
                                                 double tau = 0.0, bool* flag = nullptr) {
  float2* colb = reinterpret_cast<float2*>(lds);        // 64: sweep column broadcast
  float2* xb = reinterpret_cast<float2*>(lds + 512);    // 64: iterate broadcast
  // B = M - sigma I, and this lane's diagonal entry (real) for the pivot order
  float dg = 0.f;
#pragma unroll
  for (int j = 0; j < NM; ++j) {
    const bool dj = j == lane && lane < n;
    const float b = br.get(j) - (dj ? EMI_INV_SIGMA : 0.f);
    br.set(j, b);
    if (dj) dg = b;
  }
  // ---- X = inv(B): Hermitian sweep --------------------------------------
  // Row i:  B_ij -= f_i conj(B_jk) (row k = conj(column k)), f_i = B_ik / d,
  // d = B_kk.  Broadcasting d - 1 in place of the pivot lets the update write
  // column k itself (B_ik - f_i (d - 1) = B_ik / d); the pivot lane uses
  // f = 1 - 1/d, so its row becomes B_kj / d and its diagonal 2 - 1/d (the
  // +2 is undone in the products below).  After all steps B holds -inv(B)
  // off the diagonal and 2 - inv(B)_ii on it.
  // Static pivot order: the rows by decreasing diagonal of B (ties: smaller
  // row first), from one bitonic sort across the wave; lane s then holds the
  // row of the s-th pivot.  (B is PD here, so every diagonal entry is
  // positive and its float bits order as the values do.)
  unsigned kh = lane < n ? ~__float_as_uint(dg) : 0xffffffffu;  // ascending key: decreasing diagonal
  int kl = lane;
#pragma unroll
  for (int size = 2; size <= 64; size <<= 1) {
#pragma unroll
    for (int j = size >> 1; j > 0; j >>= 1) {
      const unsigned oh = (unsigned)__shfl_xor((int)kh, j, 64);
      const int ol = __shfl_xor(kl, j, 64);
      const bool other_less = oh < kh || (oh == kh && ol < kl);
      const bool keep_min = ((lane & size) == 0) == ((lane & j) == 0);
      if (keep_min == other_less) {
        kh = oh;
        kl = ol;
      }
    }
  }
  for (int s = 0; s < n; ++s) {
    int k = __builtin_amdgcn_readlane(kl, s);
    __asm__ volatile("" : "+s"(k));  // an opaque index keeps the row in VGPRs (M0-relative read, no scratch copy)
    const float cr = br.dyn(k), ci = bi.dyn(k);  // B[lane][k]
    const float d = rl_f(cr, k);
    if (!(d > 0.f)) return false;  // sigma >= lambda_min (or B singular): Householder path
    colb[lane] = (lane == k) ? make_float2(cr - 1.f, 0.f) : make_float2(cr, ci);
    __builtin_amdgcn_wave_barrier();
    float rd = __builtin_amdgcn_rcpf(d);
    rd = fmaf(rd, fmaf(-d, rd, 1.f), rd);
    const float fr = (lane == k) ? 1.f - rd : cr * rd;
    const float fi = (lane == k) ? 0.f : ci * rd;
    int nu_ = n;
    __asm__ volatile("" : "+s"(nu_));  // column groups at or beyond n hold zeros: skipped by a scalar test
    {
      float2 cur[4], nxt[4];
#pragma unroll
      for (int t = 0; t < 4; ++t) cur[t] = colb[t];
#pragma unroll
      for (int g = 0; g < NM / 4; ++g) {
        if (g + 1 < NM / 4) {
#pragma unroll
          for (int t = 0; t < 4; ++t) nxt[t] = colb[(g + 1) * 4 + t];
        }
#pragma unroll
        for (int t = 0; t < 4; ++t) {
          const int j = g * 4 + t;
          float re = br.get(j), im = bi.get(j);
          re = fmaf(-fr, cur[t].x, re);  // -= Re(f conj(c))
          re = fmaf(-fi, cur[t].y, re);
          im = fmaf(-fi, cur[t].x, im);  // -= Im(f conj(c))
          im = fmaf(fr, cur[t].y, im);
          br.set(j, re);
          bi.set(j, im);
        }
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int t = 0; t < 4; ++t) cur[t] = nxt[t];
      }
    }
    __builtin_amdgcn_wave_barrier();
  }
  // ---- inverse iteration: y = X x = 2 x - (B x) ---------------------------
  float xr = 0.f, xi = 0.f;
  if (lane == 0) {
    xr = 1.f;
  } else if (lane < n) {
    const float2 g = coh_at(0, lane);
    const float a = sqrtf(g.x * g.x + g.y * g.y);
    xr = a > 0.f ? g.x / a : 1.f;
    xi = a > 0.f ? -g.y / a : 0.f;
  }
  {
    const float nrm = sqrtf(wsum_f(xr * xr + xi * xi));
    xr /= nrm;
    xi /= nrm;
  }
  float mu = 0.f;
  bool conv = false;
  float d1 = 0.f, d2 = 0.f;  // REF: the two squared changes before the last
  int step = -1;             // REF: refinement step, -1 while iterating
  double x_r = 0.0, x_i = 0.0, amp = 0.0, lam = 0.0;
  bool refined = false;
  constexpr int kSweeps = EMI_INV_MAX + (REF ? EMI_REF_STEPS + 1 : 0);
  for (int it = 0; it < kSweeps; ++it) {
    float vxr = xr, vxi = xi;  // the vector multiplied: the iterate, or (refining) the FP64 residual
    if constexpr (REF) {
      if (step >= 0) {
        double2* xd = reinterpret_cast<double2*>(lds);  // 64 x 16 bytes over colb and xb
        __builtin_amdgcn_wave_barrier();
        xd[lane] = make_double2(x_r, x_i);
        __builtin_amdgcn_wave_barrier();
        double yr64, yi64;
        emi_m64_product(coh_at, n, lane, xs, xd, x_r, x_i, yr64, yi64);
        lam = wsum_d(lane < n ? x_r * yr64 + x_i * yi64 : 0.0);  // x^H M x (real: M Hermitian)
        vxr = lane < n ? (float)(yr64 - lam * x_r) : 0.f;
        vxi = lane < n ? (float)(yi64 - lam * x_i) : 0.f;
        __builtin_amdgcn_wave_barrier();
      }
    }
    xb[lane] = make_float2(vxr, vxi);
    __builtin_amdgcn_wave_barrier();
    float ar = 0.f, ai = 0.f;
    int nu_ = n;
    __asm__ volatile("" : "+s"(nu_));
    {
      float2 cur[4], nxt[4];
#pragma unroll
      for (int t = 0; t < 4; ++t) cur[t] = xb[t];
#pragma unroll
      for (int g = 0; g < NM / 4; ++g) {
        if (g + 1 < NM / 4) {
#pragma unroll
          for (int t = 0; t < 4; ++t) nxt[t] = xb[(g + 1) * 4 + t];
        }
        if (g * 4 < nu_)
#pragma unroll
        for (int t = 0; t < 4; ++t) {
          const int j = g * 4 + t;
          ar = fmaf(br.get(j), cur[t].x, ar);
          ar = fmaf(-bi.get(j), cur[t].y, ar);
          ai = fmaf(br.get(j), cur[t].y, ai);
          ai = fmaf(bi.get(j), cur[t].x, ai);
        }
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int t = 0; t < 4; ++t) cur[t] = nxt[t];
      }
    }
    if constexpr (REF) {
      if (step >= 0) {
        // d = X' r, projected off x; the estimated linked-phase error
        double d_r = lane < n ? (double)(2.f * vxr - ar) : 0.0, d_i = lane < n ? (double)(2.f * vxi - ai) : 0.0;
        const double cr = wsum_d(x_r * d_r + x_i * d_i), ci = wsum_d(x_r * d_i - x_i * d_r);  // x^H d
        d_r -= cr * x_r - ci * x_i;
        d_i -= cr * x_i + ci * x_r;
        const double ax = sqrt(x_r * x_r + x_i * x_i);
        const double est = amp * wmax_d(lane < n ? sqrt(d_r * d_r + d_i * d_i) / ax : 0.0);
        if (step == 0 && est <= tau) break;  // within tolerance: the plain outputs
        const bool done = est <= tau * (1.0 / 16.0);
        if (!done && step == EMI_REF_STEPS) {
          *flag = true;
          break;
        }
        x_r -= d_r;
        x_i -= d_i;
        const double s = 1.0 / sqrt(wsum_d(x_r * x_r + x_i * x_i));
        x_r *= s;
        x_i *= s;
        if (done) {
          refined = true;
          break;
        }
        ++step;
        continue;
      }
    }
    const float yr = lane < n ? 2.f * xr - ar : 0.f, yi = lane < n ? 2.f * xi - ai : 0.f;
    mu = wsum_f(xr * yr + xi * yi);  // x^H X x (x unit, X Hermitian: real)
    const float nrm = sqrtf(wsum_f(yr * yr + yi * yi));
    const float ny = 1.f / nrm;
    const float zr = yr * ny, zi = yi * ny;
    const float dr = zr - xr, di = zi - xi;
    const float dmax = wmax_f(dr * dr + di * di);
    xr = zr;
    xi = zi;
    __builtin_amdgcn_wave_barrier();
    if (it + 1 >= EMI_INV_MIN && dmax < EMI_INV_TOL * EMI_INV_TOL) {
      conv = true;
      if constexpr (REF) {
        // rho^2 from the larger of the last two ratios of the squared changes,
        // each floored at FP32 resolution (a change at rounding level gives a
        // larger rho)
        constexpr float fl = 1.4210855e-14f;  // (2^-23)^2
        const float r2 = fmaxf(fmaxf(dmax, fl) / fmaxf(d2, fl), fmaxf(d2, fl) / fmaxf(d1, fl));
        const double rho = sqrt((double)r2);
        if (!(rho < 1.0)) {
          *flag = true;
          break;
        }
        amp = 2.0 / (1.0 - rho);
        x_r = lane < n ? (double)xr : 0.0;
        x_i = lane < n ? (double)xi : 0.0;
        const double s = 1.0 / sqrt(wsum_d(x_r * x_r + x_i * x_i));
        x_r *= s;
        x_i *= s;
        step = 0;
        continue;
      }
      break;
    }
    if (it + 1 >= EMI_INV_MAX) break;
    d1 = d2;
    d2 = dmax;
  }
  if (!conv || !(mu > 0.f) || !isfinite(mu)) return false;
  if constexpr (REF) {
    if (refined) {
      vr = lane < n ? (float)x_r : 0.f;
      vi = lane < n ? (float)x_i : 0.f;
      rq = lam;
      return true;
    }
  }
  vr = lane < n ? xr : 0.f;
  vi = lane < n ? xi : 0.f;
  rq = (double)EMI_INV_SIGMA + 1.0 / (double)mu;
  return true;
}

// One wave, 32 < n <= 64 plain path, first pass: the shift-invert formulation
// when |Gamma| is positive definite.  Returns false when the point needs the
// Householder formulation instead (|Gamma| indefinite, sigma not below the
// smallest eigenvalue, or no convergence): the caller lists it for a second
// launch running emi_fast_point (one kernel holding both formulations spills
// the register-resident matrix to scratch).
template <int NM, class Src>
__device__ __forceinline__ bool emi_inv_point(const Src& coh_at, int n, int lane, unsigned char* lds, float& vr,
                                              float& vi, double& rq) {
  RegArr<float, NM> mr_, mi_;
  if (!emi_fast_form_m<NM>(coh_at, n, lane, lds, mr_, mi_)) return false;
  return emi_shift_invert<NM>(coh_at, n, lane, lds, mr_, mi_, vr, vi, rq);
}

// The precision form's shift-invert point (REF above; xs: this wave's slot of
// n(n+1)/2 doubles for inv(|Gamma|)); flag is set when the point needs the
// FP64 re-solve.
template <int NM, class Src>
__device__ __forceinline__ bool emi_inv_point_checked(const Src& coh_at, int n, int lane, unsigned char* lds,
                                                      float& vr, float& vi, double& rq, double* xs, double tau,
                                                      bool& flag) {
  RegArr<float, NM> mr_, mi_;
  if (!emi_fast_form_m<NM>(coh_at, n, lane, lds, mr_, mi_, xs)) return false;
  return emi_shift_invert<NM, true>(coh_at, n, lane, lds, mr_, mi_, vr, vi, rq, xs, tau, &flag);
}

// ph_i = v_i conj(v_ref) / |v_i conj(v_ref)|
__device__ __forceinline__ float2 emi_link_phase(float vr, float vi, int ref) {
  const float rr = rl_f(vr, ref), ri = rl_f(vi, ref);
  const float pr = vr * rr + vi * ri;
  const float pim = vi * rr - vr * ri;
  const float am = sqrtf(pr * pr + pim * pim);
  return make_float2(pr / am, pim / am);
}

struct CohPacked {  // compressed coherence row (triu pair order) in LDS
  const float2* cs;
  int n;
  __device__ __forceinline__ float2 operator()(int a, int b) const { return cs[pair_index(a, b, n)]; }
};

}  // namespace mr
