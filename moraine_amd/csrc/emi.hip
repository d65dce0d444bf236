// EMI phase linking for gfx950.  Replaces moraine.emi (moraine/pl.py:68-77;
// _emi_cp :18-42 = cupy inv + eigh; _emi_numba :45-65 = LAPACK inv + eigh).
//
// Per point:  Gamma = uncompress(coh);  A = |Gamma|;  M = inv(A) o Gamma;
// (lambda_min, v) = smallest eigenpair of M;  ph = v conj(v_ref)/|.|;
// quality = lambda_min.
//
// Algorithm (same in both paths):
//   1. inv(A) in FP64 by the symmetric sweep operator with largest-|diagonal|
//      pivoting.  A is symmetric but indefinite whenever the SHP count is below
//      N (|Gamma| of a rank-deficient coherence matrix), so Cholesky cannot be
//      used; the sweep keeps symmetry, needs no row exchanges and its result
//      does not depend on the pivot order.
//   2. M = inv(A) o Gamma.
//   3. Householder reduction of the Hermitian M to real symmetric tridiagonal
//      form T (unitary diagonal scaling absorbs the sub-diagonal phases).
//   4. Smallest eigenvalue of T by multisection on Sturm counts (FP64, 64
//      shifts per round, one per lane), eigenvector of T (register kernels: a
//      twisted factorisation, emi_tri.hpp; general path: inverse iteration
//      with a pivoted tridiagonal LU), back-transformation through the stored
//      reflectors.
// Fast path (32 < n <= 64): one wave per point, lane i owns row i of the
// matrix in VGPRs (FP64 for step 1, FP32 complex for steps 2-4); a uniform
// runtime column index is read with M0-relative moves (RegArr in
// emi_common.hpp), never through scratch memory.  n <= 32 and 64 < n <= 200: the same algorithm with
// two lanes per row (emi_wide.hip; a single wave for n <= 32, where one lane
// per row would leave half the wave idle).  General path (n > 200): one
// 256-thread block per point, everything FP64 in a per-block global workspace.
#include "emi_fast.hpp"
#include "tcoh.hpp"

namespace mr {

// Sturm count: number of eigenvalues of the real symmetric tridiagonal (d, e2=|e|^2) below x.
__device__ __forceinline__ int sturm_count(const double* d, const double* e2, int n, double x, double pivmin) {
  double q = d[0] - x;
  int c = q < 0.0;
  for (int i = 1; i < n; ++i) {
    if (fabs(q) < pivmin) q = -pivmin;
    q = (d[i] - x) - e2[i - 1] / q;
    c += q < 0.0;
  }
  return c;
}

// Smallest eigenpair of the real symmetric tridiagonal (d, ae) -> lambda, y (written to y[]).
// Executed by one lane; all arrays are in this wave's LDS.
__device__ void tridiag_min_eigpair(const double* d, const double* ae, const double* e2, double* y, double* wk,
                                    int n, double lambda, double scale) {
  // inverse iteration with a partially pivoted LU of (T - lambda I), two sweeps from y = 1.
  double* u0 = wk;          // U diagonal
  double* u1 = wk + n;      // U first super-diagonal
  double* u2 = wk + 2 * n;  // U second super-diagonal (fill-in from pivoting)
  double* lm = wk + 3 * n;  // multipliers
  int* sw = reinterpret_cast<int*>(wk + 4 * n);  // row swapped at step i
  const double tiny = scale * 2.2e-16 + 1e-300;
  // factor
  double a = d[0] - lambda, b = n > 1 ? ae[0] : 0.0, cc = 0.0;  // current row i: [a b cc]
  for (int i = 0; i < n - 1; ++i) {
    const double sub = ae[i];                 // T[i+1][i]
    const double nd = d[i + 1] - lambda;      // T[i+1][i+1]
    const double nsup = i + 1 < n - 1 ? ae[i + 1] : 0.0;
    if (fabs(a) >= fabs(sub)) {
      const double l = (a != 0.0) ? sub / a : 0.0;
      u0[i] = a;
      u1[i] = b;
      u2[i] = cc;
      lm[i] = l;
      sw[i] = 0;
      a = nd - l * b;
      b = nsup - l * cc;
      cc = 0.0;
    } else {
      const double l = a / sub;
      u0[i] = sub;
      u1[i] = nd;
      u2[i] = nsup;
      lm[i] = l;
      sw[i] = 1;
      const double na = b - l * nd;
      const double nb = cc - l * nsup;
      a = na;
      b = nb;
      cc = 0.0;
    }
  }
  u0[n - 1] = a;
  u1[n - 1] = 0.0;
  u2[n - 1] = 0.0;
  for (int i = 0; i < n; ++i) {
    if (fabs(u0[i]) < tiny) u0[i] = copysign(tiny, u0[i] == 0.0 ? 1.0 : u0[i]);
    y[i] = 1.0;
  }
  for (int it = 0; it < 3; ++it) {
    // forward: apply row swaps and multipliers
    for (int i = 0; i < n - 1; ++i) {
      if (sw[i]) {
        const double t = y[i];
        y[i] = y[i + 1];
        y[i + 1] = t;
      }
      y[i + 1] -= lm[i] * y[i];
    }
    // backward
    for (int i = n - 1; i >= 0; --i) {
      double s = y[i];
      if (i + 1 < n) s -= u1[i] * y[i + 1];
      if (i + 2 < n) s -= u2[i] * y[i + 2];
      y[i] = s / u0[i];
    }
    double nrm = 0.0;
    for (int i = 0; i < n; ++i) nrm += y[i] * y[i];
    nrm = 1.0 / sqrt(nrm);
    for (int i = 0; i < n; ++i) y[i] *= nrm;
  }
}

// ----------------------------------------------------------------------------
// fast path: one wave per point, NM >= n, n <= 64
// ----------------------------------------------------------------------------
constexpr int EMI_FAST_THREADS = 64;

// TC: also the DS temporal coherence of each point from the staged row
// (moraine/pl.py:126-145), as mr_ds_temp_coh computes it from (coh, ph, pairs).
// CHK: points whose conditioning test fails (tau) are appended to
// flag_list[*flag_count] for the FP64 re-solve.
// MODE 0: the Householder formulation for every point (with CHK: the precision
// form's flags).  MODE 1: the shift-invert formulation; a point it declines is
// appended to flag_list[*flag_count] and left unwritten.  MODE 2: the
// Householder formulation for the points of that list.
template <int NM, bool TC, bool CHK = false, int MODE = 0>
__global__ __launch_bounds__(EMI_FAST_THREADS) void emi_fast_kernel(const float2* __restrict__ coh, long long npts,
                                                                    int n, int ref, float2* __restrict__ ph,
                                                                    float* __restrict__ quality,
                                                                    const int32_t* __restrict__ pairs,
                                                                    float* __restrict__ t_coh, double tau = 0.0,
                                                                    unsigned* __restrict__ flag_count = nullptr,
                                                                    int* __restrict__ flag_list = nullptr) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int P = n * (n - 1) / 2;
  const int lane0 = threadIdx.x;
  float2* cs = reinterpret_cast<float2*>(smem);                          // staged coh, P entries
  unsigned char* work = smem + ((P * 8 + 15) & ~15);                      // EMI_FAST_LDS_BYTES

  const long long nit = MODE == 2 ? (long long)*flag_count : npts;
  for (long long it = blockIdx.x; it < nit; it += gridDim.x) {
    const long long pt = MODE == 2 ? (long long)flag_list[it] : it;
    // Opaque per-point copy of the lane id: keeps the compiler from hoisting
    // NM per-lane gather addresses out of the point loop (they would stay live
    // in VGPRs for the whole kernel).
    int lane = lane0;
    __asm__ volatile("" : "+v"(lane));
    const float2* src = coh + pt * (long long)P;
    // stage the row: sixteen loads per lane are in flight before their LDS
    // stores (a plain copy loop waits for every load before the next one)
    for (int t0 = 0; t0 < P; t0 += 64 * 16) {
      float2 v[16];
#pragma unroll
      for (int u = 0; u < 16; ++u) {
        const int t = t0 + 64 * u + lane;
        v[u] = t < P ? src[t] : make_float2(0.f, 0.f);
      }
#pragma unroll
      for (int u = 0; u < 16; ++u) {
        const int t = t0 + 64 * u + lane;
        if (t < P) cs[t] = v[u];
      }
    }
    __builtin_amdgcn_wave_barrier();
    float vr, vi;
    double rq;
    bool flag = false;
    if constexpr (MODE == 1) {
      if (!emi_inv_point<NM>(CohPacked{cs, n}, n, lane, work, vr, vi, rq)) {
        if (lane == 0) flag_list[atomicAdd(flag_count, 1u)] = (int)pt;
        __builtin_amdgcn_wave_barrier();
        continue;
      }
    } else {
      emi_fast_point<NM, CHK>(CohPacked{cs, n}, n, lane, work, vr, vi, rq, tau, &flag);
    }
    // ---- output: ph = v conj(v_ref) / |.|, quality --------------------------
    const float2 pv = emi_link_phase(vr, vi, ref);
    if (lane < n) ph[pt * n + lane] = pv;
    if (lane == 0) quality[pt] = (float)rq;
    if (CHK && flag && lane == 0) flag_list[atomicAdd(flag_count, 1u)] = (int)pt;
    if (TC) {
      float2* phb = reinterpret_cast<float2*>(work);  // the EMI scratch is free again
      __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
      __builtin_amdgcn_wave_barrier();
      phb[lane] = pv;
      __builtin_amdgcn_wave_barrier();
      const float tc = tcoh_point<true>([&](int p, int, int) { return cs[p]; }, [&](int i) { return phb[i]; }, pairs,
                                        P, lane);
      if (lane == 0) t_coh[pt] = tc;
    }
    __builtin_amdgcn_wave_barrier();
  }
}

// ----------------------------------------------------------------------------
// general path: one 256-thread block per point, FP64 in a global workspace
// ----------------------------------------------------------------------------
constexpr int EMI_GEN_THREADS = 256;
constexpr int EMI_GEN_BLOCKS = 1024;
constexpr int EMI_REF_BLOCKS = 1024;  // FP64 re-solve blocks for n > 64, four per CU (each with a global slice)

__host__ __device__ inline size_t emi_gen_slice_bytes(int n) {
  // A (n*n f64) + M (n*n c128) + vectors (col, u, q, p: n c128 each) + 12 n-double arrays
  // (dg, td, tae, te2, ty, dphr, dphi, 5n LU work) + n ints (swept)
  return ((size_t)n * n * 8 + (size_t)n * n * 16 + (size_t)n * 16 * 4 + (size_t)n * 8 * 13 + 255) & ~(size_t)255;
}

__device__ __forceinline__ double block_sum(double v, double* red) {
  v = wave_sum(v);
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  __syncthreads();
  if (l == 0) red[w] = v;
  __syncthreads();
  double s = 0.0;
  for (int i = 0; i < (int)(blockDim.x >> 6); ++i) s += red[i];
  return s;
}

// LIST: the FP64 re-solve of flagged points: the points are list[0 .. *count);
// for n <= 64 the block's slice lives in LDS, and with t_coh set each
// re-solved point's DS temporal coherence is recomputed from its new phases.
// Z = double2: the complex128 form (coherence and phases complex128; every
// stage FP64 as below; n <= 64 with the slice in LDS).
template <bool LIST, typename Z = float2>
__global__ __launch_bounds__(EMI_GEN_THREADS) void emi_general_kernel(const Z* __restrict__ coh,
                                                                      long long npts, int n, int ref,
                                                                      Z* __restrict__ ph,
                                                                      float* __restrict__ quality,
                                                                      unsigned char* __restrict__ ws,
                                                                      const int* __restrict__ list = nullptr,
                                                                      const unsigned* __restrict__ count = nullptr,
                                                                      const int32_t* __restrict__ pairs = nullptr,
                                                                      float* __restrict__ t_coh = nullptr) {
  __shared__ double red[EMI_GEN_THREADS / 64];
  __shared__ double sh_key[EMI_GEN_THREADS];
  __shared__ int sh_idx[EMI_GEN_THREADS];
  // T for the one-wave eigen-solve of the LDS-slice path (LIST or complex128, n <= 64)
  constexpr bool kSmallPath = LIST || sizeof(Z) == 16;
  __shared__ double2 tri_de[kSmallPath ? 64 : 1];
  __shared__ double tri_e[kSmallPath ? 64 : 1];
  extern __shared__ __align__(16) unsigned char gen_lds[];
  const int P = n * (n - 1) / 2;
  const int tid = threadIdx.x;
  // LIST with n <= 64: the slice in LDS, pivot and eigen-solve by one wave
  const bool small = kSmallPath && n <= 64;
  unsigned char* base = small ? gen_lds : ws + (size_t)blockIdx.x * emi_gen_slice_bytes(n);
  double* A = reinterpret_cast<double*>(base);
  c128* M = reinterpret_cast<c128*>(A + (size_t)n * n);
  c128* col = M + (size_t)n * n;  // n
  c128* u = col + n;
  c128* q = u + n;
  c128* pv = q + n;
  double* dg = reinterpret_cast<double*>(pv + n);
  double* td = dg + n;
  double* tae = td + n;
  double* te2 = tae + n;
  double* ty = te2 + n;
  double* dphr = ty + n;
  double* dphi = dphr + n;
  double* wk = dphi + n;  // 5n
  int* swept = reinterpret_cast<int*>(wk + 5 * n);

  const long long nit = LIST ? (long long)*count : npts;
  for (long long it = blockIdx.x; it < nit; it += gridDim.x) {
    const long long pt = LIST ? (long long)list[it] : it;
    const Z* src = coh + pt * (long long)P;
    for (int t = tid; t < n * n; t += EMI_GEN_THREADS) {
      const int i = t / n, j = t - (t / n) * n;
      c128 g;
      if (i == j) {
        g = {1.0, 0.0};
      } else if (j > i) {
        const Z c = src[pair_index(i, j, n)];
        g = {c.x, c.y};
      } else {
        const Z c = src[pair_index(j, i, n)];
        g = {c.x, -c.y};
      }
      M[t] = g;
      A[t] = sqrt(g.re * g.re + g.im * g.im);
    }
    for (int i = tid; i < n; i += EMI_GEN_THREADS) {
      dg[i] = 1.0;
      swept[i] = 0;
    }
    __syncthreads();
    // sweep inverse
    for (int s = 0; s < n; ++s) {
      if (small) {
        // n <= 64: the pivot (largest |d| not yet swept, ties to the smallest
        // row) from one wave
        if (tid < 64) {
          const int k1 = wargmax_d(tid < n && !swept[tid] ? fabs(dg[tid]) : -1.0, tid);
          if (tid == 0) sh_idx[0] = k1;
        }
        __syncthreads();
      } else {
        double key = -1.0;
        int idx = 0x7fffffff;
        for (int i = tid; i < n; i += EMI_GEN_THREADS)
          if (!swept[i] && (fabs(dg[i]) > key || (fabs(dg[i]) == key && i < idx))) {
            key = fabs(dg[i]);
            idx = i;
          }
        sh_key[tid] = key;
        sh_idx[tid] = idx;
        __syncthreads();
        for (int st = EMI_GEN_THREADS / 2; st > 0; st >>= 1) {
          if (tid < st) {
            const double ok = sh_key[tid + st];
            const int oi = sh_idx[tid + st];
            if (ok > sh_key[tid] || (ok == sh_key[tid] && oi < sh_idx[tid])) {
              sh_key[tid] = ok;
              sh_idx[tid] = oi;
            }
          }
          __syncthreads();
        }
      }
      const int k = sh_idx[0];
      for (int i = tid; i < n; i += EMI_GEN_THREADS) col[i].re = A[(size_t)i * n + k];
      __syncthreads();
      const double rd = 1.0 / col[k].re;
      for (int t = tid; t < n * n; t += EMI_GEN_THREADS) {
        const int i = t / n, j = t - (t / n) * n;
        const double f = (i == k) ? (1.0 - rd) : col[i].re * rd;
        A[t] = fma(-f, col[j].re, A[t]);
      }
      __syncthreads();
      for (int i = tid; i < n; i += EMI_GEN_THREADS) {
        A[(size_t)i * n + k] = (i == k) ? -rd : col[i].re * rd;
        const double f = (i == k) ? (1.0 - rd) : col[i].re * rd;
        dg[i] = fma(-f, col[i].re, dg[i]);
        if (i == k) swept[i] = 1;
      }
      __syncthreads();
    }
    // M = inv(A) o Gamma  (A holds -inv)
    for (int t = tid; t < n * n; t += EMI_GEN_THREADS) {
      const double ai = -A[t];
      M[t] = {ai * M[t].re, ai * M[t].im};
    }
    __syncthreads();
    // Householder tridiagonalisation (FP64); reflector k kept in column k below the diagonal
    double dcr = 1.0, dci = 0.0;
    if (tid == 0) {
      dphr[0] = 1.0;
      dphi[0] = 0.0;
    }
    for (int k = 0; k + 2 < n; ++k) {
      double part = 0.0;
      for (int i = k + 1 + tid; i < n; i += EMI_GEN_THREADS) {
        const c128 x = M[(size_t)i * n + k];
        part += x.re * x.re + x.im * x.im;
      }
      const double nrm2 = block_sum(part, red);
      const c128 x0 = M[(size_t)(k + 1) * n + k];
      const double nx = sqrt(nrm2);
      const double ax0 = sqrt(x0.re * x0.re + x0.im * x0.im);
      const double phr = ax0 > 0 ? x0.re / ax0 : 1.0, phi = ax0 > 0 ? x0.im / ax0 : 0.0;
      const double alr = -phr * nx, ali = -phi * nx;
      const double den = sqrt(2.0 * nx * (nx + ax0));
      const double inv = den > 0 ? 1.0 / den : 0.0;
      const double dkk = M[(size_t)k * n + k].re;
      __syncthreads();
      for (int i = tid; i < n; i += EMI_GEN_THREADS) {
        c128 v = {0.0, 0.0};
        if (i > k) {
          const c128 x = M[(size_t)i * n + k];
          v = (i == k + 1) ? c128{(x.re - alr) * inv, (x.im - ali) * inv} : c128{x.re * inv, x.im * inv};
        }
        u[i] = v;
      }
      __syncthreads();
      for (int i = tid; i < n; i += EMI_GEN_THREADS) {
        c128 p = {0.0, 0.0};
        if (i > k)
          for (int j = k + 1; j < n; ++j) {
            const c128 m = M[(size_t)i * n + j];
            p.re += m.re * u[j].re - m.im * u[j].im;
            p.im += m.re * u[j].im + m.im * u[j].re;
          }
        pv[i] = p;
      }
      __syncthreads();
      double cp = 0.0;
      for (int i = tid; i < n; i += EMI_GEN_THREADS) cp += u[i].re * pv[i].re + u[i].im * pv[i].im;
      const double cval = block_sum(cp, red);
      for (int i = tid; i < n; i += EMI_GEN_THREADS)
        q[i] = (i > k) ? c128{2.0 * (pv[i].re - cval * u[i].re), 2.0 * (pv[i].im - cval * u[i].im)} : c128{0.0, 0.0};
      __syncthreads();
      const int m = n - k - 1;
      for (int t = tid; t < m * m; t += EMI_GEN_THREADS) {
        const int i = k + 1 + t / m, j = k + 1 + (t - (t / m) * m);
        c128 v = M[(size_t)i * n + j];
        // -= u_i conj(q_j) + q_i conj(u_j)
        v.re -= u[i].re * q[j].re + u[i].im * q[j].im + q[i].re * u[j].re + q[i].im * u[j].im;
        v.im -= u[i].im * q[j].re - u[i].re * q[j].im + q[i].im * u[j].re - q[i].re * u[j].im;
        M[(size_t)i * n + j] = v;
      }
      for (int i = tid; i < n; i += EMI_GEN_THREADS)
        if (i > k) M[(size_t)i * n + k] = u[i];
      if (tid == 0) {
        td[k] = dkk;
        tae[k] = nx;
        te2[k] = nx * nx;
        const double sr = nx > 0 ? alr / nx : 1.0, si = nx > 0 ? ali / nx : 0.0;
        const double r = dcr * sr - dci * si, im = dcr * si + dci * sr;
        dcr = r;
        dci = im;
        dphr[k + 1] = dcr;
        dphi[k + 1] = dci;
      }
      __syncthreads();
    }
    if (tid == 0) {
      const c128 e = M[(size_t)(n - 1) * n + (n - 2)];
      const double ae = sqrt(e.re * e.re + e.im * e.im);
      td[n - 2] = M[(size_t)(n - 2) * n + (n - 2)].re;
      td[n - 1] = M[(size_t)(n - 1) * n + (n - 1)].re;
      tae[n - 2] = ae;
      te2[n - 2] = ae * ae;
      tae[n - 1] = 0.0;
      te2[n - 1] = 0.0;
      const double sr = ae > 0 ? e.re / ae : 1.0, si = ae > 0 ? e.im / ae : 0.0;
      dphr[n - 1] = dcr * sr - dci * si;
      dphi[n - 1] = dcr * si + dci * sr;
      if (!small) {
        // bisection (serial, FP64) + inverse iteration
        double lo = 1e308, hi = -1e308, sc = 0.0;
        for (int i = 0; i < n; ++i) {
          const double off = (i > 0 ? tae[i - 1] : 0.0) + tae[i];
          lo = fmin(lo, td[i] - off);
          hi = fmax(hi, td[i] + off);
          sc = fmax(sc, fabs(td[i]) + off);
        }
        const double scale = sc > 0 ? sc : 1.0;
        lo -= 1e-14 * scale;
        hi += 1e-14 * scale;
        const double pivmin = 1e-300 + scale * scale * 1e-290;
        for (int it = 0; it < 200 && hi - lo > 2e-15 * scale; ++it) {
          const double mid = 0.5 * (lo + hi);
          if (mid <= lo || mid >= hi) break;
          if (sturm_count(td, te2, n, mid, pivmin) >= 1)
            hi = mid;
          else
            lo = mid;
        }
        const double lam = 0.5 * (lo + hi);
        tridiag_min_eigpair(td, tae, te2, ty, wk, n, lam, scale);
        double rq = 0.0;
        for (int i = 0; i < n; ++i) {
          double t = td[i] * ty[i];
          if (i > 0) t += tae[i - 1] * ty[i - 1];
          if (i + 1 < n) t += tae[i] * ty[i + 1];
          rq += ty[i] * t;
        }
        quality[pt] = (float)rq;
      }
    }
    __syncthreads();
    if (small) {
      // n <= 64: one wave finds the smallest eigenpair of T by multisection
      // (run to FP64 resolution) and a twisted factorisation
      if (tid < 64) {
        double rq;
        const double z = emi_tri_smallest(n, tid, tid < n ? td[tid] : 0.0, tid < n ? tae[tid] : 0.0, tri_de, tri_e,
                                          rq, 12);
        if (tid < n) ty[tid] = z;
        if (tid == 0) quality[pt] = (float)rq;
      }
      __syncthreads();
    }
    // back-transformation v = H_0 ... H_{n-3} D y
    for (int i = tid; i < n; i += EMI_GEN_THREADS) pv[i] = {dphr[i] * ty[i], dphi[i] * ty[i]};
    __syncthreads();
    for (int k = n - 3; k >= 0; --k) {
      double sr = 0.0, si = 0.0;
      for (int i = k + 1 + tid; i < n; i += EMI_GEN_THREADS) {
        const c128 uu = M[(size_t)i * n + k];
        sr += uu.re * pv[i].re + uu.im * pv[i].im;
        si += uu.re * pv[i].im - uu.im * pv[i].re;
      }
      sr = block_sum(sr, red);
      si = block_sum(si, red);
      for (int i = k + 1 + tid; i < n; i += EMI_GEN_THREADS) {
        const c128 uu = M[(size_t)i * n + k];
        pv[i].re -= 2.0 * (uu.re * sr - uu.im * si);
        pv[i].im -= 2.0 * (uu.re * si + uu.im * sr);
      }
      __syncthreads();
    }
    const c128 vr = pv[ref];
    for (int i = tid; i < n; i += EMI_GEN_THREADS) {
      const double pr = pv[i].re * vr.re + pv[i].im * vr.im;
      const double pim = pv[i].im * vr.re - pv[i].re * vr.im;
      const double am = sqrt(pr * pr + pim * pim);
      const Z v = mkz<Z>((typename ZT<Z>::S)(pr / am), (typename ZT<Z>::S)(pim / am));
      ph[pt * n + i] = v;
      if constexpr (LIST) reinterpret_cast<float2*>(wk)[i] = v;  // LU work space is free again
    }
    __syncthreads();
    if constexpr (LIST) {
      if (t_coh != nullptr && tid < 64) {
        const float2* phl = reinterpret_cast<const float2*>(wk);
        const float tc = tcoh_point<true>([&](int p, int, int) { return src[p]; }, [&](int i) { return phl[i]; },
                                          pairs, P, tid);
        if (tid == 0) t_coh[pt] = tc;
      }
    }
    __syncthreads();
  }
}

int emi_wide_launch(const float2* coh, long long npts, int n, int ref, float2* ph, float* quality, hipStream_t st,
                    double tau = 0.0, unsigned* flag_count = nullptr, int* flag_list = nullptr);

__global__ void emi_zero_count_kernel(unsigned* c) { *c = 0u; }

// count + list of the points the shift-invert pass leaves to the Householder pass
inline size_t emi_list_bytes(long long npts) { return 256 + (((size_t)(npts > 0 ? npts : 0) * 4 + 255) & ~(size_t)255); }

static size_t emi_fast_lds(int n, int nm) {
  const int P = n * (n - 1) / 2;
  return (((size_t)P * 8 + 15) & ~(size_t)15) + EMI_FAST_LDS_BYTES;
}

// 32 < n <= 64: the shift-invert pass over every point, then the Householder
// pass over the points it left (count and list at the start of ws).
constexpr int EMI_FALLBACK_BLOCKS = 8192;
template <bool TC>
static int emi_two_pass(const float2* c, long long n_pts, int n, int ref, float2* p, float* quality,
                        const int32_t* pairs, float* t_coh, unsigned char* ws, hipStream_t st) {
  unsigned* count = reinterpret_cast<unsigned*>(ws);
  int* list = reinterpret_cast<int*>(ws + 256);
  const size_t lds = emi_fast_lds(n, 64);
  hipLaunchKernelGGL(emi_zero_count_kernel, dim3(1), dim3(1), 0, st, count);
  if (int rc = launch_status("emi count init")) return rc;
  const int nb = (int)std::min<long long>(n_pts, 1 << 20);
  hipLaunchKernelGGL((emi_fast_kernel<64, TC, false, 1>), dim3(nb), dim3(EMI_FAST_THREADS), lds, st, c, n_pts, n,
                     ref, p, quality, pairs, t_coh, 0.0, count, list);
  if (int rc = launch_status("emi_fast shift-invert")) return rc;
  const int nb2 = (int)std::min<long long>(n_pts, EMI_FALLBACK_BLOCKS);
  hipLaunchKernelGGL((emi_fast_kernel<64, TC, false, 2>), dim3(nb2), dim3(EMI_FAST_THREADS), lds, st, c, n_pts, n,
                     ref, p, quality, pairs, t_coh, 0.0, count, list);
  return launch_status("emi_fast householder");
}

}  // namespace mr

extern "C" {

size_t mr_emi_workspace_bytes(long long n_pts, int n) {
  if (n > 32 && n <= 64) return mr::emi_list_bytes(n_pts);  // the Householder-fallback list
  if (n <= mr::EMI_WIDE_MAX_N) return 0;
  const long long nb = n_pts < mr::EMI_GEN_BLOCKS ? n_pts : mr::EMI_GEN_BLOCKS;
  return (size_t)(nb > 0 ? nb : 1) * mr::emi_gen_slice_bytes(n);
}

int mr_emi(const float* coh, long long n_pts, int n, int ref, float* ph, float* quality, void* workspace,
           size_t workspace_bytes, mr_stream_t stream) {
  MR_CHECK_ARG(n >= 2, "emi: need at least 2 images (n=%d)", n);
  MR_CHECK_ARG(ref >= 0 && ref < n, "emi: ref=%d outside [0,%d)", ref, n);
  MR_CHECK_ARG(n_pts >= 0, "emi: negative point count");
  if (n_pts == 0) return MR_OK;
  MR_CHECK_ARG(coh && ph && quality, "emi: null pointer");
  hipStream_t st = mr::as_stream(stream);
  const float2* c = reinterpret_cast<const float2*>(coh);
  float2* p = reinterpret_cast<float2*>(ph);
  if (n > 32 && n <= 64) {
    const size_t need = mr_emi_workspace_bytes(n_pts, n);
    MR_CHECK_ARG(workspace != nullptr && workspace_bytes >= need, "emi: workspace %zu < %zu bytes", workspace_bytes,
                 need);
    return mr::emi_two_pass<false>(c, n_pts, n, ref, p, quality, nullptr, nullptr,
                                   reinterpret_cast<unsigned char*>(workspace), st);
  }
  if (n <= mr::EMI_WIDE_MAX_N) return mr::emi_wide_launch(c, n_pts, n, ref, p, quality, st);
  const size_t need = mr_emi_workspace_bytes(n_pts, n);
  MR_CHECK_ARG(workspace != nullptr && workspace_bytes >= need, "emi: workspace %zu < %zu bytes", workspace_bytes,
               need);
  const int nb = (int)std::min<long long>(n_pts, mr::EMI_GEN_BLOCKS);
  hipLaunchKernelGGL(mr::emi_general_kernel<false>, dim3(nb), dim3(mr::EMI_GEN_THREADS), 0, st, c, n_pts, n, ref, p,
                     quality, reinterpret_cast<unsigned char*>(workspace));
  return mr::launch_status("emi_general");
}

size_t mr_emi_c128_workspace_bytes(long long n_pts, int n) {
  if (n <= 64) return 0;  // the slice of a point lives in LDS
  const long long nb = n_pts < mr::EMI_GEN_BLOCKS ? n_pts : mr::EMI_GEN_BLOCKS;
  return (size_t)(nb > 0 ? nb : 1) * mr::emi_gen_slice_bytes(n);
}

int mr_emi_c128(const double* coh, long long n_pts, int n, int ref, double* ph, float* quality, void* workspace,
                size_t workspace_bytes, mr_stream_t stream) {
  MR_CHECK_ARG(n >= 2, "emi_c128: need at least 2 images (n=%d)", n);
  MR_CHECK_ARG(ref >= 0 && ref < n, "emi_c128: ref=%d outside [0,%d)", ref, n);
  MR_CHECK_ARG(n_pts >= 0, "emi_c128: negative point count");
  if (n_pts == 0) return MR_OK;
  MR_CHECK_ARG(coh && ph && quality, "emi_c128: null pointer");
  const size_t need = mr_emi_c128_workspace_bytes(n_pts, n);
  MR_CHECK_ARG(need == 0 || (workspace != nullptr && workspace_bytes >= need), "emi_c128: workspace %zu < %zu bytes",
               workspace_bytes, need);
  const int nb = (int)std::min<long long>(n_pts, mr::EMI_GEN_BLOCKS);
  hipLaunchKernelGGL((mr::emi_general_kernel<false, double2>), dim3(nb), dim3(mr::EMI_GEN_THREADS),
                     n <= 64 ? mr::emi_gen_slice_bytes(n) : 0, mr::as_stream(stream),
                     reinterpret_cast<const double2*>(coh), n_pts, n, ref, reinterpret_cast<double2*>(ph), quality,
                     reinterpret_cast<unsigned char*>(workspace), nullptr, nullptr, nullptr, nullptr);
  return mr::launch_status("emi_c128");
}

int mr_emi_ds_temp_coh(const float* coh, long long n_pts, int n, int ref, const int32_t* pairs, int n_pairs,
                       float* ph, float* quality, float* t_coh, void* workspace, size_t workspace_bytes,
                       mr_stream_t stream) {
  MR_CHECK_ARG(n >= 2, "emi_ds_temp_coh: need at least 2 images (n=%d)", n);
  MR_CHECK_ARG(ref >= 0 && ref < n, "emi_ds_temp_coh: ref=%d outside [0,%d)", ref, n);
  MR_CHECK_ARG(n_pts >= 0, "emi_ds_temp_coh: negative point count");
  MR_CHECK_ARG(n_pairs == n * (n - 1) / 2, "emi_ds_temp_coh: coherence rows hold %d pairs, n=%d needs %d", n_pairs, n,
               n * (n - 1) / 2);
  if (n_pts == 0) return MR_OK;
  MR_CHECK_ARG(coh && pairs && ph && quality && t_coh, "emi_ds_temp_coh: null pointer");
  if (n > 32 && n <= 64) {
    const size_t need = mr_emi_workspace_bytes(n_pts, n);
    MR_CHECK_ARG(workspace != nullptr && workspace_bytes >= need, "emi_ds_temp_coh: workspace %zu < %zu bytes",
                 workspace_bytes, need);
    return mr::emi_two_pass<true>(reinterpret_cast<const float2*>(coh), n_pts, n, ref, reinterpret_cast<float2*>(ph),
                                  quality, pairs, t_coh, reinterpret_cast<unsigned char*>(workspace),
                                  mr::as_stream(stream));
  }
  const int rc = mr_emi(coh, n_pts, n, ref, ph, quality, workspace, workspace_bytes, stream);
  if (rc != MR_OK) return rc;
  return mr_ds_temp_coh(coh, ph, n_pts, n, pairs, n_pairs, t_coh, stream);
}

size_t mr_emi_refined_workspace_bytes(long long n_pts, int n) {
  const size_t list = 256 + (((size_t)(n_pts > 0 ? n_pts : 0) * 4 + 255) & ~(size_t)255);
  if (n <= 64) return list;
  // n > 64: the FP64 re-solve's per-block slices in global memory
  const long long nb = n_pts < mr::EMI_REF_BLOCKS ? n_pts : mr::EMI_REF_BLOCKS;
  return list + (size_t)(nb > 0 ? nb : 1) * mr::emi_gen_slice_bytes(n);
}

int mr_emi_refined(const float* coh, long long n_pts, int n, int ref, double tol, const int32_t* pairs, int n_pairs,
                   float* ph, float* quality, float* t_coh, void* workspace, size_t workspace_bytes,
                   mr_stream_t stream) {
  MR_CHECK_ARG(n >= 2 && n <= mr::EMI_WIDE_MAX_N,
               "emi_refined: covers 2 <= n <= %d (n=%d; above, mr_emi is FP64 throughout)", mr::EMI_WIDE_MAX_N, n);
  MR_CHECK_ARG(ref >= 0 && ref < n, "emi_refined: ref=%d outside [0,%d)", ref, n);
  MR_CHECK_ARG(n_pts >= 0 && n_pts < (1ll << 31), "emi_refined: point count %lld outside [0, 2^31)", n_pts);
  MR_CHECK_ARG(tol > 0.0, "emi_refined: tol must be positive");
  MR_CHECK_ARG(t_coh == nullptr || n_pairs == n * (n - 1) / 2,
               "emi_refined: coherence rows hold %d pairs, n=%d needs %d", n_pairs, n, n * (n - 1) / 2);
  if (n_pts == 0) return MR_OK;
  MR_CHECK_ARG(coh && ph && quality && (t_coh == nullptr || pairs), "emi_refined: null pointer");
  const size_t need = mr_emi_refined_workspace_bytes(n_pts, n);
  MR_CHECK_ARG(workspace != nullptr && workspace_bytes >= need, "emi_refined: workspace %zu < %zu bytes",
               workspace_bytes, need);
  hipStream_t st = mr::as_stream(stream);
  unsigned char* w = reinterpret_cast<unsigned char*>(workspace);
  unsigned* count = reinterpret_cast<unsigned*>(w);
  int* list = reinterpret_cast<int*>(w + 256);
  hipLaunchKernelGGL(mr::emi_zero_count_kernel, dim3(1), dim3(1), 0, st, count);
  if (int rc = mr::launch_status("emi_refined")) return rc;
  const float2* c = reinterpret_cast<const float2*>(coh);
  float2* p = reinterpret_cast<float2*>(ph);
  const bool fused = t_coh != nullptr && n > 32 && n <= 64;
  if (n > 32 && n <= 64) {
    const size_t lds = mr::emi_fast_lds(n, 64);
    const int nb = (int)std::min<long long>(n_pts, 1 << 20);
    if (fused)
      hipLaunchKernelGGL((mr::emi_fast_kernel<64, true, true>), dim3(nb), dim3(mr::EMI_FAST_THREADS), lds, st, c,
                         n_pts, n, ref, p, quality, pairs, t_coh, tol, count, list);
    else
      hipLaunchKernelGGL((mr::emi_fast_kernel<64, false, true>), dim3(nb), dim3(mr::EMI_FAST_THREADS), lds, st, c,
                         n_pts, n, ref, p, quality, nullptr, nullptr, tol, count, list);
    if (int rc = mr::launch_status("emi_refined")) return rc;
  } else {
    if (int rc = mr::emi_wide_launch(c, n_pts, n, ref, p, quality, st, tol, count, list)) return rc;
  }
  // FP64 re-solve of the flagged points, one block per point (slice in LDS for
  // n <= 64, in the workspace after the list above)
  if (n <= 64) {
    const int nb2 = (int)std::min<long long>(n_pts, mr::EMI_GEN_BLOCKS);
    hipLaunchKernelGGL(mr::emi_general_kernel<true>, dim3(nb2), dim3(mr::EMI_GEN_THREADS),
                       mr::emi_gen_slice_bytes(n), st, c, n_pts, n, ref, p, quality, nullptr, list, count, pairs,
                       fused ? t_coh : nullptr);
  } else {
    const int nb2 = (int)std::min<long long>(n_pts, mr::EMI_REF_BLOCKS);
    unsigned char* slices = w + 256 + (((size_t)n_pts * 4 + 255) & ~(size_t)255);
    hipLaunchKernelGGL(mr::emi_general_kernel<true>, dim3(nb2), dim3(mr::EMI_GEN_THREADS), 0, st, c, n_pts, n, ref,
                       p, quality, slices, list, count, nullptr, nullptr);
  }
  if (int rc = mr::launch_status("emi_refined")) return rc;
  if (t_coh && !fused) return mr_ds_temp_coh(coh, ph, n_pts, n, pairs, n_pairs, t_coh, stream);
  return MR_OK;
}

}  // extern "C"
