// FP64 EMI of one point per 256-thread block (moraine/pl.py:45-65 restated,
// every stage in FP64): the general path (n > 200), the complex128 form, and
// the FP64 re-solve of points the FP32 kernels flag.  Shared by emi.hip and
// the batched form (batch.hip), which feeds it the coherence from LDS tiles.
#pragma once
#include "emi_tri.hpp"
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

constexpr int EMI_GEN_THREADS = 256;
constexpr int EMI_GEN_BLOCKS = 1024;
constexpr int EMI_REF_BLOCKS = 1024;  // FP64 re-solve blocks for n > 64, four per CU (each with a global slice)

__host__ __device__ inline size_t emi_gen_slice_bytes(int n) {
  // M (n*n c128; A, n*n f64, in its first half until M is formed) + vectors (col, u, q, p:
  // n c128 each) + 12 n-double arrays (dg, td, tae, te2, ty, dphr, dphi, 5n LU work) + n ints
  // (swept).  n = 64: 76 288 bytes, two blocks per CU.
  return ((size_t)n * n * 16 + (size_t)n * 16 * 4 + (size_t)n * 8 * 13 + 255) & ~(size_t)255;
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

// Two block sums in one exchange (each the same sum, in the same order, as block_sum's).
__device__ __forceinline__ void block_sum2(double& a, double& b, double* red2) {
  a = wave_sum(a);
  b = wave_sum(b);
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  __syncthreads();
  if (l == 0) {
    red2[2 * w] = a;
    red2[2 * w + 1] = b;
  }
  __syncthreads();
  double sa = 0.0, sb = 0.0;
  for (int i = 0; i < (int)(blockDim.x >> 6); ++i) {
    sa += red2[2 * i];
    sb += red2[2 * i + 1];
  }
  a = sa;
  b = sb;
}

// Static LDS of a block running emi_general_solve.
struct GenShared {
  double* red;      // EMI_GEN_THREADS / 64
  double* sh_key;   // EMI_GEN_THREADS
  int* sh_idx;      // EMI_GEN_THREADS
  double2* tri_de;  // 64 (small path)
  double* tri_e;    // 64 (small path)
};

// Compressed coherence row in global memory (triu pair order).
template <typename Z>
struct GenRow {
  const Z* row;
  int n;
  __device__ __forceinline__ Z operator()(int a, int b) const { return row[pair_index(a, b, n)]; }
  __device__ __forceinline__ Z pair(int p, int, int) const { return row[p]; }
};

// The FP64 solve of one point by one EMI_GEN_THREADS block.  src(a, b), a < b,
// returns the coherence of pair (a, b); src.pair(p, a, b) the entry of pairs[p]
// = (a, b) for the temporal coherence.  small (n <= 64): the slice `base`
// (emi_gen_slice_bytes(n)) is in LDS and one wave finds the pivots and the
// eigenpair of T; otherwise `base` is a global slice and the block reduces
// the pivot search, one thread bisects.  Writes ph_row[0..n) and *q_out; TC
// (n <= 64, Z = float2): also the DS temporal coherence of the new phases
// over the n(n-1)/2 pairs into *t_out.  src is read twice (|Gamma| for the
// sweep, then Gamma to form M in A's place).  Element loops run rows over the
// block's waves and columns over lanes (no index division).
template <bool TC, typename Z, class Src>
__device__ void emi_general_solve(const Src& src, int n, int ref, bool small, unsigned char* base,
                                  const GenShared& sh, Z* ph_row, float* q_out, const int32_t* pairs, float* t_out) {
  static_assert(!TC || sizeof(Z) == 8, "the fused temporal coherence takes complex64 phases");
  double* red = sh.red;
  double* sh_key = sh.sh_key;
  int* sh_idx = sh.sh_idx;
  double2* tri_de = sh.tri_de;
  double* tri_e = sh.tri_e;
  const int P = n * (n - 1) / 2;
  const int tid = threadIdx.x;
  const int wv = tid >> 6, ln = tid & 63;
  constexpr int NW = EMI_GEN_THREADS / 64;
  double* A = reinterpret_cast<double*>(base);  // the first half of M's space
  c128* M = reinterpret_cast<c128*>(base);
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
  for (int i = wv; i < n; i += NW)
    for (int j = ln; j < n; j += 64) {
      double a = 1.0;
      if (i != j) {
        const Z c = i < j ? src(i, j) : src(j, i);
        const double gr = c.x, gi = c.y;
        a = sqrt(gr * gr + gi * gi);
      }
      A[(size_t)i * n + j] = a;
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
    for (int i = wv; i < n; i += NW) {
      const double f = (i == k) ? (1.0 - rd) : col[i].re * rd;
      for (int j = ln; j < n; j += 64) A[(size_t)i * n + j] = fma(-f, col[j].re, A[(size_t)i * n + j]);
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
  // M = inv(A) o Gamma (A holds -inv), written over A: chunks of EMI_GEN_THREADS
  // entries from the last down; chunk c's entries of M cover A's entries 2t, 2t + 1,
  // in chunks above c, so every A entry is read before it is overwritten
  {
    const int nn = n * n;
    for (int c = (nn - 1) / EMI_GEN_THREADS; c >= 0; --c) {
      const int t = c * EMI_GEN_THREADS + tid;
      double ai = 0.0, gr = 0.0, gi = 0.0;
      if (t < nn) {
        const int i = t / n, j = t - (t / n) * n;
        ai = -A[t];
        if (i == j) {
          gr = 1.0;
        } else if (j > i) {
          const Z g = src(i, j);
          gr = g.x;
          gi = g.y;
        } else {
          const Z g = src(j, i);
          gr = g.x;
          gi = -g.y;
        }
      }
      __syncthreads();
      if (t < nn) M[t] = {ai * gr, ai * gi};
    }
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
    // p = M u: four threads per row (columns k+1+r, k+5+r, ...), summed across the four
    for (int i = tid >> 2; i < n; i += EMI_GEN_THREADS / 4) {
      c128 p = {0.0, 0.0};
      if (i > k)
        for (int j = k + 1 + (tid & 3); j < n; j += 4) {
          const c128 m = M[(size_t)i * n + j];
          p.re += m.re * u[j].re - m.im * u[j].im;
          p.im += m.re * u[j].im + m.im * u[j].re;
        }
      p.re += __shfl_xor(p.re, 1);
      p.im += __shfl_xor(p.im, 1);
      p.re += __shfl_xor(p.re, 2);
      p.im += __shfl_xor(p.im, 2);
      if ((tid & 3) == 0) pv[i] = p;
    }
    __syncthreads();
    double cp = 0.0;
    for (int i = tid; i < n; i += EMI_GEN_THREADS) cp += u[i].re * pv[i].re + u[i].im * pv[i].im;
    const double cval = block_sum(cp, red);
    for (int i = tid; i < n; i += EMI_GEN_THREADS)
      q[i] = (i > k) ? c128{2.0 * (pv[i].re - cval * u[i].re), 2.0 * (pv[i].im - cval * u[i].im)} : c128{0.0, 0.0};
    __syncthreads();
    for (int i = k + 1 + wv; i < n; i += NW)
      for (int j = k + 1 + ln; j < n; j += 64) {
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
      *q_out = (float)rq;
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
      if (tid == 0) *q_out = (float)rq;
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
    block_sum2(sr, si, sh_key);  // sh_key: free after the sweep
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
    ph_row[i] = v;
    if constexpr (TC) reinterpret_cast<float2*>(wk)[i] = v;  // LU work space is free again
  }
  __syncthreads();
  if constexpr (TC) {
    if (tid < 64) {
      const float2* phl = reinterpret_cast<const float2*>(wk);
      const float tc = tcoh_point<true>([&](int p, int a, int b) { return src.pair(p, a, b); }, [&](int i) { return phl[i]; },
                                        pairs, P, tid);
      if (tid == 0) *t_out = tc;
    }
  }
  __syncthreads();
}

}  // namespace mr
