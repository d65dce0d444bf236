// EMI phase linking of one point for n <= 32 and 64 < n <= 200 on gfx950: one
// workgroup per point, two lanes per matrix row.  Same algorithm and precision
// split as the one-wave kernel (moraine/pl.py:45-65 restated: inv(|Gamma|) by
// the FP64 symmetric sweep with largest-|diagonal| pivoting, M = inv(|Gamma|) o
// Gamma, FP32 Householder tridiagonalisation, FP64 multisection + twisted
// factorisation for the smallest eigenpair, back-transformation,
// ph = v conj(v_ref)/|.|).
//
// Layout: thread t holds row i = t / R and the interleaved columns
// j = R s + (t % R), s = 0..E-1, in VGPRs (FP64 during the sweep, FP32 complex
// afterwards), so a workgroup of R n lanes keeps the whole matrix in registers:
// R = 2 with E = 16, 48, 64, 80, 100 for n <= 32, 96, 128, 160, 200 (n <= 32 is a
// single wave, the others 3-7 waves).  R = 4 (E = 40, 10 waves at 156 VGPRs,
// three per SIMD) was measured at n = 150: 6% slower than R = 2 (one 5-wave
// block per CU), the extra barriers costing more than the occupancy gained, so
// it is not dispatched.
// The lanes of a row are adjacent (one DPP quad): row partial sums combine with
// DPP exchanges; column vectors are broadcast through LDS; reductions over rows
// are wave DPP sums plus one LDS slot per wave.  Two workgroup barriers per
// sweep step and per Householder step, one per back-transformation step.
//
// The coherence of the point comes through an accessor src(a, b), a < b: a
// triu row in global memory (EwPairRow; the drop-in kernel in emi_wide.hip) or
// the coherence the batched form's fused kernels keep on chip (batch.hip).
// Either way the same values enter the same arithmetic.
#pragma once
#include "emi_common.hpp"

namespace mr {

constexpr int EW_MAX_WAVES = 10;
static_assert(EMI_WIDE_MAX_N <= 256, "pivot keys carry the row index in 8 bits");

// Rows are held by R adjacent lanes (R = 2 or 4, inside one DPP quad).
// rowsum: sum over the R lanes of a row; rowpick: the value held by lane h of
// the row (h uniform; every lane active -- a DPP read of a masked-off lane
// returns 0).  quad_perm [h,h,h,h] = 0x55 h, [h,h,2+h,2+h] = 0xA0 / 0xF5.
template <int R>
__device__ __forceinline__ float rowsum_f(float v) {
  v += dpp_f<DPP_XOR1>(v);
  if (R == 4) v += dpp_f<DPP_XOR2>(v);
  return v;
}
template <int R>
__device__ __forceinline__ int rowpick_i(int v, int h) {
  if (R == 2) return h == 0 ? dpp_i<0xA0>(v) : dpp_i<0xF5>(v);
  switch (h) {
    case 0:
      return dpp_i<0x00>(v);
    case 1:
      return dpp_i<0x55>(v);
    case 2:
      return dpp_i<0xAA>(v);
    default:
      return dpp_i<0xFF>(v);
  }
}
template <int R>
__device__ __forceinline__ float rowpick_f(float v, int h) {
  return __int_as_float(rowpick_i<R>(__float_as_int(v), h));
}
template <int R>
__device__ __forceinline__ double rowpick_d(double v, int h) {
  const long long b = __double_as_longlong(v);
  const int lo = rowpick_i<R>((int)(b & 0xffffffffll), h);
  const int hi = rowpick_i<R>((int)(b >> 32), h);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// LDS layout for NM = 2E columns (bytes):
//   doubles: colb[NM] td[NM] te[NM] dpl[NM] dmi[NM] zz[NM] red[2*16] misc[8]
//   float2 : xb[NM] ub[NM] pb[NM] dph[NM] hal[NM]
//   float  : hinv[NM] redf[2*32]
//   int    : redi[2*16]
__host__ __device__ inline size_t ew_lds_bytes(int nm) {
  return (size_t)8 * (6 * nm + 32 + 8) + (size_t)8 * 5 * nm + (size_t)4 * (nm + 64) + 4 * 32;
}

// LDS of the per-point body, carved from the workgroup's dynamic LDS base
// `smem` (offsets as in ew_lds_bytes).  Every pointer is formed from smem
// where it is used, so the compiler keeps them as LDS addresses.
#define MR_EW_LDS(NM, smem)                                                     \
  double* colb = reinterpret_cast<double*>(smem);                               \
  double* td = colb + (NM);                                                     \
  double* te = td + (NM);                                                       \
  double* dpl = te + (NM);                                                      \
  double* dmi = dpl + (NM);                                                     \
  double* zz = dmi + (NM);                                                      \
  double* red = zz + (NM);  /* [2][16] */                                       \
  double* misc = red + 32;  /* [8] */                                           \
  float2* xb = reinterpret_cast<float2*>(misc + 8);                             \
  float2* ub = xb + (NM);                                                       \
  float2* pb = ub + (NM);                                                       \
  float2* dph = pb + (NM);                                                      \
  float2* hal = dph + (NM);                                                     \
  float* hinv = reinterpret_cast<float*>(hal + (NM));                           \
  float* redf = hinv + (NM);                     /* [2][32] */                  \
  int* redi = reinterpret_cast<int*>(redf + 64); /* [2][16] */                  \
  (void)colb, (void)td, (void)te, (void)dpl, (void)dmi, (void)zz, (void)red, (void)misc, (void)xb, (void)ub, \
      (void)pb, (void)dph, (void)hal, (void)hinv, (void)redf, (void)redi

// Zeroes the column broadcast buffers, which the body reads beyond n (rows >= n
// never write them).  Once per kernel, before the first point; ends with a
// workgroup barrier.
template <int NM>
__device__ __forceinline__ void ew_lds_init(unsigned char* smem) {
  MR_EW_LDS(NM, smem);
  for (int t = threadIdx.x; t < NM; t += blockDim.x) {
    colb[t] = 0.0;
    td[t] = 0.0;  // second column buffer of the sweep
    xb[t] = make_float2(0.f, 0.f);
    ub[t] = make_float2(0.f, 0.f);
    pb[t] = make_float2(0.f, 0.f);
  }
  __syncthreads();
}

// Outputs of point pt of a dense batch: phases ph[pt * n + i], quality[pt].
struct EwDenseOut {
  float2* phs;
  float* qs;
  long long pt;
  int n;
  __device__ __forceinline__ void ph(int i, float2 v) const { phs[pt * n + i] = v; }
  __device__ __forceinline__ void quality(float v) const { qs[pt] = v; }
  __device__ __forceinline__ int id() const { return (int)pt; }
};

// A coherence row in triu pair order (moraine/tnet.py:15-20) in global memory.
struct EwPairRow {
  const float2* p;
  int n;
  __device__ __forceinline__ float2 operator()(int a, int b) const { return p[pair_index(a, b, n)]; }
};

// GS / GH: columns per software-pipelined LDS group in the sweep / Householder
// updates (smaller groups trade latency hiding for registers at large E).
// E <= 64 (n <= 128, blocks of three or four waves) is held to 168 VGPRs so
// three waves run per SIMD: 12 waves per CU instead of 8 (n = 66..96 1.6x
// faster; n = 97..128 1.15x, with a few registers spilled outside the
// Householder loops).  E = 16 (n <= 32, one wave per point) is held to four
// waves per SIMD.
// CHK: points whose conditioning test fails (tau; the test of the one-wave
// kernel, with the Gershgorin bound of T for ||T||) are appended to
// flag_list[*flag_count] (as out.id()) for the FP64 re-solve.
// Outputs go through `out`: out.ph(i, phase) for the linked phases (also
// written to ph_lds when given), out.quality(q).  The output addresses are
// formed only there, from values the caller keeps anyway (kernel arguments,
// the point index), so no extra pointer stays live through the solve.
template <int R, int E, int GS, int GH, bool CHK, class Src, class Out>
__device__ __forceinline__ void ew_point(const Src& src, int n, int ref, unsigned char* smem, const Out& out,
                                         double tau, unsigned* __restrict__ flag_count, int* __restrict__ flag_list,
                                         float2* ph_lds = nullptr) {
  constexpr int NM = R * E;
  static_assert(E % GS == 0 && E % GH == 0, "groups must tile the row");
  MR_EW_LDS(NM, smem);
  const int W = blockDim.x >> 6;
  const int wv = threadIdx.x >> 6;
  int tid = threadIdx.x;
  __asm__ volatile("" : "+v"(tid));  // keeps per-lane addresses from being hoisted out of the point loop
  const int lane = tid & 63;
  // Row held by this lane (R adjacent lanes per row slot).  In blocks of more
  // than four waves, wave w + 4 shares a SIMD with wave w; those extra waves
  // take the first n - S4 rows (reduced first) and waves 0..3 the last S4
  // rows, so a shared SIMD issues for two active waves during fewer
  // Householder steps.  Slots past the last row stay dead (row >= n).
  const int slot = tid / R, c = tid % R;
  constexpr int S4 = 4 * (64 / R);  // row slots of waves 0..3
  const int nx = W > 4 ? n - S4 : 0;
  const int row = slot < S4 ? slot + nx : (slot < S4 + nx ? slot - S4 : slot);
  const bool live = row < n;

  // ---- step 1: A = |Gamma| (FP64), sweep inverse --------------------------
  RegArr<double, E> a;
#pragma unroll
  for (int s = 0; s < E; ++s) {
    const int j = R * s + c;
    double v = 0.0;
    if (live && j < n) {
      if (j == row) {
        v = 1.0;
      } else {
        const float2 g = (j > row ? src(row, j) : src(j, row));
        v = sqrt((double)g.x * g.x + (double)g.y * g.y);
      }
    }
    a.set(s, v);
    if ((s & 7) == 7) __builtin_amdgcn_sched_barrier(0);  // bounds the loads in flight
  }
  const bool dlane = live && (c == row % R);  // this lane holds the diagonal entry of its row
  double dg = dlane ? 1.0 : 0.0;
  // Pivot: the unswept row with the largest |diagonal| (ties -> smallest
  // row).  Every wave finds the next pivot itself: the diagonal after step
  // st is d_i - (a_ik / a_kk) a_ik, computed from the broadcast column and
  // the diagonal (both in LDS) with the same operations the diagonal's
  // holder uses, so one workgroup barrier per step suffices.  The column,
  // diagonal and pivot buffers alternate between two copies by step parity
  // (td / te / dpl are free until the tridiagonal is formed); a wave can only
  // overwrite a copy after every wave has passed the next step's barrier.
  // The first pivot is row 0 (A's diagonal is all ones).
  unsigned long long sw0 = 0, sw1 = 0, sw2 = 0, sw3 = 0;  // swept rows (uniform)
  int k = 0;
  for (int st = 0; st < n; ++st) {
    const int b = st & 1;
    double* cb = b ? td : colb;
    double* dgl = b ? dpl : te;
    const bool holdk = c == k % R;
    const double own = a.dyn(k / R);
    const double aik = rowpick_d<R>(own, k % R);  // A[row][k] on every lane of the row
    // d-1 in place of the pivot: the rank-1 update then writes column k itself
    // (c - (c/d)(d-1) = c/d); the pivot's own diagonal becomes 2 - 1/d, which
    // the sweep never reads again and which is corrected when M is formed.
    if (live && holdk) cb[row] = (row == k) ? aik - 1.0 : aik;
    if (live && holdk && row == k) misc[6 + b] = aik;
    if (dlane) dgl[row] = dg;
    __syncthreads();
    const double rd = rcp_d(misc[6 + b]);
    const double f = !live ? 0.0 : (row == k) ? (1.0 - rd) : aik * rd;
    {
      const unsigned long long bit = 1ull << (k & 63);
      const int wk = k >> 6;
      sw0 |= wk == 0 ? bit : 0ull;
      sw1 |= wk == 1 ? bit : 0ull;
      sw2 |= wk == 2 ? bit : 0ull;
      sw3 |= wk == 3 ? bit : 0ull;
    }
    int key = -1;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int i = lane + 64 * j;
      const unsigned long long swj = j == 0 ? sw0 : j == 1 ? sw1 : j == 2 ? sw2 : sw3;
      if (64 * j < n && i < n && !((swj >> lane) & 1ull)) {
        const double ac = cb[i];
        key = max(key, pivot_key<8>(fabs(fma(-(ac * rd), ac, dgl[i])), true, i));
      }
    }
    const int kn = wmax_i(key);
    {
      double cur[GS], nxt[GS];
#pragma unroll
      for (int t = 0; t < GS; ++t) cur[t] = cb[R * t + c];
#pragma unroll
      for (int g = 0; g < E / GS; ++g) {
        if (g + 1 < E / GS) {
#pragma unroll
          for (int t = 0; t < GS; ++t) nxt[t] = cb[R * ((g + 1) * GS + t) + c];
        }
#pragma unroll
        for (int t = 0; t < GS; ++t) a.set(g * GS + t, fma(-f, cur[t], a.get(g * GS + t)));
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int t = 0; t < GS; ++t) cur[t] = nxt[t];
      }
    }
    dg = fma(-f, aik, dg);
    k = min(__builtin_amdgcn_readfirstlane(255 - (kn & 255)), n - 1);
  }
  // a holds -inv(A) (diagonal entries offset by +2)

  // ---- step 2: M = inv(A) o Gamma (FP32 complex) ---------------------------
  RegArr<float, E> mr_, mi_;
  {
  // fresh opaque copies of row and column parity: the E column indices and
  // pair offsets of step 1 are recomputed here instead of being kept live in
  // VGPRs through the sweep
  int row2 = row, c2 = c;
  __asm__ volatile("" : "+v"(row2), "+v"(c2));
#pragma unroll
  for (int s = 0; s < E; ++s) {
    const int j = R * s + c2;
    float re = 0.f, im = 0.f;
    if (live && j < n) {
      const double ai = (j == row2) ? 2.0 - a.get(s) : -a.get(s);
      if (j == row2) {
        re = (float)ai;
      } else if (j > row2) {
        const float2 g = src(row2, j);
        re = (float)(ai * g.x);
        im = (float)(ai * g.y);
      } else {
        const float2 g = src(j, row2);
        re = (float)(ai * g.x);
        im = (float)(-ai * g.y);
      }
    }
    mr_.set(s, re);
    mi_.set(s, im);
    if ((s & 7) == 7) __builtin_amdgcn_sched_barrier(0);
  }
  }

  // ---- step 3: Householder tridiagonalisation -----------------------------
  // p = M u is formed from x before ||x|| is known:
  //   u = inv (x - alpha e_{k+1})  =>  M u = inv (M x - alpha M[:, k+1]),
  // so one barrier publishes x and the norm partials, the next publishes
  // u, 2p and the u^H p partials; q = 2p - 2 (u^H p) u is formed per column.
  float dcr = 1.f, dci = 0.f;  // running phase product D_k (uniform)
  if (threadIdx.x == 0) dph[0] = make_float2(1.f, 0.f);
  double* const L_misc = misc;
  float* const L_redf = redf;
  // last matrix row held by this wave: once it is reduced (<= k) the wave's
  // product and update terms are all zero, so it skips both loops and only
  // takes part in the reductions and barriers
  const int wave_last = wmax_i(live ? row : -1);
  for (int k = 0; k + 2 < n; ++k) {
    // per-step opaque copy of n: the column-group activity tests below are
    // recomputed each step (scalar compares) instead of being hoisted out of
    // the loop as E/GH live masks
    int ns = n;
    __asm__ volatile("" : "+s"(ns));
    // per-step base of the step's broadcast slots (an opaque zero: formed here
    // each step rather than held through the loop, where the register budget
    // of the E = 64 kernel would push it to scratch)
    int z0 = 0;
    __asm__ volatile("" : "+v"(z0));
    double* misc = L_misc + z0;
    float* redf = L_redf + z0;
    const float ownr = mr_.dyn(k / R), owni = mi_.dyn(k / R);
    const float mrk = rowpick_f<R>(ownr, k % R), mik = rowpick_f<R>(owni, k % R);  // M[row][k]
    const bool below = live && row > k;
    const float xr = below ? mrk : 0.f, xi = below ? mik : 0.f;
    {
      const float ws = wsum_f(c == 0 ? xr * xr + xi * xi : 0.f);
      if (lane == 0) redf[wv] = ws;
    }
    if (live && c == 0) {
      xb[row] = make_float2(xr, xi);
      if (row == k) misc[1] = (double)mrk;
      if (row == k + 1) {
        misc[2] = (double)xr;
        misc[3] = (double)xi;
      }
    }
    __syncthreads();
    float nrm2 = redf[0];
    for (int w2 = 1; w2 < W; ++w2) nrm2 += redf[w2];
    const float dkk = (float)misc[1], x0r = (float)misc[2], x0i = (float)misc[3];
    const float nx = __builtin_amdgcn_sqrtf(nrm2);
    const float ax0 = __builtin_amdgcn_sqrtf(x0r * x0r + x0i * x0i);
    const float rax0 = ax0 > 0.f ? __builtin_amdgcn_rcpf(ax0) : 0.f;
    const float phr = ax0 > 0.f ? x0r * rax0 : 1.f, phi = ax0 > 0.f ? x0i * rax0 : 0.f;
    const float alr = -phr * nx, ali = -phi * nx;  // alpha = -phase(x0) ||x||
    const float den2 = 2.f * nx * (nx + ax0);
    const float inv = den2 > 0.f ? __builtin_amdgcn_rsqf(den2) : 0.f;
    {
      const float sr = nx > 0.f ? -phr : 1.f, si = nx > 0.f ? -phi : 0.f;  // alpha/|alpha|
      const float r = dcr * sr - dci * si, i = dcr * si + dci * sr;
      dcr = r;
      dci = i;
    }
    if (threadIdx.x == 0) {
      td[k] = dkk;
      te[k] = nx;
      hal[k] = make_float2(alr, ali);
      hinv[k] = inv;
      dph[k + 1] = make_float2(dcr, dci);
    }
    // s = partial of M x over this lane's columns; m1 = M[row][k+1] (from its holder)
    float sr = 0.f, si = 0.f;
    if (wave_last > k) {
      float2 cu[GH], nu[GH];
#pragma unroll
      for (int t = 0; t < GH; ++t) cu[t] = xb[R * t + c];
#pragma unroll
      for (int g = 0; g < E / GH; ++g) {
        const bool act = (R * GH * g + R * GH - 1 > k) && (R * GH * g < ns);
        if (g + 1 < E / GH) {
#pragma unroll
          for (int t = 0; t < GH; ++t) nu[t] = xb[R * ((g + 1) * GH + t) + c];
        }
        if (act) {
#pragma unroll
          for (int t = 0; t < GH; ++t) {
            const int s = g * GH + t;
            sr = fmaf(mr_.get(s), cu[t].x, sr);
            sr = fmaf(-mi_.get(s), cu[t].y, sr);
            si = fmaf(mr_.get(s), cu[t].y, si);
            si = fmaf(mi_.get(s), cu[t].x, si);
          }
        }
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int t = 0; t < GH; ++t) cu[t] = nu[t];
      }
    }
    const float m1o_r = mr_.dyn((k + 1) / R), m1o_i = mi_.dyn((k + 1) / R);
    const float m1r = rowpick_f<R>(m1o_r, (k + 1) % R), m1i = rowpick_f<R>(m1o_i, (k + 1) % R);
    sr = rowsum_f<R>(sr);
    si = rowsum_f<R>(si);
    float pr = inv * (sr - (alr * m1r - ali * m1i));
    float pi = inv * (si - (alr * m1i + ali * m1r));
    float ur = (row == k + 1) ? (xr - alr) : xr;
    float ui = (row == k + 1) ? (xi - ali) : xi;
    ur = below ? ur * inv : 0.f;
    ui = below ? ui * inv : 0.f;
    if (!below) {
      pr = 0.f;
      pi = 0.f;
    }
    {
      const float ws = wsum_f(c == 0 ? ur * pr + ui * pi : 0.f);  // real(u^H p)
      if (lane == 0) redf[16 + wv] = ws;
    }
    if (live && c == 0) {
      ub[row] = make_float2(ur, ui);
      pb[row] = make_float2(2.f * pr, 2.f * pi);
    }
    __syncthreads();
    float cval = redf[16];
    for (int w2 = 1; w2 < W; ++w2) cval += redf[16 + w2];
    const float c2 = 2.f * cval;
    const float qr = 2.f * pr - c2 * ur, qi = 2.f * pi - c2 * ui;
    // M22 -= u q^H + q u^H  over this lane's active columns
    if (wave_last > k) {
      float2 cu[GH], cp[GH], nu[GH], np[GH];
#pragma unroll
      for (int t = 0; t < GH; ++t) {
        cu[t] = ub[R * t + c];
        cp[t] = pb[R * t + c];
      }
#pragma unroll
      for (int g = 0; g < E / GH; ++g) {
        const bool act = (R * GH * g + R * GH - 1 > k) && (R * GH * g < ns);
        if (g + 1 < E / GH) {
#pragma unroll
          for (int t = 0; t < GH; ++t) {
            nu[t] = ub[R * ((g + 1) * GH + t) + c];
            np[t] = pb[R * ((g + 1) * GH + t) + c];
          }
        }
        if (act) {
#pragma unroll
          for (int t = 0; t < GH; ++t) {
            const int s = g * GH + t;
            const float cqr = fmaf(-c2, cu[t].x, cp[t].x);
            const float cqi = fmaf(-c2, cu[t].y, cp[t].y);
            float r = mr_.get(s), i = mi_.get(s);
            r = fmaf(-ur, cqr, r);
            r = fmaf(-ui, cqi, r);
            i = fmaf(ur, cqi, i);
            i = fmaf(-ui, cqr, i);
            r = fmaf(-qr, cu[t].x, r);
            r = fmaf(-qi, cu[t].y, r);
            i = fmaf(qr, cu[t].y, i);
            i = fmaf(-qi, cu[t].x, i);
            mr_.set(s, r);
            mi_.set(s, i);
          }
        }
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int t = 0; t < GH; ++t) {
          cu[t] = nu[t];
          cp[t] = np[t];
        }
      }
    }
  }
  // trailing 2x2 block: rows n-2, n-1
  {
    const int ka = n - 2, kb = n - 1;
    const float va = mr_.dyn(ka / R);  // M[row][n-2] (real part) on its holder
    const float vai = mi_.dyn(ka / R);
    const float vb = mr_.dyn(kb / R);  // M[row][n-1] on its holder
    if (row == ka && c == ka % R) td[ka] = va;
    if (row == kb && c == kb % R) td[kb] = vb;
    if (row == kb && c == ka % R) {
      const float ae = sqrtf(va * va + vai * vai);
      const float sr = ae > 0.f ? va / ae : 1.f, si = ae > 0.f ? vai / ae : 0.f;
      te[ka] = ae;
      te[kb] = 0.0;
      dph[kb] = make_float2(dcr * sr - dci * si, dcr * si + dci * sr);
    }
  }
  __syncthreads();

  // ---- step 4a: smallest eigenvalue of T by multisection (64 W shifts) ---
  double lo, hi, scale;
  {
    double plo = 1e308, phi_ = -1e308, psc = 0.0;
    for (int i = lane; i < n; i += 64) {
      const double off = (i > 0 ? te[i - 1] : 0.0) + te[i];
      plo = fmin(plo, td[i] - off);
      phi_ = fmax(phi_, td[i] + off);
      psc = fmax(psc, fabs(td[i]) + off);
    }
    lo = wmin_d(plo);
    hi = wmax_d(phi_);
    const double sc = wmax_d(psc);
    scale = sc > 0.0 ? sc : 1.0;
    lo -= 1e-14 * scale;
    hi += 1e-14 * scale;
  }
  const double pivmin = 1e-300 + scale * scale * 1e-290;
  {
    const int G = 64 * W;
    const double rg = 1.0 / (double)(G + 1);
    // rounds until the bracket has shrunk by at least 65^5 ~ 1.2e9 (the
    // one-wave kernels' five rounds of 64 shifts): five for one or two
    // waves, four from three waves on
    const int rounds = W >= 3 ? 4 : 5;
    for (int round = 0; round < rounds; ++round) {
      if (hi - lo <= 2e-15 * scale) break;
      const int g = wv * 64 + lane;
      const double x = lo + (hi - lo) * (double)(g + 1) * rg;
      double q = td[0] - x;
      int cnt = q < 0.0;
      for (int i = 1; i < n; ++i) {
        if (fabs(q) < pivmin) q = -pivmin;
        double r = __builtin_amdgcn_rcp(q);
        r = fma(r, fma(-q, r, 1.0), r);
        q = (td[i] - x) - (te[i - 1] * te[i - 1]) * r;
        cnt += q < 0.0;
      }
      const unsigned long long m = __ballot(cnt >= 1);
      if (lane == 0) redi[(round & 1) * 16 + wv] = m ? wv * 64 + __ffsll((long long)m) - 1 : G;
      __syncthreads();
      int L = G;
      for (int w2 = 0; w2 < W; ++w2) L = min(L, redi[(round & 1) * 16 + w2]);
      const double xl = lo + (hi - lo) * (double)(L + 1) * rg;
      const double xp = lo + (hi - lo) * (double)L * rg;  // shift L-1 (or lo itself for L = 0)
      if (L == G) {
        lo = lo + (hi - lo) * (double)G * rg;
      } else {
        if (L > 0) lo = xp;
        hi = xl;
      }
    }
  }
  const double lam = 0.5 * (lo + hi);

  // ---- step 4b: eigenvector of T by a twisted factorisation ---------------
  const double tiny = 2.2e-16 * scale + 1e-300;
  // the two serial recurrences run on lanes of different waves (same wave if W = 1)
  const int t2 = W > 1 ? 64 : 32;
  if (threadIdx.x == 0) {
    double dp = td[0] - lam;
    if (fabs(dp) < tiny) dp = tiny;
    dpl[0] = dp;
    for (int i = 1; i < n; ++i) {
      dp = (td[i] - lam) - te[i - 1] * te[i - 1] * rcp_d(dp);
      if (fabs(dp) < tiny) dp = tiny;
      dpl[i] = dp;
    }
  } else if (threadIdx.x == t2) {
    double dm = td[n - 1] - lam;
    if (fabs(dm) < tiny) dm = tiny;
    dmi[n - 1] = dm;
    for (int i = n - 2; i >= 0; --i) {
      dm = (td[i] - lam) - te[i] * te[i] * rcp_d(dm);
      if (fabs(dm) < tiny) dm = tiny;
      dmi[i] = dm;
    }
  }
  __syncthreads();
  int rtw;
  {
    const int i = tid;  // one row per thread for the twist search
    const double gam = i < n ? fabs(dpl[i] + dmi[i] - (td[i] - lam)) : 1e308;
    const int bl = wargmax_d(-gam, lane);
    if (lane == 0) {
      red[16 + wv] = -rl_d(gam, bl);
      redi[wv] = wv * 64 + bl;
    }
    __syncthreads();
    double best = red[16];
    rtw = redi[0];
    for (int w2 = 1; w2 < W; ++w2)
      if (red[16 + w2] > best) {
        best = red[16 + w2];
        rtw = redi[w2];
      }
    rtw = min(__builtin_amdgcn_readfirstlane(rtw), n - 1);
  }
  // z_r = 1; z_i = -(e_i / D+_i) z_{i+1} below r; z_{i+1} = -(e_i / D-_{i+1}) z_i above r
  if (threadIdx.x == 0) {
    double z = 1.0;
    zz[rtw] = 1.0;
    for (int i = rtw - 1; i >= 0; --i) {
      z *= -te[i] * rcp_d(dpl[i]);
      zz[i] = z;
    }
  } else if (threadIdx.x == t2) {
    double z = 1.0;
    for (int i = rtw + 1; i < n; ++i) {
      z *= -te[i - 1] * rcp_d(dmi[i]);
      zz[i] = z;
    }
  }
  __syncthreads();
  double rq, znorm;
  {
    const int i = tid;
    double zs = 0.0, ztz = 0.0;
    if (i < n) {
      const double zi = zz[i];
      double tz = td[i] * zi;
      if (i > 0) tz += te[i - 1] * zz[i - 1];
      if (i + 1 < n) tz += te[i] * zz[i + 1];
      zs = zi * zi;
      ztz = zi * tz;
    }
    zs = wsum_d(zs);
    ztz = wsum_d(ztz);
    if (lane == 0) {
      red[wv] = zs;
      red[16 + wv] = ztz;
    }
    __syncthreads();
    double s1 = 0.0, s2 = 0.0;
    for (int w2 = 0; w2 < W; ++w2) {
      s1 += red[w2];
      s2 += red[16 + w2];
    }
    rq = s2 / s1;
    znorm = 1.0 / sqrt(s1);
  }

  // ---- back-transformation v = H_0 ... H_{n-3} D z ------------------------
  float vr = 0.f, vi = 0.f;
  if (live) {
    const float zi = (float)(zz[row] * znorm);
    const float2 d = dph[row];
    vr = d.x * zi;
    vi = d.y * zi;
  }
  for (int k = n - 3; k >= 0; --k) {
    const float ownr = mr_.dyn(k / R), owni = mi_.dyn(k / R);
    const bool below = live && row > k;
    const float2 al = hal[k];
    const float hv = hinv[k];
    float ar = rowpick_f<R>(ownr, k % R), ai = rowpick_f<R>(owni, k % R);
    if (row == k + 1) {
      ar -= al.x;
      ai -= al.y;
    }
    ar = below ? ar * hv : 0.f;
    ai = below ? ai * hv : 0.f;
    const int b = k & 1;
    {
      const float s1 = wsum_f(c == 0 ? ar * vr + ai * vi : 0.f);
      const float s2 = wsum_f(c == 0 ? ar * vi - ai * vr : 0.f);
      if (lane == 0) {
        redf[b * 32 + wv] = s1;
        redf[b * 32 + 16 + wv] = s2;
      }
    }
    __syncthreads();
    float sr = redf[b * 32], si = redf[b * 32 + 16];
    for (int w2 = 1; w2 < W; ++w2) {
      sr += redf[b * 32 + w2];
      si += redf[b * 32 + 16 + w2];
    }
    vr -= 2.f * (ar * sr - ai * si);
    vi -= 2.f * (ar * si + ai * sr);
  }
  if constexpr (CHK) {
    // phase error estimate eps32 ||T|| c(n) / (gap min_i |v_i|) > tau  <=>
    // T has a second eigenvalue below lambda + eps32 ||T|| c(n) / (tau min_i |v_i|);
    // T is still in td / te, dpl is free since the twisted factorisation
    const double wm = wmin_d(live && c == 0 ? (double)sqrtf(vr * vr + vi * vi) : 1e300);
    if (lane == 0) dpl[wv] = wm;
    __syncthreads();
    double vmin = dpl[0];
    for (int w2 = 1; w2 < W; ++w2) vmin = fmin(vmin, dpl[w2]);
    const double x = rq + 5.9604644775390625e-8 * scale * emi_flag_scale(n) / (tau * fmax(vmin, 1e-300));
    if (wv == 0) {
      double q = td[0] - x;
      int cnt = q < 0.0;
      for (int i = 1; i < n; ++i) {
        if (fabs(q) < pivmin) q = -pivmin;
        double r = __builtin_amdgcn_rcp(q);
        r = fma(r, fma(-q, r, 1.0), r);
        q = (td[i] - x) - (te[i - 1] * te[i - 1]) * r;
        cnt += q < 0.0;
      }
      if (lane == 0 && cnt >= 2 && isfinite(x)) flag_list[atomicAdd(flag_count, 1u)] = out.id();
    }
  }
  // ---- output ---------------------------------------------------------------
  if (row == ref && c == 0) {
    misc[4] = vr;
    misc[5] = vi;
  }
  __syncthreads();
  const float rr = (float)misc[4], ri = (float)misc[5];
  if (live && c == 0) {
    const float pr = vr * rr + vi * ri;
    const float pim = vi * rr - vr * ri;
    const float am = sqrtf(pr * pr + pim * pim);
    const float2 pv = make_float2(pr / am, pim / am);
    out.ph(row, pv);
    if (ph_lds != nullptr) ph_lds[row] = pv;
  }
  if (threadIdx.x == 0) out.quality((float)rq);
  __syncthreads();
}

}  // namespace mr
