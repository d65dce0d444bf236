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
#include "co_mfma.hpp"
#include "emi_fast.hpp"
#include "emi_general.hpp"
#include "tcoh.hpp"

namespace mr {

// ----------------------------------------------------------------------------
// fast path: one wave per point, NM >= n, n <= 64
// ----------------------------------------------------------------------------
constexpr int EMI_FAST_THREADS = 64;

// TC: also the DS temporal coherence of each point from the staged row
// (moraine/pl.py:126-145), as mr_ds_temp_coh computes it from (coh, ph, pairs).
// CHK: points whose check fails (tau) are appended to flag_list[*flag_count]
// (MODE 1, 2: res_list[*res_count]) for the FP64 re-solve; MODE 0 and 2 test
// the conditioning of the Householder solve, MODE 1 the FP64 eigen-residual
// (each block keeps inv(|Gamma|) in its slot of slot_doubles doubles at xslots).
// MODE 0: the Householder formulation for every point.  MODE 1: the shift-invert
// formulation; a point it declines is appended to flag_list[*flag_count] and
// left unwritten.  MODE 2: the Householder formulation for the points of that
// list.
template <int NM, bool TC, bool CHK = false, int MODE = 0>
__global__ __launch_bounds__(EMI_FAST_THREADS) void emi_fast_kernel(const float2* __restrict__ coh, long long npts,
                                                                    int n, int ref, float2* __restrict__ ph,
                                                                    float* __restrict__ quality,
                                                                    const int32_t* __restrict__ pairs,
                                                                    float* __restrict__ t_coh, double tau = 0.0,
                                                                    unsigned* __restrict__ flag_count = nullptr,
                                                                    int* __restrict__ flag_list = nullptr,
                                                                    unsigned* __restrict__ res_count = nullptr,
                                                                    int* __restrict__ res_list = nullptr,
                                                                    double* __restrict__ xslots = nullptr,
                                                                    int slot_doubles = 0) {
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
      bool ok;
      if constexpr (CHK)
        ok = emi_inv_point_checked<NM>(CohPacked{cs, n}, n, lane, work, vr, vi, rq,
                                       xslots + (size_t)blockIdx.x * slot_doubles, tau, flag);
      else
        ok = emi_inv_point<NM>(CohPacked{cs, n}, n, lane, work, vr, vi, rq);
      if (!ok) {
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
    if (CHK && flag && lane == 0) {
      if constexpr (MODE != 0)
        res_list[atomicAdd(res_count, 1u)] = (int)pt;
      else
        flag_list[atomicAdd(flag_count, 1u)] = (int)pt;
    }
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

// The FP64 solve (emi_general.hpp), one block per point.  LIST: the FP64
// re-solve of flagged points: the points are list[0 .. *count); for n <= 64
// the block's slice lives in LDS, and with t_coh set each re-solved point's DS
// temporal coherence is recomputed from its new phases.  Z = double2: the
// complex128 form (coherence and phases complex128; every stage FP64; n <= 64
// with the slice in LDS).  Otherwise each block has a global slice in ws.
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
  const GenShared sh{red, sh_key, sh_idx, tri_de, tri_e};
  const int P = n * (n - 1) / 2;
  const bool small = kSmallPath && n <= 64;
  unsigned char* base = small ? gen_lds : ws + (size_t)blockIdx.x * emi_gen_slice_bytes(n);
  const long long nit = LIST ? (long long)*count : npts;
  for (long long it = blockIdx.x; it < nit; it += gridDim.x) {
    const long long pt = LIST ? (long long)list[it] : it;
    const GenRow<Z> src{coh + pt * (long long)P, n};
    if constexpr (LIST && sizeof(Z) == 8) {
      if (t_coh != nullptr) {
        emi_general_solve<true, Z>(src, n, ref, small, base, sh, ph + pt * n, quality + pt, pairs, t_coh + pt);
        continue;
      }
    }
    emi_general_solve<false, Z>(src, n, ref, small, base, sh, ph + pt * n, quality + pt, nullptr, nullptr);
  }
}

int emi_wide_launch(const float2* coh, long long npts, int n, int ref, float2* ph, float* quality, hipStream_t st,
                    double tau = 0.0, unsigned* flag_count = nullptr, int* flag_list = nullptr);

__global__ void emi_zero_count_kernel(unsigned* c) {
  c[0] = 0u;
  c[1] = 0u;
}

// Workspace of the two-pass path: two counts in the first 256 bytes, then the
// list of points the shift-invert pass leaves to the Householder pass, then
// the list of those the Householder pass flags for the FP64 re-solve.
inline size_t emi_list_span(long long npts) { return ((size_t)(npts > 0 ? npts : 0) * 4 + 255) & ~(size_t)255; }
inline size_t emi_list_bytes(long long npts) { return 256 + 2 * emi_list_span(npts); }

static size_t emi_fast_lds(int n, int nm) {
  const int P = n * (n - 1) / 2;
  return (((size_t)P * 8 + 15) & ~(size_t)15) + EMI_FAST_LDS_BYTES;
}

// 32 < n <= 64: the shift-invert pass over every point, then the Householder
// pass over the points it left (|Gamma| indefinite, sigma not below lambda_min,
// or no convergence: the poorly conditioned points), then the FP64 re-solve
// (emi_general_kernel<true>, slice in LDS) of the points the checks flag.  The
// Householder pass tests its conditioning at tau2 (EMI_HH_TOL in the plain
// EMI, the caller's tolerance in the precision form); with xslots (precision
// form) the shift-invert pass also checks its FP64 eigen-residual at tau1.
constexpr int EMI_FALLBACK_BLOCKS = 8192;
constexpr double EMI_HH_TOL = 1e-3;
constexpr int EMI_SI_SLOTS = 8192;  // blocks of the checked shift-invert pass, one slot each
inline int emi_slot_doubles(int n) { return ((n * (n + 1) / 2) + 31) & ~31; }
inline size_t emi_slots_bytes(long long npts, int n) {
  return (size_t)std::min<long long>(npts > 0 ? npts : 1, EMI_SI_SLOTS) * emi_slot_doubles(n) * 8;
}
template <bool TC>
static int emi_hh_passes(const float2* c, long long n_pts, int n, int ref, float2* p, float* quality,
                         const int32_t* pairs, float* t_coh, unsigned* count, int* list, unsigned* res_count,
                         int* res_list, double tau2, hipStream_t st);
template <bool TC>
static int emi_two_pass(const float2* c, long long n_pts, int n, int ref, float2* p, float* quality,
                        const int32_t* pairs, float* t_coh, unsigned* count, int* list, unsigned* res_count,
                        int* res_list, double* xslots, double tau1, double tau2, hipStream_t st) {
  const size_t lds = emi_fast_lds(n, 64);
  hipLaunchKernelGGL(emi_zero_count_kernel, dim3(1), dim3(1), 0, st, count < res_count ? count : res_count);
  if (int rc = launch_status("emi count init")) return rc;
  // the shift-invert pass keeps NM columns per lane: 60 up to n = 60 (the
  // padded columns would cost 1/16 of its FMAs), else 64
#define MR_EMI_SI(NM)                                                                                               \
  if (xslots != nullptr) {                                                                                          \
    const int nb = (int)std::min<long long>(n_pts, EMI_SI_SLOTS);                                                   \
    hipLaunchKernelGGL((emi_fast_kernel<NM, TC, true, 1>), dim3(nb), dim3(EMI_FAST_THREADS), lds, st, c, n_pts, n,  \
                       ref, p, quality, pairs, t_coh, tau1, count, list, res_count, res_list, xslots,              \
                       emi_slot_doubles(n));                                                                        \
  } else {                                                                                                          \
    const int nb = (int)std::min<long long>(n_pts, 1 << 20);                                                        \
    hipLaunchKernelGGL((emi_fast_kernel<NM, TC, false, 1>), dim3(nb), dim3(EMI_FAST_THREADS), lds, st, c, n_pts, n, \
                       ref, p, quality, pairs, t_coh, 0.0, count, list, nullptr, nullptr, nullptr, 0);             \
  }
  if (n <= 60) {
    MR_EMI_SI(60)
  } else {
    MR_EMI_SI(64)
  }
#undef MR_EMI_SI
  if (int rc = launch_status("emi_fast shift-invert")) return rc;
  return emi_hh_passes<TC>(c, n_pts, n, ref, p, quality, pairs, t_coh, count, list, res_count, res_list, tau2, st);
}

// Passes 2 and 3 of the two-pass path: the Householder pass over the points
// the shift-invert pass listed (list[0 .. *count)), then the FP64 re-solve of
// those its conditioning test (tau2) flags (res_list[0 .. *res_count)).
template <bool TC>
static int emi_hh_passes(const float2* c, long long n_pts, int n, int ref, float2* p, float* quality,
                         const int32_t* pairs, float* t_coh, unsigned* count, int* list, unsigned* res_count,
                         int* res_list, double tau2, hipStream_t st) {
  const size_t lds = emi_fast_lds(n, 64);
  const int nb2 = (int)std::min<long long>(n_pts, EMI_FALLBACK_BLOCKS);
  hipLaunchKernelGGL((emi_fast_kernel<64, TC, true, 2>), dim3(nb2), dim3(EMI_FAST_THREADS), lds, st, c, n_pts, n,
                     ref, p, quality, pairs, t_coh, tau2, count, list, res_count, res_list, nullptr, 0);
  if (int rc = launch_status("emi_fast householder")) return rc;
  const int nb3 = (int)std::min<long long>(n_pts, EMI_GEN_BLOCKS);
  hipLaunchKernelGGL(emi_general_kernel<true>, dim3(nb3), dim3(EMI_GEN_THREADS), emi_gen_slice_bytes(n), st, c,
                     n_pts, n, ref, p, quality, nullptr, res_list, res_count, TC ? pairs : nullptr,
                     TC ? t_coh : nullptr);
  return launch_status("emi FP64 re-solve");
}

// The plain two-pass path's workspace: the Householder list's count at ws[0],
// the re-solve count at ws[4], the lists after the first 256 bytes.
template <bool TC>
static int emi_two_pass_plain(const float2* c, long long n_pts, int n, int ref, float2* p, float* quality,
                              const int32_t* pairs, float* t_coh, unsigned char* ws, hipStream_t st) {
  unsigned* count = reinterpret_cast<unsigned*>(ws);
  return emi_two_pass<TC>(c, n_pts, n, ref, p, quality, pairs, t_coh, count, reinterpret_cast<int*>(ws + 256),
                          count + 1, reinterpret_cast<int*>(ws + 256 + emi_list_span(n_pts)), nullptr, 0.0, EMI_HH_TOL,
                          st);
}

// ----------------------------------------------------------------------------
// The batch driver's step in one launch (32 < n <= 64; moraine/pl.py:148-168:
// emperical_co_pc -> emi -> ds_temp_coh on the full pair network): one wave
// per point computes the point's coherence on MFMA (the point-cloud coherence
// kernel's body, co_mfma.hpp) straight into an LDS row in triu order -- the
// values that kernel stores in HBM -- then runs the shift-invert EMI pass and
// the DS temporal coherence on that row, as emi_fast_kernel<NM, true, false, 1>
// does after staging the row from HBM.  Points the pass declines get their row
// written to coh (the only coherence that reaches HBM) and go to the
// Householder pass and FP64 re-solve (emi_hh_passes), which read it there.
// The outputs equal emperical_co_pc -> emi -> ds_temp_coh bit for bit.  The
// waves of a CU interleave the coherence's gathers with other points' EMI
// arithmetic.  Point order: the coherence kernel's XCD ranges and strips, so
// neighbouring windows share an XCD's L2.  LDS per wave: the coherence row
// (which holds the SHP list while the coherence accumulates), scales, EMI
// scratch.
// ----------------------------------------------------------------------------
__host__ __device__ inline size_t coe_row_bytes(int n, int win) {
  const size_t row = (size_t)(n * (n - 1) / 2) * 8, lst = (size_t)win * 4;
  return ((row > lst ? row : lst) + 15) & ~(size_t)15;
}
inline size_t coe_lds(int n, int win) { return coe_row_bytes(n, win) + 128 * 4 + EMI_FAST_LDS_BYTES; }

template <int NM>
__global__ __launch_bounds__(EMI_FAST_THREADS) void co_emi_pc_kernel(
    const float2* __restrict__ rslc, int nl, int w, int n, const int32_t* __restrict__ idx,
    const uint8_t* __restrict__ shp, long long npts, int aw, int rw, int ref, const int32_t* __restrict__ pairs,
    float2* __restrict__ coh, float2* __restrict__ ph, float* __restrict__ quality, float* __restrict__ t_coh,
    unsigned* __restrict__ count, int* __restrict__ list) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int P = n * (n - 1) / 2;
  const int win = aw * rw, az = (aw - 1) / 2, r = (rw - 1) / 2;
  float2* cs = reinterpret_cast<float2*>(smem);  // coherence row (the SHP list before it)
  int* shl = reinterpret_cast<int*>(smem);
  float* sc = reinterpret_cast<float*>(smem + coe_row_bytes(n, win));  // scales, 1/sqrt(diag)
  unsigned char* work = reinterpret_cast<unsigned char*>(sc + 128);     // EMI_FAST_LDS_BYTES
  // XCD x (block b runs on XCD b % 8) takes the x-th of eight contiguous
  // ranges of points, in the coherence kernel's strip order (grid: a multiple of 8)
  const unsigned nbx = gridDim.x / 8, xcd = blockIdx.x % 8, jb = blockIdx.x / 8;
  const long long per = (npts + 7) / 8;
  const long long beg = (long long)xcd * per, end = beg + per < npts ? beg + per : npts;
  const CoStrips st = co_strips(idx, beg, end, w);
  for (long long sl = jb; sl < st.slots(end - beg); sl += nbx) {
    const long long u = st.point(sl, end - beg);
    if (u < 0) continue;  // a slot of the strip grid past the run
    const long long pt = beg + u;
    int lane = threadIdx.x;
    __asm__ volatile("" : "+v"(lane));  // per-point opaque lane (see emi_fast_kernel)
    const int ia = idx[2 * pt], ir = idx[2 * pt + 1];
    const int K = co_compact_shp(shp + pt * win, win, rw, ia, ir, az, r, nl, w, lane, shl);
    __builtin_amdgcn_wave_barrier();
    const CoAcc acc = co_mfma_accumulate<true>(rslc, RowList{shl}, K, n, lane);
    __builtin_amdgcn_wave_barrier();  // the row below overwrites the list
    co_store_scales<true>(acc, n, lane, sc);
    {
      int ln = lane;
      __asm__ volatile("" : "+v"(ln));
      co_write_tile_direct(acc.r00, acc.i00, 0, 0, n, ln, sc + 64, cs);
      co_write_tile_direct(acc.r01, acc.i01, 0, 1, n, ln, sc + 64, cs);
      co_write_tile_direct(acc.r11, acc.i11, 1, 1, n, ln, sc + 64, cs);
    }
    __builtin_amdgcn_wave_barrier();
    float vr, vi;
    double rq;
    if (!emi_inv_point<NM>(CohPacked{cs, n}, n, lane, work, vr, vi, rq)) {
      // declined: the row goes to HBM for the Householder pass
      float2* dst = coh + pt * (long long)P;
      for (int t = lane; t < P; t += 64) dst[t] = cs[t];
      if (lane == 0) list[atomicAdd(count, 1u)] = (int)pt;
      __builtin_amdgcn_wave_barrier();
      continue;
    }
    const float2 pv = emi_link_phase(vr, vi, ref);
    if (lane < n) ph[pt * n + lane] = pv;
    if (lane == 0) quality[pt] = (float)rq;
    float2* phb = reinterpret_cast<float2*>(work);  // the EMI scratch is free again
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    __builtin_amdgcn_wave_barrier();
    phb[lane] = pv;
    __builtin_amdgcn_wave_barrier();
    const float tc = tcoh_point<true>([&](int q, int, int) { return cs[q]; }, [&](int i) { return phb[i]; }, pairs, P,
                                      lane);
    if (lane == 0) t_coh[pt] = tc;
    __builtin_amdgcn_wave_barrier();
  }
}

// the full pair network of n images in triu order (moraine/tnet.py:15-20)
__global__ void full_pairs_kernel(int n, int32_t* __restrict__ pairs) {
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= n) return;
  int p = a * (2 * n - a - 1) / 2;
  for (int b = a + 1; b < n; ++b, ++p) {
    pairs[2 * p] = a;
    pairs[2 * p + 1] = b;
  }
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
    return mr::emi_two_pass_plain<false>(c, n_pts, n, ref, p, quality, nullptr, nullptr,
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
    return mr::emi_two_pass_plain<true>(reinterpret_cast<const float2*>(coh), n_pts, n, ref,
                                        reinterpret_cast<float2*>(ph), quality, pairs, t_coh,
                                        reinterpret_cast<unsigned char*>(workspace), mr::as_stream(stream));
  }
  const int rc = mr_emi(coh, n_pts, n, ref, ph, quality, workspace, workspace_bytes, stream);
  if (rc != MR_OK) return rc;
  return mr_ds_temp_coh(coh, ph, n_pts, n, pairs, n_pairs, t_coh, stream);
}

static size_t coe_round(size_t b) { return (b + 255) & ~(size_t)255; }

size_t mr_co_emi_ds_temp_coh_pc_workspace_bytes(long long n_pc, int n) {
  if (n < 2 || n_pc < 0) return 0;
  const size_t P = (size_t)n * (n - 1) / 2;
  return coe_round(mr_emi_workspace_bytes(n_pc, n)) + coe_round(P * 8) + (size_t)n_pc * P * 8;
}

int mr_co_emi_ds_temp_coh_pc(const float* rslc, int nl, int w, int n, const int32_t* idx, const uint8_t* pc_is_shp,
                             long long n_pc, int aw, int rw, int ref, float* ph, float* quality, float* t_coh,
                             void* workspace, size_t workspace_bytes, mr_stream_t stream) {
  MR_CHECK_ARG(n >= 2, "co_emi_ds_temp_coh_pc: need at least 2 images (n=%d)", n);
  MR_CHECK_ARG(ref >= 0 && ref < n, "co_emi_ds_temp_coh_pc: ref=%d outside [0,%d)", ref, n);
  MR_CHECK_ARG(n_pc >= 0 && n_pc < (1ll << 31), "co_emi_ds_temp_coh_pc: point count %lld outside [0, 2^31)", n_pc);
  MR_CHECK_ARG(nl >= 0 && w >= 0, "co_emi_ds_temp_coh_pc: negative raster size");
  MR_CHECK_ARG(aw >= 1 && rw >= 1 && aw * rw <= 4096, "co_emi_ds_temp_coh_pc: window %dx%d unsupported (<= 4096 entries)",
               aw, rw);
  if (n_pc == 0) return MR_OK;
  MR_CHECK_ARG(nl > 0 && w > 0, "co_emi_ds_temp_coh_pc: empty raster (%d x %d) with %lld points", nl, w, n_pc);
  MR_CHECK_ARG(rslc && idx && pc_is_shp && ph && quality && t_coh, "co_emi_ds_temp_coh_pc: null pointer");
  const size_t need = mr_co_emi_ds_temp_coh_pc_workspace_bytes(n_pc, n);
  MR_CHECK_ARG(workspace != nullptr && workspace_bytes >= need, "co_emi_ds_temp_coh_pc: workspace %zu < %zu bytes",
               workspace_bytes, need);
  hipStream_t st = mr::as_stream(stream);
  const int P = n * (n - 1) / 2;
  unsigned char* ws = reinterpret_cast<unsigned char*>(workspace);
  const size_t ew = mr_emi_workspace_bytes(n_pc, n);
  int32_t* pairs = reinterpret_cast<int32_t*>(ws + coe_round(ew));
  float2* coh = reinterpret_cast<float2*>(ws + coe_round(ew) + coe_round((size_t)P * 8));
  hipLaunchKernelGGL(mr::full_pairs_kernel, dim3((n + 63) / 64), dim3(64), 0, st, n, pairs);
  if (int rc = mr::launch_status("co_emi_ds_temp_coh_pc pairs")) return rc;
  if (!(n > 32 && n <= 64)) {
    // the stages one after the other through the coherence buffer
    if (int rc = mr_emperical_co_pc(rslc, nl, w, n, idx, pc_is_shp, (int)n_pc, aw, rw, pairs, P,
                                    reinterpret_cast<float*>(coh), stream))
      return rc;
    return mr_emi_ds_temp_coh(reinterpret_cast<const float*>(coh), n_pc, n, ref, pairs, P, ph, quality, t_coh, ws, ew,
                              stream);
  }
  unsigned* count = reinterpret_cast<unsigned*>(ws);
  int* list = reinterpret_cast<int*>(ws + 256);
  int* res_list = reinterpret_cast<int*>(ws + 256 + mr::emi_list_span(n_pc));
  hipLaunchKernelGGL(mr::emi_zero_count_kernel, dim3(1), dim3(1), 0, st, count);
  if (int rc = mr::launch_status("co_emi_ds_temp_coh_pc count init")) return rc;
  // one wave per slot of the coherence kernel's strip order (a multiple of 8 blocks)
  const int win = aw * rw;
  const long long per = (n_pc + 7) / 8;
  const long long waves = per < 2 * (long long)w ? n_pc : 8 * mr::co_max_slots(per, w);
  const long long nb = (std::min<long long>(waves, 1 << 20) + 7) / 8 * 8;
  const float2* z = reinterpret_cast<const float2*>(rslc);
  float2* p = reinterpret_cast<float2*>(ph);
  if (n <= 60)
    hipLaunchKernelGGL((mr::co_emi_pc_kernel<60>), dim3((unsigned)nb), dim3(mr::EMI_FAST_THREADS), mr::coe_lds(n, win),
                       st, z, nl, w, n, idx, pc_is_shp, n_pc, aw, rw, ref, pairs, coh, p, quality, t_coh, count, list);
  else
    hipLaunchKernelGGL((mr::co_emi_pc_kernel<64>), dim3((unsigned)nb), dim3(mr::EMI_FAST_THREADS), mr::coe_lds(n, win),
                       st, z, nl, w, n, idx, pc_is_shp, n_pc, aw, rw, ref, pairs, coh, p, quality, t_coh, count, list);
  if (int rc = mr::launch_status("co_emi_pc")) return rc;
  return mr::emi_hh_passes<true>(coh, n_pc, n, ref, p, quality, pairs, t_coh, count, list, count + 1, res_list,
                                 mr::EMI_HH_TOL, st);
}

size_t mr_emi_refined_workspace_bytes(long long n_pts, int n) {
  const size_t list = 256 + (((size_t)(n_pts > 0 ? n_pts : 0) * 4 + 255) & ~(size_t)255);
  // 32 < n <= 64: the Householder-pass list and the shift-invert pass's FP64 slots
  if (n > 32 && n <= 64) return list + mr::emi_list_span(n_pts) + mr::emi_slots_bytes(n_pts, n);
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
  // the re-solve count at w[0] and its list after the first 256 bytes (every n)
  unsigned* count = reinterpret_cast<unsigned*>(w);
  int* list = reinterpret_cast<int*>(w + 256);
  const float2* c = reinterpret_cast<const float2*>(coh);
  float2* p = reinterpret_cast<float2*>(ph);
  if (n > 32 && n <= 64) {
    // the two-pass path with both checks at tol (the shift-invert pass's FP64
    // eigen-residual, the Householder pass's conditioning test)
    const size_t span = mr::emi_list_span(n_pts);
    unsigned* hh_count = count + 1;
    int* hh_list = reinterpret_cast<int*>(w + 256 + span);
    double* xslots = reinterpret_cast<double*>(w + 256 + 2 * span);
    // (the Householder pass at min(tol, EMI_HH_TOL): never looser than the plain EMI, so
    // the points this leaves unflagged keep the plain call's outputs)
    const double tol2 = std::min(tol, mr::EMI_HH_TOL);
    if (t_coh != nullptr)
      return mr::emi_two_pass<true>(c, n_pts, n, ref, p, quality, pairs, t_coh, hh_count, hh_list, count, list,
                                    xslots, tol, tol2, st);
    return mr::emi_two_pass<false>(c, n_pts, n, ref, p, quality, nullptr, nullptr, hh_count, hh_list, count, list,
                                   xslots, tol, tol2, st);
  }
  hipLaunchKernelGGL(mr::emi_zero_count_kernel, dim3(1), dim3(1), 0, st, count);
  if (int rc = mr::launch_status("emi_refined")) return rc;
  if (int rc = mr::emi_wide_launch(c, n_pts, n, ref, p, quality, st, tol, count, list)) return rc;
  // FP64 re-solve of the flagged points, one block per point (slice in LDS for
  // n <= 32, in the workspace after the list above for n > 64)
  if (n <= 64) {
    const int nb2 = (int)std::min<long long>(n_pts, mr::EMI_GEN_BLOCKS);
    hipLaunchKernelGGL(mr::emi_general_kernel<true>, dim3(nb2), dim3(mr::EMI_GEN_THREADS),
                       mr::emi_gen_slice_bytes(n), st, c, n_pts, n, ref, p, quality, nullptr, list, count, nullptr,
                       nullptr);
  } else {
    const int nb2 = (int)std::min<long long>(n_pts, mr::EMI_REF_BLOCKS);
    unsigned char* slices = w + 256 + (((size_t)n_pts * 4 + 255) & ~(size_t)255);
    hipLaunchKernelGGL(mr::emi_general_kernel<true>, dim3(nb2), dim3(mr::EMI_GEN_THREADS), 0, st, c, n_pts, n, ref,
                       p, quality, slices, list, count, nullptr, nullptr);
  }
  if (int rc = mr::launch_status("emi_refined")) return rc;
  if (t_coh) return mr_ds_temp_coh(coh, ph, n_pts, n, pairs, n_pairs, t_coh, stream);
  return MR_OK;
}

}  // extern "C"
