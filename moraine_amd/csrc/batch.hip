// Batched DS form for gfx950: many independent DS pixels, each with its own
// complex64 sample matrix (K_i rows of N images; the SHP stacks of the pixel),
// taken through coherence -> EMI -> temporal coherence.  This is the batch
// driver of moraine (moraine/pl.py:148-168: emperical_co_pc -> emi ->
// ds_temp_coh per batch) over ragged per-pixel inputs, grouped by N by the
// caller (one call per image count).
//
// Fused launches (blocks >= 0), one per stage chain, for every n <= 200:
//  * 32 < n <= 64: as in the plain EMI, a shift-invert pass, a Householder
//    pass over the pixels it declines (with the conditioning test) and an FP64
//    re-solve of the pixels that test flags.  Each pass computes the pixel's
//    coherence on MFMA (co_mfma.hpp) into LDS -- the shift-invert pass as the
//    packed triu row the plain EMI reads, the others as tiles -- and runs EMI
//    (emi_fast.hpp) and the temporal coherence (tcoh.hpp) from it.
//  * n <= 32: one wave per pixel, coherence into one LDS tile, EMI by the
//    two-lanes-per-row body (emi_wide.hpp, the drop-in's EMI for this n) reading
//    the tile, temporal coherence from the tile.
//  * 64 < n <= 200: one workgroup of ceil(n / 32) waves per pixel.  The waves
//    accumulate the coherence tiles (the per-stage coherence kernel's body,
//    co_mfma.hpp) and write the row to the workgroup's own slot of the
//    workspace; the EMI body (emi_wide.hpp) and the temporal coherence read it
//    there.  The slot is re-used for every pixel of the workgroup, so the row
//    stays in the L2 / Infinity Cache instead of going through an HBM buffer
//    of the whole chunk.
// Every stage shares its code with the drop-in kernels, so for a sample matrix
// made of a pixel's SHP stacks in window order the outputs equal
// emperical_co_pc -> emi -> ds_temp_coh exactly.
// n > 200, or any n with blocks < 0 (the per-stage mode): per chunk of pixels,
// the coherence kernel in sample-matrix mode (the drop-in's choice for n), EMI
// and the temporal-coherence kernel as separate launches through an HBM
// coherence buffer, then a scatter into the ragged outputs.  The per-stage mode
// is what the drop-in functions do for one group; it gives the same results as
// the fused launches.
#include "co_mfma.hpp"
#include "emi_fast.hpp"
#include "emi_general.hpp"
#include "emi_wide.hpp"
#include "tcoh.hpp"

namespace mr {

int co_samples_launch(const float2* samples, const long long* row_off, const int32_t* n_samp, const int32_t* pix,
                      long long pix0, long long cn, int n, const int32_t* pairs, int npairs, float2* coh,
                      hipStream_t st);
int co_mfma_samples_launch(const float2* samples, const long long* row_off, const int32_t* n_samp,
                           const int32_t* pix, long long pix0, long long cn, int n, const int32_t* pairs,
                           int npairs, float2* coh, hipStream_t st);
int co_mfma_wide_samples_launch(const float2* samples, const long long* row_off, const int32_t* n_samp,
                                const int32_t* pix, long long pix0, long long cn, int n, const int32_t* pairs,
                                int npairs, float2* coh, hipStream_t st, int waves);
constexpr int BATCH_DEFAULT_BLOCKS = 1 << 15;
constexpr int BATCH_FALLBACK_BLOCKS = 8192;

__global__ void batch_zero_count_kernel(unsigned* c) {
  c[0] = 0u;
  c[1] = 0u;
}
constexpr double BATCH_HH_TOL = 1e-3;  // the plain EMI's Householder-pass tolerance

__host__ __device__ inline size_t batch_fused_lds(bool two) {
  return (size_t)com_tile_floats(two) * 4 + 128 * 4 /* scales */ + EMI_FAST_LDS_BYTES;
}

// MODE 2 (the only mode): the Householder formulation for the pixels the
// shift-invert pass (ds_batch_si_kernel) listed, rebuilding their coherence
// tiles, appending those its conditioning test flags to res_list[*res_count]
// (the plain kernels' passes, so the outputs equal the staged path's bit for
// bit).
template <int NM, bool TWO, int MODE>
__global__ __launch_bounds__(64) void ds_batch_fused_kernel(const float2* __restrict__ samples,
                                                            const long long* __restrict__ row_off,
                                                            const int32_t* __restrict__ n_samp,
                                                            const long long* __restrict__ ph_off,
                                                            const int32_t* __restrict__ pix, long long npix, int n,
                                                            int ref, const int32_t* __restrict__ pairs, int npairs,
                                                            float2* __restrict__ ph, float* __restrict__ quality,
                                                            float* __restrict__ tcoh, unsigned* __restrict__ count,
                                                            int* __restrict__ list, unsigned* __restrict__ res_count,
                                                            int* __restrict__ res_list) {
  static_assert(MODE == 2, "the shift-invert pass is ds_batch_si_kernel");
  extern __shared__ __align__(16) unsigned char smem[];
  float2* tiles = reinterpret_cast<float2*>(smem);
  float* sc = reinterpret_cast<float*>(smem + (size_t)com_tile_floats(TWO) * 4);  // scales, 1/sqrt(diag)
  unsigned char* work = reinterpret_cast<unsigned char*>(sc + 128);  // EMI_FAST_LDS_BYTES
  float2* ub = reinterpret_cast<float2*>(work + 512);               // linked phases after EMI
  const long long nit = (long long)*count;
  for (long long it = blockIdx.x; it < nit; it += gridDim.x) {
    const long long t = (long long)list[it];
    int lane = threadIdx.x;
    __asm__ volatile("" : "+v"(lane));  // per-pixel opaque lane (see emi.hip)
    const int id = pix[t];
    co_mfma_point<TWO>(samples + row_off[id], RowIdentity{}, n_samp[id], n, lane, tiles, sc);
    const CohTiles<TWO> ct{tiles, sc + 64};
    float vr, vi;
    double rq;
    bool flag = false;
    emi_fast_point<NM, true>(ct, n, lane, work, vr, vi, rq, BATCH_HH_TOL, &flag);
    if (flag && lane == 0) res_list[atomicAdd(res_count, 1u)] = (int)t;
    const float2 pv = emi_link_phase(vr, vi, ref);
    __builtin_amdgcn_wave_barrier();
    ub[lane] = pv;  // linked phases for the temporal coherence
    if (lane < n) ph[ph_off[id] + lane] = pv;
    if (lane == 0) quality[id] = (float)rq;
    __builtin_amdgcn_wave_barrier();
    const float tc = tcoh_point<true>([&](int, int a, int b) { return ct(a, b); }, [&](int i) { return ub[i]; },
                                      pairs, npairs, lane);
    if (lane == 0) tcoh[id] = tc;
    __builtin_amdgcn_wave_barrier();
  }
}

// The shift-invert pass of the fused launch (32 < n <= 64), every pixel: the
// coherence accumulates on MFMA and its output pass writes the pixel's triu
// row straight from the accumulators into LDS -- the values the tiles give,
// in the layout the plain EMI reads (CohPacked) -- then the shift-invert EMI
// and the temporal coherence read that row.  A pixel the pass declines goes to
// list[*count] for the Householder pass (ds_batch_fused_kernel<.., 2>, which
// rebuilds its tiles).  LDS per wave: the row (16 KB at n = 64), scales, EMI
// scratch, against the three 8 KB tiles of the Householder pass: four waves
// per SIMD's worth of LDS per CU instead of three.
__host__ __device__ inline size_t batch_si_lds(int n) {
  return (((size_t)(n * (n - 1) / 2) * 8 + 15) & ~(size_t)15) + 128 * 4 + EMI_FAST_LDS_BYTES;
}

template <int NM>
__global__ __launch_bounds__(64) void ds_batch_si_kernel(const float2* __restrict__ samples,
                                                         const long long* __restrict__ row_off,
                                                         const int32_t* __restrict__ n_samp,
                                                         const long long* __restrict__ ph_off,
                                                         const int32_t* __restrict__ pix, long long npix, int n,
                                                         int ref, const int32_t* __restrict__ pairs, int npairs,
                                                         float2* __restrict__ ph, float* __restrict__ quality,
                                                         float* __restrict__ tcoh, unsigned* __restrict__ count,
                                                         int* __restrict__ list) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int P = n * (n - 1) / 2;
  float2* cs = reinterpret_cast<float2*>(smem);                                    // the pixel's triu row
  float* sc = reinterpret_cast<float*>(smem + ((P * 8 + 15) & ~15));               // scales, 1/sqrt(diag)
  unsigned char* work = reinterpret_cast<unsigned char*>(sc + 128);                // EMI_FAST_LDS_BYTES
  float2* ub = reinterpret_cast<float2*>(work + 512);                              // linked phases after EMI
  for (long long t = blockIdx.x; t < npix; t += gridDim.x) {
    int lane = threadIdx.x;
    __asm__ volatile("" : "+v"(lane));  // per-pixel opaque lane (see emi.hip)
    const int id = pix[t];
    const CoAcc acc = co_mfma_accumulate<true>(samples + row_off[id], RowIdentity{}, n_samp[id], n, lane);
    __builtin_amdgcn_wave_barrier();
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
      if (lane == 0) list[atomicAdd(count, 1u)] = (int)t;
      __builtin_amdgcn_wave_barrier();
      continue;
    }
    const float2 pv = emi_link_phase(vr, vi, ref);
    __builtin_amdgcn_wave_barrier();
    ub[lane] = pv;  // linked phases for the temporal coherence
    if (lane < n) ph[ph_off[id] + lane] = pv;
    if (lane == 0) quality[id] = (float)rq;
    __builtin_amdgcn_wave_barrier();
    // pairs is the full triu network (mr_ds_batch's contract): row entry q is pair q
    const float tc = tcoh_point<true>([&](int q, int, int) { return cs[q]; }, [&](int i) { return ub[i]; }, pairs,
                                      npairs, lane);
    if (lane == 0) tcoh[id] = tc;
    __builtin_amdgcn_wave_barrier();
  }
}

// Coherence tiles of one pixel as the FP64 solve's source.
struct BatchTileSrc {
  CohTiles<true> ct;
  __device__ __forceinline__ float2 operator()(int a, int b) const { return ct(a, b); }
  __device__ __forceinline__ float2 pair(int, int a, int b) const { return ct(a, b); }
};

// The FP64 re-solve of the pixels res_list[0 .. *res_count): wave 0 rebuilds
// the pixel's coherence tiles, then the block runs emi_general_solve (slice in
// LDS) with the temporal coherence: the same values the plain EMI's re-solve
// computes from the stored coherence row.
__global__ __launch_bounds__(EMI_GEN_THREADS) void ds_batch_resolve_kernel(
    const float2* __restrict__ samples, const long long* __restrict__ row_off, const int32_t* __restrict__ n_samp,
    const long long* __restrict__ ph_off, const int32_t* __restrict__ pix, int n, int ref,
    const int32_t* __restrict__ pairs, float2* __restrict__ ph, float* __restrict__ quality,
    float* __restrict__ tcoh, const unsigned* __restrict__ res_count, const int* __restrict__ res_list) {
  __shared__ double red[EMI_GEN_THREADS / 64];
  __shared__ double sh_key[EMI_GEN_THREADS];
  __shared__ int sh_idx[EMI_GEN_THREADS];
  __shared__ double2 tri_de[64];
  __shared__ double tri_e[64];
  extern __shared__ __align__(16) unsigned char smem[];
  float2* tiles = reinterpret_cast<float2*>(smem);
  float* sc = reinterpret_cast<float*>(smem + (size_t)com_tile_floats(true) * 4);
  unsigned char* slice = reinterpret_cast<unsigned char*>(sc + 128);
  const GenShared sh{red, sh_key, sh_idx, tri_de, tri_e};
  const long long nit = *res_count;
  for (long long it = blockIdx.x; it < nit; it += gridDim.x) {
    const int id = pix[res_list[it]];
    if (threadIdx.x < 64) {
      int lane = threadIdx.x;
      __asm__ volatile("" : "+v"(lane));
      co_mfma_point<true>(samples + row_off[id], RowIdentity{}, n_samp[id], n, lane, tiles, sc);
    }
    __syncthreads();
    const BatchTileSrc src{CohTiles<true>{tiles, sc + 64}};
    emi_general_solve<true, float2>(src, n, ref, true, slice, sh, ph + ph_off[id], quality + id, pairs, tcoh + id);
  }
}

__host__ __device__ inline size_t batch_resolve_lds(int n) {
  return (size_t)com_tile_floats(true) * 4 + 128 * 4 + emi_gen_slice_bytes(n);
}

// Outputs of pixel id: phases ph[off + i], quality[id] (ragged batch).
struct BatchOut {
  float2* phs;
  float* qs;
  long long off;
  int pid;
  __device__ __forceinline__ void ph(int i, float2 v) const { phs[off + i] = v; }
  __device__ __forceinline__ void quality(float v) const { qs[pid] = v; }
  __device__ __forceinline__ int id() const { return pid; }
};

// n <= 32: pixel t's coherence into one LDS tile, EMI by the drop-in's body
// for this n reading the tile, temporal coherence from the tile and the linked
// phases (kept in LDS by the EMI body).
__host__ __device__ inline size_t batch_small_lds() {
  return ((ew_lds_bytes(32) + 15) & ~(size_t)15) + (size_t)com_tile_floats(false) * 4 + 128 * 4 + 32 * 8;
}

__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(4))) void ds_batch_fused_small_kernel(
    const float2* __restrict__ samples, const long long* __restrict__ row_off, const int32_t* __restrict__ n_samp,
    const long long* __restrict__ ph_off, const int32_t* __restrict__ pix, long long npix, int n, int ref,
    const int32_t* __restrict__ pairs, int npairs, float2* __restrict__ ph, float* __restrict__ quality,
    float* __restrict__ tcoh) {
  extern __shared__ __align__(16) unsigned char smem[];
  ew_lds_init<32>(smem);
  unsigned char* rest = smem + ((ew_lds_bytes(32) + 15) & ~(size_t)15);
  float2* tiles = reinterpret_cast<float2*>(rest);
  float* sc = reinterpret_cast<float*>(rest + (size_t)com_tile_floats(false) * 4);  // scales, 1/sqrt(diag)
  float2* phl = reinterpret_cast<float2*>(sc + 128);                                 // linked phases
  for (long long t = blockIdx.x; t < npix; t += gridDim.x) {
    int lane = threadIdx.x;
    __asm__ volatile("" : "+v"(lane));  // per-pixel opaque lane (see emi.hip)
    const int id = pix[t];
    co_mfma_point<false>(samples + row_off[id], RowIdentity{}, n_samp[id], n, lane, tiles, sc);
    const CohTiles<false> ct{tiles, sc + 64};
    ew_point<2, 16, 8, 4, false>(ct, n, ref, smem, BatchOut{ph, quality, ph_off[id], id}, 0.0, nullptr, nullptr, phl);
    const float tc = tcoh_point<true>([&](int, int a, int b) { return ct(a, b); }, [&](int i) { return phl[i]; },
                                      pairs, npairs, lane);
    if (lane == 0) tcoh[id] = tc;
    __syncthreads();  // the next pixel overwrites the tile
  }
}

// 64 < n <= 200: NB = ceil(n / 32) waves per workgroup, as the EMI body's two
// lanes per row need.  Per pixel: coherence row into the workgroup's slot
// (slots + blockIdx.x * npairs), EMI from the slot, DS temporal coherence from
// the slot on wave 0.
__host__ __device__ inline size_t batch_wide_lds(int e, int nb) {
  return ((ew_lds_bytes(2 * e) + 15) & ~(size_t)15) + (size_t)2 * 32 * nb * 4 + (size_t)2 * e * 8 +
         (size_t)nb * 32 * 32 * 8;
}

template <int E, int GS, int GH, int NB>
__global__ __launch_bounds__(64 * NB) __attribute__((amdgpu_waves_per_eu(E <= 64 ? 3 : 1))) void
ds_batch_fused_wide_kernel(const float2* __restrict__ samples, const long long* __restrict__ row_off,
                           const int32_t* __restrict__ n_samp, const long long* __restrict__ ph_off,
                           const int32_t* __restrict__ pix, long long npix, int n, int ref,
                           const int32_t* __restrict__ pairs, int npairs, float2* __restrict__ ph,
                           float* __restrict__ quality, float* __restrict__ tcoh, float2* __restrict__ slots) {
  extern __shared__ __align__(16) unsigned char smem[];
  ew_lds_init<2 * E>(smem);
  unsigned char* rest = smem + ((ew_lds_bytes(2 * E) + 15) & ~(size_t)15);
  float* sc = reinterpret_cast<float*>(rest);                 // 32 NB scales
  float* rs = sc + 32 * NB;                                   // 32 NB 1/sqrt(diag)
  float2* phl = reinterpret_cast<float2*>(rs + 32 * NB);      // 2E linked phases
  float2* stage = phl + 2 * E;                                // NB 32x32 tiles (other pair orders)
  const bool canon = co_pairs_canonical(pairs, npairs, n);
  float2* slot = slots + (size_t)blockIdx.x * npairs;
  for (long long t = blockIdx.x; t < npix; t += gridDim.x) {
    const int id = pix[t];
    cow_point<true, NB, NB, true>(samples + row_off[id], nullptr, n_samp[id], n, pairs, npairs, canon, sc, rs, stage,
                            slot);
    __syncthreads();  // the row is complete in the slot
    ew_point<2, E, GS, GH, false>(EwPairRow{slot, n}, n, ref, smem, BatchOut{ph, quality, ph_off[id], id}, 0.0,
                                  nullptr, nullptr, phl);
    if (threadIdx.x < 64) {
      int lane = threadIdx.x;
      __asm__ volatile("" : "+v"(lane));
      const float tc = tcoh_point<true>([&](int p, int, int) { return slot[p]; }, [&](int i) { return phl[i]; },
                                        pairs, npairs, lane);
      if (lane == 0) tcoh[id] = tc;
    }
    __syncthreads();  // the next pixel overwrites the slot
  }
}

// Grid of the 64 < n <= 200 fused launch: `blocks` if positive, else 2048
// workgroups (several rounds of the resident count at every n, for balance),
// never more than the pixels.
constexpr int BATCH_WIDE_DEFAULT_BLOCKS = 2048;
inline long long batch_wide_grid(long long n_pix, int blocks) {
  const long long nb = blocks > 0 ? blocks : BATCH_WIDE_DEFAULT_BLOCKS;
  return n_pix < nb ? n_pix : nb;
}

template <int E, int GS, int GH, int NB>
static int batch_wide_launch(const float2* smp, const long long* row_off, const int32_t* n_samples,
                             const long long* ph_off, const int32_t* pix, long long n_pix, int n, int ref,
                             const int32_t* pairs, int n_pairs, float2* ph, float* quality, float* t_coh, int nb,
                             float2* slots, hipStream_t st) {
  hipLaunchKernelGGL((ds_batch_fused_wide_kernel<E, GS, GH, NB>), dim3(nb), dim3(64 * NB), batch_wide_lds(E, NB), st,
                     smp, row_off, n_samples, ph_off, pix, n_pix, n, ref, pairs, n_pairs, ph, quality, t_coh, slots);
  return launch_status("ds_batch fused (64 < n <= 200)");
}

__global__ void ds_batch_scatter_kernel(const float2* __restrict__ ph_t, const float* __restrict__ q_t,
                                        const float* __restrict__ t_t, const int32_t* __restrict__ pix, long long c0,
                                        long long cn, int n, const long long* __restrict__ ph_off,
                                        float2* __restrict__ ph, float* __restrict__ quality,
                                        float* __restrict__ tcoh) {
  const long long tot = cn * n;
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < tot; t += (long long)gridDim.x * blockDim.x) {
    const long long pt = t / n;
    const int j = (int)(t - pt * n);
    const int id = pix[c0 + pt];
    ph[ph_off[id] + j] = ph_t[t];
    if (j == 0) {
      quality[id] = q_t[pt];
      tcoh[id] = t_t[pt];
    }
  }
}

static size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

}  // namespace mr

extern "C" {

size_t mr_ds_batch_fused_workspace_bytes(long long n_pix, int n, int blocks) {
  if (n_pix <= 0) return 0;
  if (n > 32 && n <= 64) return 256 + 2 * ((((size_t)n_pix * 4 + 255) & ~(size_t)255));
  if (n > 64 && n <= mr::EMI_WIDE_MAX_N)
    return (size_t)mr::batch_wide_grid(n_pix, blocks) * ((size_t)n * (n - 1) / 2) * sizeof(float2);
  return 0;  // n <= 32: on chip; n > 200: no fused launch
}

size_t mr_ds_batch_workspace_bytes(long long chunk, int n) {
  if (n < 2 || chunk <= 0) return 0;
  const long long P = (long long)n * (n - 1) / 2;
  return mr::align256((size_t)chunk * P * 8) + mr::align256((size_t)chunk * n * 8) +
         2 * mr::align256((size_t)chunk * 4) + mr::align256(mr_emi_workspace_bytes(chunk, n));
}

int mr_ds_batch(const float* samples, const long long* row_off, const int32_t* n_samples, const long long* ph_off,
                const int32_t* pix, long long n_pix, int n, int ref, const int32_t* pairs, int n_pairs, float* ph,
                float* quality, float* t_coh, int blocks, long long chunk, int coh_waves, void* workspace,
                size_t workspace_bytes, mr_stream_t stream) {
  MR_CHECK_ARG(n >= 2 && n <= 1024, "ds_batch: n=%d outside [2, 1024]", n);
  MR_CHECK_ARG(ref >= 0 && ref < n, "ds_batch: ref=%d outside [0,%d)", ref, n);
  MR_CHECK_ARG(n_pix >= 0, "ds_batch: negative pixel count");
  MR_CHECK_ARG(n_pairs == n * (n - 1) / 2, "ds_batch: needs the full pair network (%d pairs for n=%d, got %d)",
               n * (n - 1) / 2, n, n_pairs);
  if (n_pix == 0) return MR_OK;
  MR_CHECK_ARG(samples && row_off && n_samples && ph_off && pix && pairs && ph && quality && t_coh,
               "ds_batch: null pointer");
  hipStream_t st = mr::as_stream(stream);
  const float2* smp = reinterpret_cast<const float2*>(samples);
  float2* phc = reinterpret_cast<float2*>(ph);
  if (n <= 32 && blocks >= 0) {
    const int nb = (int)std::min<long long>(n_pix, blocks > 0 ? blocks : mr::BATCH_DEFAULT_BLOCKS);
    hipLaunchKernelGGL(mr::ds_batch_fused_small_kernel, dim3(nb), dim3(64), mr::batch_small_lds(), st, smp, row_off,
                       n_samples, ph_off, pix, n_pix, n, ref, pairs, n_pairs, phc, quality, t_coh);
    return mr::launch_status("ds_batch fused (n <= 32)");
  }
  if (n > 64 && n <= mr::EMI_WIDE_MAX_N && blocks >= 0) {
    const size_t need = mr_ds_batch_fused_workspace_bytes(n_pix, n, blocks);
    MR_CHECK_ARG(workspace && workspace_bytes >= need, "ds_batch: fused launch workspace %zu < %zu bytes",
                 workspace_bytes, need);
    const int nb = (int)mr::batch_wide_grid(n_pix, blocks);
    float2* slots = reinterpret_cast<float2*>(workspace);
#define MR_BW(E, GS, GH, NB)                                                                                       \
  mr::batch_wide_launch<E, GS, GH, NB>(smp, row_off, n_samples, ph_off, pix, n_pix, n, ref, pairs, n_pairs, phc,  \
                                       quality, t_coh, nb, slots, st)
    if (n <= 96) return MR_BW(48, 4, 2, 3);
    if (n <= 128) return MR_BW(64, 4, 2, 4);
    if (n <= 160) return MR_BW(80, 8, 4, 5);
    if (n <= 192) return MR_BW(100, 4, 2, 6);
    return MR_BW(100, 4, 2, 7);
#undef MR_BW
  }
  if (n > 32 && n <= 64 && blocks >= 0) {
    const size_t need = mr_ds_batch_fused_workspace_bytes(n_pix, n, blocks);
    MR_CHECK_ARG(workspace && workspace_bytes >= need, "ds_batch: fused launch workspace %zu < %zu bytes",
                 workspace_bytes, need);
    unsigned* count = reinterpret_cast<unsigned*>(workspace);
    unsigned* res_count = count + 1;
    const size_t span = ((size_t)n_pix * 4 + 255) & ~(size_t)255;
    int* list = reinterpret_cast<int*>(reinterpret_cast<unsigned char*>(workspace) + 256);
    int* res_list = reinterpret_cast<int*>(reinterpret_cast<unsigned char*>(workspace) + 256 + span);
    hipLaunchKernelGGL(mr::batch_zero_count_kernel, dim3(1), dim3(1), 0, st, count);
    if (int rc = mr::launch_status("ds_batch count init")) return rc;
    const int nb = (int)std::min<long long>(n_pix, blocks > 0 ? blocks : mr::BATCH_DEFAULT_BLOCKS);
    // the shift-invert pass keeps 60 columns per lane up to n = 60 (as the plain EMI)
    if (n <= 60)
      hipLaunchKernelGGL((mr::ds_batch_si_kernel<60>), dim3(nb), dim3(64), mr::batch_si_lds(n), st, smp, row_off,
                         n_samples, ph_off, pix, n_pix, n, ref, pairs, n_pairs, phc, quality, t_coh, count, list);
    else
      hipLaunchKernelGGL((mr::ds_batch_si_kernel<64>), dim3(nb), dim3(64), mr::batch_si_lds(n), st, smp, row_off,
                         n_samples, ph_off, pix, n_pix, n, ref, pairs, n_pairs, phc, quality, t_coh, count, list);
    if (int rc = mr::launch_status("ds_batch shift-invert")) return rc;
    const int nb2 = (int)std::min<long long>(n_pix, mr::BATCH_FALLBACK_BLOCKS);
    hipLaunchKernelGGL((mr::ds_batch_fused_kernel<64, true, 2>), dim3(nb2), dim3(64), mr::batch_fused_lds(true), st,
                       smp, row_off, n_samples, ph_off, pix, n_pix, n, ref, pairs, n_pairs, phc, quality, t_coh, count,
                       list, res_count, res_list);
    if (int rc = mr::launch_status("ds_batch householder")) return rc;
    const int nb3 = (int)std::min<long long>(n_pix, mr::EMI_GEN_BLOCKS);
    hipLaunchKernelGGL(mr::ds_batch_resolve_kernel, dim3(nb3), dim3(mr::EMI_GEN_THREADS), mr::batch_resolve_lds(n),
                       st, smp, row_off, n_samples, ph_off, pix, n, ref, pairs, phc, quality, t_coh, res_count,
                       res_list);
    return mr::launch_status("ds_batch FP64 re-solve");
  }
  MR_CHECK_ARG(chunk > 0, "ds_batch: chunk must be positive for n > 200 or the per-stage mode");
  const size_t need = mr_ds_batch_workspace_bytes(chunk, n);
  MR_CHECK_ARG(workspace && workspace_bytes >= need, "ds_batch: workspace %zu < %zu bytes", workspace_bytes, need);
  const long long P = (long long)n * (n - 1) / 2;
  unsigned char* w = reinterpret_cast<unsigned char*>(workspace);
  float2* coh = reinterpret_cast<float2*>(w);
  w += mr::align256((size_t)chunk * P * 8);
  float2* ph_t = reinterpret_cast<float2*>(w);
  w += mr::align256((size_t)chunk * n * 8);
  float* q_t = reinterpret_cast<float*>(w);
  w += mr::align256((size_t)chunk * 4);
  float* t_t = reinterpret_cast<float*>(w);
  w += mr::align256((size_t)chunk * 4);
  void* emi_ws = w;
  const size_t emi_ws_bytes = mr::align256(mr_emi_workspace_bytes(chunk, n));
  for (long long c0 = 0; c0 < n_pix; c0 += chunk) {
    const long long cn = std::min(chunk, n_pix - c0);
    // the coherence kernel the drop-in uses for this n (MFMA up to 224, VALU above)
    const bool wide = n > 64 && n <= 224;
    const int rc0 =
        n <= 64 ? mr::co_mfma_samples_launch(smp, row_off, n_samples, pix, c0, cn, n, pairs, n_pairs, coh, st)
        : wide  ? mr::co_mfma_wide_samples_launch(smp, row_off, n_samples, pix, c0, cn, n, pairs, n_pairs, coh, st,
                                                  coh_waves)
                : mr::co_samples_launch(smp, row_off, n_samples, pix, c0, cn, n, pairs, n_pairs, coh, st);
    if (rc0) return rc0;
    // EMI and the temporal coherence of the chunk: one kernel for 32 < n <= 64 (equal, bit for bit, to
    // mr_emi followed by mr_ds_temp_coh)
    if (int rc = mr_emi_ds_temp_coh(reinterpret_cast<const float*>(coh), cn, n, ref, pairs, n_pairs,
                                    reinterpret_cast<float*>(ph_t), q_t, t_t, emi_ws, emi_ws_bytes, stream))
      return rc;
    const int nb = (int)std::min<long long>(mr::div_up((unsigned long long)cn * n, 256), 1 << 16);
    hipLaunchKernelGGL(mr::ds_batch_scatter_kernel, dim3(nb), dim3(256), 0, st, ph_t, q_t, t_t, pix, c0, cn, n,
                       ph_off, phc, quality, t_coh);
    if (int rc = mr::launch_status("ds_batch scatter")) return rc;
  }
  return MR_OK;
}

}  // extern "C"
