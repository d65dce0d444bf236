// Coherence estimation on MFMA (gfx950), n <= 64: raster (moraine/co.py:117-137)
// and point-cloud (moraine/co.py:230-266) forms.  One wave per point compacts
// the in-bounds SHPs of its window, then runs the shared per-point MFMA
// accumulation (co_mfma.hpp) and writes coh (and cov) from the tiles.
#include "co_mfma.hpp"

namespace mr {

constexpr int COM_WAVES = 2;  // points per block (one per wave); LDS per wave: com_wave_bytes (~9 KiB)

// LDS per wave: the window list (while accumulating), then one 32x32 complex
// tile at a time for the output pass, plus scales and 1/sqrt(diag).
__host__ __device__ inline size_t com_wave_bytes(int win, bool) {
  const size_t list = ((size_t)win * 4 + 15) & ~(size_t)15;
  const size_t tile = 32 * 32 * 8;
  return (list > tile ? list : tile) + 2 * 64 * 4;
}

template <bool RASTER, bool TWO>
__global__ __launch_bounds__(64 * COM_WAVES) void co_mfma_kernel(
    const float2* __restrict__ rslc, int nl, int w, int n, const int32_t* __restrict__ idx,
    const uint8_t* __restrict__ shp, long long npts, int aw, int rw, long long pix0,
    const int32_t* __restrict__ pairs, int npairs, float2* __restrict__ cov, float2* __restrict__ coh) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // uniform: per-point pointers stay scalar
  const int lane = threadIdx.x & 63;
  const int win = aw * rw;
  const int az = (aw - 1) / 2, r = (rw - 1) / 2;
  unsigned char* base = smem + (size_t)wv * com_wave_bytes(win, TWO);
  int* list = reinterpret_cast<int*>(base);
  float2* tiles = reinterpret_cast<float2*>(base);  // reused after the main loop
  float* sc = reinterpret_cast<float*>(base + (com_wave_bytes(win, TWO) - 2 * 64 * 4));
  const float* rs = sc + 64;

  // XCD-aware order: workgroups are dispatched round-robin over the 8 XCDs, so
  // block b runs on XCD b % 8.  XCD x takes the x-th of eight contiguous ranges
  // of points and its workgroups sweep that range in order, round after
  // round, so overlapping SHP windows share that XCD's L2 (the grid is a
  // multiple of 8).
  const unsigned nbx = gridDim.x / 8, xcd = blockIdx.x % 8, jb = blockIdx.x / 8;
  const long long per = (npts + 7) / 8;
  const long long beg = (long long)xcd * per, end = beg + per < npts ? beg + per : npts;
  const bool canon = !RASTER && co_pairs_canonical(pairs, npairs, n);
  const CoStrips st = RASTER ? CoStrips{} : co_strips(idx, beg, end, w);
  for (long long sl = (long long)jb * COM_WAVES + wv; sl < st.slots(end - beg); sl += (long long)nbx * COM_WAVES) {
    const long long u = st.point(sl, end - beg);
    if (u < 0) continue;  // a slot of the strip grid past the run
    const long long pt = beg + u;
    int ia, ir;
    const uint8_t* ms;
    if (RASTER) {
      const long long px = pix0 + pt;
      ia = (int)(px / w);
      ir = (int)(px - (long long)ia * w);
      ms = shp + px * win;
    } else {
      ia = idx[2 * pt];
      ir = idx[2 * pt + 1];
      ms = shp + pt * win;
    }
    // ---- in-bounds SHPs, compacted in window order ---------------------------
    const int K = co_compact_shp(ms, win, rw, ia, ir, az, r, nl, w, lane, list);
    __builtin_amdgcn_wave_barrier();
    const CoAcc acc = co_mfma_accumulate<TWO>(rslc, RowList{list}, K, n, lane);
    __builtin_amdgcn_wave_barrier();  // the tile below overwrites the list
    co_store_scales<TWO>(acc, n, lane, sc);
    if (canon) {
      // opaque per-point lane: keeps the 16-48 per-lane pair offsets from being
      // hoisted out of the point loop into VGPRs
      int ln = lane;
      __asm__ volatile("" : "+v"(ln));
      float2* out = coh + (size_t)pt * npairs;
      co_write_tile_direct(acc.r00, acc.i00, 0, 0, n, ln, rs, out);
      if (TWO) {
        co_write_tile_direct(acc.r01, acc.i01, 0, 1, n, ln, rs, out);
        co_write_tile_direct(acc.r11, acc.i11, 1, 1, n, ln, rs, out);
      }
      __builtin_amdgcn_wave_barrier();
      continue;
    }
    // output pass, one upper tile at a time: (0,0), then (0,1) and (1,1)
#pragma unroll
    for (int t = 0; t < (TWO ? 3 : 1); ++t) {
      co_store_tile(tiles, t == 0 ? acc.r00 : t == 1 ? acc.r01 : acc.r11, t == 0 ? acc.i00 : t == 1 ? acc.i01 : acc.i11,
                    lane);
      __builtin_amdgcn_wave_barrier();
      auto in_tile = [&](int lo, int hi) { return !TWO || ((lo >= 32) ? 2 : (hi >= 32 ? 1 : 0)) == t; };
      if (RASTER) {
        // the tile's 32x32 block of the N x N output and, off the diagonal, its
        // mirrored block (conjugated); a half-wave writes 32 consecutive entries
        const float fk = (float)K;
        const int ta = t == 2 ? 1 : 0, tb = t == 0 ? 0 : 1;
        for (int mir = 0; mir < (ta != tb ? 2 : 1); ++mir) {
          const int r0 = 32 * (mir ? tb : ta), c0 = 32 * (mir ? ta : tb);
          for (int it = 0; it < 16; ++it) {
            const int m = r0 + 2 * it + (lane >> 5), j = c0 + (lane & 31);
            if (m >= n || j >= n) continue;
            const int lo = min(m, j), hi = max(m, j);
            float2 v = tiles[(lo & 31) * 32 + (hi & 31)];
            if (m > j) v.y = -v.y;
            const float f = rs[m] * rs[j];
            const size_t o = (size_t)pt * n * n + (size_t)m * n + j;
            coh[o] = make_float2(v.x * f, v.y * f);
            if (cov) {
              const float us = 1.f / (sc[m] * sc[j]);  // exact: powers of two
              cov[o] = make_float2(v.x * us / fk, v.y * us / fk);
            }
          }
        }
      } else {
        for (int p = lane; p < npairs; p += 64) {
          const int a = pairs[2 * p], b = pairs[2 * p + 1];
          const int lo = min(a, b), hi = max(a, b);
          if (!in_tile(lo, hi)) continue;
          float2 v = tiles[(lo & 31) * 32 + (hi & 31)];
          if (a > b) v.y = -v.y;
          const float f = rs[a] * rs[b];
          coh[(size_t)pt * npairs + p] = make_float2(v.x * f, v.y * f);
        }
      }
      __builtin_amdgcn_wave_barrier();
    }
  }
}

// ---------------------------------------------------------------------------
// 64 < n <= 224 (point-cloud form): one 4- or 8-wave workgroup per point.  The
// upper tiles of the NB x NB grid of 32x32 blocks (NB = 4: 10 tiles) are split
// over the waves (tile t = wave + 4 s, at most 3 per wave); each wave accumulates its tiles with the
// same FP16 hi+lo MFMA arithmetic and per-image power-of-two scaling as the
// one-wave kernel, the diagonal tiles give 1/sqrt(S'_mm) through LDS, and each
// wave writes its tiles (straight from the accumulators for the canonical pair
// list, through an LDS tile otherwise).  NB = 3 (6 tiles) serves 65 <= n <= 96.
// With the accumulators in VGPRs (two waves per SIMD at NB = 4, three at NB = 3)
// it is 2.1x / 2.8x faster than the VALU kernel on the same SHP sets at n = 120
// (21x41 windows) / n = 90 (15x21).  PF prefetches one chunk of operands ahead
// (off: the registers cost more than it hides).
// ---------------------------------------------------------------------------
// SAMP: point pt is pixel pix[pix0 + pt], its K = n_samp[id] rows are the
// sample matrix at rslc + row_off[id] (batched form), no window.
// XCD: workgroups are dispatched round-robin over the 8 XCDs (block b runs on
// XCD b % 8); XCD x takes the x-th of eight contiguous ranges of points and
// sweeps it in order, so overlapping windows share that XCD's L2 (the grid is
// a multiple of 8).
template <bool SAMP, int NB, int W, bool XCD = false>
__global__ __launch_bounds__(64 * W) void co_mfma_wide_kernel(
    const float2* __restrict__ rslc, int nl, int w, int n, const int32_t* __restrict__ idx,
    const uint8_t* __restrict__ shp, long long npts, int aw, int rw, const int32_t* __restrict__ pairs, int npairs,
    float2* __restrict__ coh, const long long* __restrict__ row_off, const int32_t* __restrict__ n_samp,
    const int32_t* __restrict__ pix, long long pix0) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int win = aw * rw;
  const int az = (aw - 1) / 2, r = (rw - 1) / 2;
  int* list = reinterpret_cast<int*>(smem);
  float* sc = reinterpret_cast<float*>(smem + (((size_t)win * 4 + 15) & ~(size_t)15));  // 32 NB scales
  float* rs = sc + 32 * NB;                                                             // 32 NB 1/sqrt(diag)
  float2* stage = reinterpret_cast<float2*>(rs + 32 * NB);                              // W x 32x32
  int* sK = reinterpret_cast<int*>(stage + W * 1024);
  const int wv = threadIdx.x >> 6;
  const int lane0 = threadIdx.x & 63;
  const bool canon = co_pairs_canonical(pairs, npairs, n);
  // XCD order: XCD x sweeps the x-th of eight contiguous ranges of points
  const long long per = XCD ? (npts + 7) / 8 : npts;
  const long long beg = XCD ? (long long)(blockIdx.x % 8) * per : 0;
  const long long end = beg + per < npts ? beg + per : npts;
  const unsigned jb = XCD ? blockIdx.x / 8 : blockIdx.x, nbx = XCD ? gridDim.x / 8 : gridDim.x;
  for (long long pt = beg + jb; pt < end; pt += nbx) {
    int lane = lane0;
    __asm__ volatile("" : "+v"(lane));
    // ---- in-bounds SHPs in window order (wave 0) -----------------------------
    const float2* base = rslc;
    if (SAMP) {
      const int id = pix[pix0 + pt];
      base = rslc + row_off[id];
      if (threadIdx.x == 0) sK[0] = n_samp[id];
    } else if (wv == 0) {
      const int ia = idx[2 * pt], ir = idx[2 * pt + 1];
      const uint8_t* ms = shp + pt * win;
      const int K = co_compact_shp(ms, win, rw, ia, ir, az, r, nl, w, lane, list);
      if (lane == 0) sK[0] = K;
    }
    __syncthreads();
    cow_point<SAMP, NB, W>(base, list, sK[0], n, pairs, npairs, canon, sc, rs, stage, coh + (size_t)pt * npairs);
    __syncthreads();
  }
}

// Tile grid and waves per workgroup for n: one wave per tile row of the 3x3 and
// 4x4 grids (two tiles per wave), 5x5 to 7x7 on 8 waves (at most four tiles per
// wave); the point-cloud form in XCD-contiguous point order for the 3x3 and
// 4x4 grids.  Measured on 15x21 (n = 90, 150) and 21x41 (n = 120) windows with
// every layout bit-identical: n = 90 4 waves 6.14 ms -> XCD order 5.72 ->
// 3 waves 4.97 ms per 131 072 points; n = 120 13.39 -> 12.09 ms per 65 536
// points with the XCD order; n = 150 8.53 ms on 8 waves, 8.77 with the XCD
// order, 11.66 on 5 waves.
#define MR_COW_DISPATCH(SAMP, LDS_WIN, ...)                                                                        \
  do {                                                                                                          \
    if (n <= 96)                                                                                                \
      hipLaunchKernelGGL((co_mfma_wide_kernel<SAMP, 3, 3, !SAMP>), dim3(nb), dim3(64 * 3),                      \
                         cow_lds_bytes(LDS_WIN, 3, 3), st, __VA_ARGS__);                                        \
    else if (n <= 128)                                                                                          \
      hipLaunchKernelGGL((co_mfma_wide_kernel<SAMP, 4, 4, !SAMP>), dim3(nb), dim3(64 * 4),                      \
                         cow_lds_bytes(LDS_WIN, 4, 4), st, __VA_ARGS__);                                        \
    else if (n <= 160)                                                                                          \
      hipLaunchKernelGGL((co_mfma_wide_kernel<SAMP, 5, 8>), dim3(nb), dim3(64 * 8),                             \
                         cow_lds_bytes(LDS_WIN, 5, 8), st, __VA_ARGS__);                                        \
    else if (n <= 192)                                                                                          \
      hipLaunchKernelGGL((co_mfma_wide_kernel<SAMP, 6, 8>), dim3(nb), dim3(64 * 8),                             \
                         cow_lds_bytes(LDS_WIN, 6, 8), st, __VA_ARGS__);                                        \
    else                                                                                                        \
      hipLaunchKernelGGL((co_mfma_wide_kernel<SAMP, 7, 8>), dim3(nb), dim3(64 * 8),                             \
                         cow_lds_bytes(LDS_WIN, 7, 8), st, __VA_ARGS__);                                        \
  } while (0)


int co_mfma_wide_launch(const float2* rslc, int nl, int w, int n, const int32_t* idx, const uint8_t* shp,
                        long long npts, int aw, int rw, const int32_t* pairs, int npairs, float2* coh, hipStream_t st) {
  // a multiple of 8 workgroups keeps the XCD renumbering a bijection
  const int nb = (int)((std::min<long long>(npts, 1 << 20) + 7) / 8 * 8);
  MR_COW_DISPATCH(false, aw * rw, rslc, nl, w, n, idx, shp, npts, aw, rw, pairs, npairs, coh, nullptr, nullptr,
                  nullptr, 0ll);
  return launch_status("co_mfma_wide");
}

#undef MR_COW_DISPATCH

// Sample-matrix form with an explicit tile layout: the NB x NB tile grid's
// upper triangle spread over W waves (tile slots per wave ceil(NT / W)).
template <int NB, int W>
static void cow_samples(const float2* samples, const long long* row_off, const int32_t* n_samp, const int32_t* pix,
                        long long pix0, long long cn, int n, const int32_t* pairs, int npairs, float2* coh,
                        hipStream_t st) {
  const int nb = (int)std::min<long long>(cn, 1 << 20);
  hipLaunchKernelGGL((co_mfma_wide_kernel<true, NB, W>), dim3(nb), dim3(64 * W), cow_lds_bytes(1, NB, W), st,
                     samples, 0, 0, n, nullptr, nullptr, cn, 1, 1, pairs, npairs, coh, row_off, n_samp, pix, pix0);
}

template <int NB>
static int cow_samples_waves(int waves, const float2* samples, const long long* row_off, const int32_t* n_samp,
                             const int32_t* pix, long long pix0, long long cn, int n, const int32_t* pairs,
                             int npairs, float2* coh, hipStream_t st) {
  constexpr int kDefault = NB <= 4 ? NB : 8;  // the drop-in's layout for this grid
  if (waves == 0) waves = kDefault;
  if constexpr (NB == 3) {
    if (waves == 2 || waves == 3) {
      if (waves == 2)
        cow_samples<NB, 2>(samples, row_off, n_samp, pix, pix0, cn, n, pairs, npairs, coh, st);
      else
        cow_samples<NB, 3>(samples, row_off, n_samp, pix, pix0, cn, n, pairs, npairs, coh, st);
      return MR_OK;
    }
  }
  if (waves == 4) {
    cow_samples<NB, 4>(samples, row_off, n_samp, pix, pix0, cn, n, pairs, npairs, coh, st);
  } else if (waves == 8) {
    cow_samples<NB, 8>(samples, row_off, n_samp, pix, pix0, cn, n, pairs, npairs, coh, st);
  } else {
    set_error("ds_batch: %d waves per coherence tile grid unsupported for n=%d (0, %s4 or 8)", waves, n,
              NB == 3 ? "2, 3, " : "");
    return MR_ERR_UNSUPPORTED;
  }
  return MR_OK;
}

// waves: 0 = the drop-in's layout (3 waves for n <= 96, 4 for n <= 128, 8 above); 2 or 3 (n <= 96), 4 or 8.
int co_mfma_wide_samples_launch(const float2* samples, const long long* row_off, const int32_t* n_samp,
                                const int32_t* pix, long long pix0, long long cn, int n, const int32_t* pairs,
                                int npairs, float2* coh, hipStream_t st, int waves) {
  int rc;
  if (n <= 96)
    rc = cow_samples_waves<3>(waves, samples, row_off, n_samp, pix, pix0, cn, n, pairs, npairs, coh, st);
  else if (n <= 128)
    rc = cow_samples_waves<4>(waves, samples, row_off, n_samp, pix, pix0, cn, n, pairs, npairs, coh, st);
  else if (n <= 160)
    rc = cow_samples_waves<5>(waves, samples, row_off, n_samp, pix, pix0, cn, n, pairs, npairs, coh, st);
  else if (n <= 192)
    rc = cow_samples_waves<6>(waves, samples, row_off, n_samp, pix, pix0, cn, n, pairs, npairs, coh, st);
  else
    rc = cow_samples_waves<7>(waves, samples, row_off, n_samp, pix, pix0, cn, n, pairs, npairs, coh, st);
  if (rc) return rc;
  return launch_status("ds_batch coherence (MFMA tile grid)");
}

// Sample-matrix form (batched path, per-stage mode): point pt is pixel
// pix[pix0 + pt] with K = n_samp[id] rows at samples + row_off[id]; writes the
// triu coherence row of the point, as the point-cloud form does.
template <bool TWO>
__global__ __launch_bounds__(64 * COM_WAVES) void co_mfma_samples_kernel(
    const float2* __restrict__ samples, const long long* __restrict__ row_off, const int32_t* __restrict__ n_samp,
    const int32_t* __restrict__ pix, long long pix0, long long npts, int n, const int32_t* __restrict__ pairs,
    int npairs, float2* __restrict__ coh) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int wv = threadIdx.x >> 6;
  const int lane = threadIdx.x & 63;
  const size_t wb = (size_t)com_tile_floats(TWO) * 4 + 2 * 64 * 4;
  float2* tiles = reinterpret_cast<float2*>(smem + wv * wb);
  float* sc = reinterpret_cast<float*>(smem + wv * wb + (size_t)com_tile_floats(TWO) * 4);
  for (long long pt = (long long)blockIdx.x * COM_WAVES + wv; pt < npts; pt += (long long)gridDim.x * COM_WAVES) {
    const int id = pix[pix0 + pt];
    co_mfma_point<TWO>(samples + row_off[id], RowIdentity{}, n_samp[id], n, lane, tiles, sc);
    const CohTiles<TWO> ct{tiles, sc + 64};
    for (int p = lane; p < npairs; p += 64) coh[(size_t)pt * npairs + p] = ct(pairs[2 * p], pairs[2 * p + 1]);
    __builtin_amdgcn_wave_barrier();
  }
}

int co_mfma_samples_launch(const float2* samples, const long long* row_off, const int32_t* n_samp,
                           const int32_t* pix, long long pix0, long long cn, int n, const int32_t* pairs,
                           int npairs, float2* coh, hipStream_t st) {
  const bool two = n > 32;
  const size_t lds = ((size_t)com_tile_floats(two) * 4 + 2 * 64 * 4) * COM_WAVES;
  const long long nb = std::min<long long>((cn + COM_WAVES - 1) / COM_WAVES, 1 << 20);
  if (two)
    hipLaunchKernelGGL((co_mfma_samples_kernel<true>), dim3((unsigned)nb), dim3(64 * COM_WAVES), lds, st, samples,
                       row_off, n_samp, pix, pix0, cn, n, pairs, npairs, coh);
  else
    hipLaunchKernelGGL((co_mfma_samples_kernel<false>), dim3((unsigned)nb), dim3(64 * COM_WAVES), lds, st, samples,
                       row_off, n_samp, pix, pix0, cn, n, pairs, npairs, coh);
  return launch_status("ds_batch coherence (MFMA)");
}

int co_mfma_launch(bool raster, const float2* rslc, int nl, int w, int n, const int32_t* idx, const uint8_t* shp,
                   long long npts, int aw, int rw, long long pix0, const int32_t* pairs, int npairs, float2* cov,
                   float2* coh, hipStream_t st) {
  const bool two = n > 32;
  const size_t lds = com_wave_bytes(aw * rw, two) * COM_WAVES;
  // a multiple of 8 blocks keeps the XCD renumbering a bijection
  // one wave per slot: the point-cloud form's strip order has up to
  // co_max_slots slots per XCD range (the empty ones exit at once)
  const long long per = (npts + 7) / 8;
  const long long waves = raster || per < 2 * (long long)w ? npts : 8 * co_max_slots(per, w);
  const long long nb = (std::min<long long>((waves + COM_WAVES - 1) / COM_WAVES, 1 << 20) + 7) / 8 * 8;
  dim3 g((unsigned)nb), b(64 * COM_WAVES);
#define MR_COM(R, T)                                                                                              \
  hipLaunchKernelGGL((co_mfma_kernel<R, T>), g, b, lds, st, rslc, nl, w, n, idx, shp, npts, aw, rw, pix0, pairs, \
                     npairs, cov, coh)
  if (raster) {
    if (two)
      MR_COM(true, true);
    else
      MR_COM(true, false);
  } else {
    if (two)
      MR_COM(false, true);
    else
      MR_COM(false, false);
  }
#undef MR_COM
  return launch_status("co_mfma");
}

}  // namespace mr
