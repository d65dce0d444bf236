// Coherence estimation on MFMA (gfx950), n <= 64: raster (moraine/co.py:117-137)
// and point-cloud (moraine/co.py:230-266) forms.  One wave per point compacts
// the in-bounds SHPs of its window, then runs the shared per-point MFMA
// accumulation (co_mfma.hpp) and writes coh (and cov) from the tiles.
#include "co_mfma.hpp"

namespace mr {

constexpr int COM_WAVES = 2;  // points per block (one per wave); 2 x ~25 KiB LDS per block

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
  const int wv = threadIdx.x >> 6;
  const int lane = threadIdx.x & 63;
  const int win = aw * rw;
  const int az = (aw - 1) / 2, r = (rw - 1) / 2;
  unsigned char* base = smem + (size_t)wv * com_wave_bytes(win, TWO);
  int* list = reinterpret_cast<int*>(base);
  float2* tiles = reinterpret_cast<float2*>(base);  // reused after the main loop
  float* sc = reinterpret_cast<float*>(base + (com_wave_bytes(win, TWO) - 2 * 64 * 4));
  const float* rs = sc + 64;

  // XCD-aware order: workgroups are dispatched round-robin over the 8 XCDs, so
  // block b runs on XCD b % 8.  Renumbering gives every XCD a contiguous range
  // of points, whose overlapping SHP windows then share that XCD's L2.
  const unsigned nbx = (gridDim.x + 7) / 8;
  const unsigned bx = (blockIdx.x % 8) * nbx + blockIdx.x / 8;
  if (bx >= gridDim.x) return;  // only when gridDim.x is not a multiple of 8
  for (long long pt = (long long)bx * COM_WAVES + wv; pt < npts; pt += (long long)gridDim.x * COM_WAVES) {
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
    int K = 0;
    for (int t0 = 0; t0 < win; t0 += 64) {
      const int t = t0 + lane;
      bool v = false;
      int lin = 0;
      if (t < win) {
        const int k = t / rw, l = t - (t / rw) * rw;
        const int sa = ia - az + k, sr = ir - r + l;
        v = ms[t] && sa >= 0 && sa < nl && sr >= 0 && sr < w;
        lin = sa * w + sr;
      }
      const unsigned long long b = __ballot(v);
      if (v) list[K + __popcll(b & ((1ull << lane) - 1ull))] = lin;
      K += __popcll(b);
    }
    __builtin_amdgcn_wave_barrier();
    const CoAcc acc = co_mfma_accumulate<TWO>(rslc, RowList{list}, K, n, lane);
    __builtin_amdgcn_wave_barrier();  // the tile below overwrites the list
    co_store_scales<TWO>(acc, n, lane, sc);
    // output pass, one upper tile at a time: (0,0), then (0,1) and (1,1)
#pragma unroll
    for (int t = 0; t < (TWO ? 3 : 1); ++t) {
      co_store_tile(tiles, t == 0 ? acc.r00 : t == 1 ? acc.r01 : acc.r11, t == 0 ? acc.i00 : t == 1 ? acc.i01 : acc.i11,
                    lane);
      __builtin_amdgcn_wave_barrier();
      auto in_tile = [&](int lo, int hi) { return !TWO || ((lo >= 32) ? 2 : (hi >= 32 ? 1 : 0)) == t; };
      if (RASTER) {
        const float fk = (float)K;
        for (int e = lane; e < n * n; e += 64) {
          const int m = e / n, j = e - (e / n) * n;
          const int lo = min(m, j), hi = max(m, j);
          if (!in_tile(lo, hi)) continue;
          float2 v = tiles[(lo & 31) * 32 + (hi & 31)];
          if (m > j) v.y = -v.y;
          const float f = rs[m] * rs[j];
          const size_t o = (size_t)pt * n * n + e;
          coh[o] = make_float2(v.x * f, v.y * f);
          if (cov) {
            const float us = 1.f / (sc[m] * sc[j]);  // exact: powers of two
            cov[o] = make_float2(v.x * us / fk, v.y * us / fk);
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
  const long long nb = (std::min<long long>((npts + COM_WAVES - 1) / COM_WAVES, 1 << 20) + 7) / 8 * 8;
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
