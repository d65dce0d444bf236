// Coherence estimation for gfx950 (general VALU path).
//
// Replaces moraine.emperical_co (moraine/co.py:117-137, kernel :61-114) and
// moraine.emperical_co_pc (moraine/co.py:230-266, kernel :141-190, numba
// :193-227).  One 256-thread block per point: the in-bounds SHPs of the
// window are compacted in window order (the reference's k-then-l loop order),
// their stacks are staged in LDS in chunks, per-image power sums are formed
// once per point and every thread accumulates a batch of image pairs in FP32
// registers.  uncompress_coh (moraine/co.py:281-299) lives here too.
#include "common.hpp"

namespace mr {

int co_mfma_launch(bool raster, const float2* rslc, int nl, int w, int n, const int32_t* idx, const uint8_t* shp,
                   long long npts, int aw, int rw, long long pix0, const int32_t* pairs, int npairs, float2* cov,
                   float2* coh, hipStream_t st);
constexpr int CO_MFMA_MAX_N = 64;
int co_mfma_wide_launch(const float2* rslc, int nl, int w, int n, const int32_t* idx, const uint8_t* shp,
                        long long npts, int aw, int rw, const int32_t* pairs, int npairs, float2* coh, hipStream_t st);
constexpr int CO_MFMA_WIDE_MIN_N = 65;   // point-cloud form on the 4-wave MFMA kernel
constexpr int CO_MFMA_WIDE_MAX_N = 224;  // for n in [65, 224]

constexpr int CO_THREADS = 256;
constexpr int CO_QB = 8;             // pairs per thread per batch
constexpr int CO_MAX_WIN = 4096;     // window entries
constexpr int CO_STAGE_BYTES = 32768;

// SAMP: batched form, point pt is pixel pix[pix0 + pt] whose K = n_samp[id]
// stacks are the rows of its sample matrix at rslc + row_off[id] (no window).
template <bool RASTER, bool SAMP = false>
__global__ __launch_bounds__(CO_THREADS) void co_kernel(const float2* __restrict__ rslc, int nl, int w, int n,
                                                        const int32_t* __restrict__ idx,
                                                        const uint8_t* __restrict__ shp, long long npts, int aw,
                                                        int rw, long long pix0, const int32_t* __restrict__ pairs,
                                                        int npairs, float2* __restrict__ cov,
                                                        float2* __restrict__ coh,
                                                        const long long* __restrict__ row_off = nullptr,
                                                        const int32_t* __restrict__ n_samp = nullptr,
                                                        const int32_t* __restrict__ pix = nullptr) {
  __shared__ int list[CO_MAX_WIN];
  __shared__ int wtot[CO_THREADS / 64];
  __shared__ int s_k;
  extern __shared__ __align__(16) unsigned char smem[];
  float2* stage = reinterpret_cast<float2*>(smem);
  float* pw = reinterpret_cast<float*>(smem + CO_STAGE_BYTES);
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int win = aw * rw;
  const int az = (aw - 1) / 2, r = (rw - 1) / 2;
  const int kc = CO_STAGE_BYTES / (8 * n);  // stacks per LDS chunk
  const int np = RASTER ? n * n : npairs;

  for (long long pt = blockIdx.x; pt < npts; pt += gridDim.x) {
    int ia = 0, ir = 0;
    const uint8_t* ms = nullptr;
    const float2* base = rslc;
    if (SAMP) {
      const int id = pix[pix0 + pt];
      base = rslc + row_off[id];
      if (tid == 0) s_k = min((int)n_samp[id], CO_MAX_WIN);
    } else if (RASTER) {
      const long long px = pix0 + pt;
      ia = (int)(px / w);
      ir = (int)(px - (long long)ia * w);
      ms = shp + px * win;
    } else {
      ia = idx[2 * pt];
      ir = idx[2 * pt + 1];
      ms = shp + pt * win;
    }
    // ---- compact in-bounds SHPs in window order ----
    if (!SAMP && tid == 0) s_k = 0;
    __syncthreads();
    for (int t0 = 0; t0 < (SAMP ? 0 : win); t0 += CO_THREADS) {
      const int t = t0 + tid;
      bool v = false;
      int lin = 0;
      if (t < win) {
        const int k = t / rw, l = t - (t / rw) * rw;
        const int sa = ia - az + k, sr = ir - r + l;
        v = ms[t] && sa >= 0 && sa < nl && sr >= 0 && sr < w;
        lin = sa * w + sr;
      }
      const unsigned long long b = __ballot(v);
      const int before = __popcll(b & ((1ull << lane) - 1ull));
      if (lane == 0) wtot[wv] = __popcll(b);
      __syncthreads();
      int off = s_k;
      for (int i = 0; i < wv; ++i) off += wtot[i];
      if (v) list[off + before] = lin;
      __syncthreads();
      if (tid == 0) {
        int s = s_k;
        for (int i = 0; i < CO_THREADS / 64; ++i) s += wtot[i];
        s_k = s;
      }
      __syncthreads();
    }
    const int K = s_k;
    const int nchunk = (K + kc - 1) / kc;
    float my_pw[4] = {0.f, 0.f, 0.f, 0.f};  // power sums for images tid + 256*q
    for (int pb = 0; pb < np; pb += CO_THREADS * CO_QB) {
      float accr[CO_QB], acci[CO_QB];
      int pa[CO_QB], pbb[CO_QB];
#pragma unroll
      for (int q = 0; q < CO_QB; ++q) {
        accr[q] = 0.f;
        acci[q] = 0.f;
        const int p = pb + q * CO_THREADS + tid;
        if (p < np) {
          if (RASTER) {
            pa[q] = p / n;
            pbb[q] = p - (p / n) * n;
          } else {
            pa[q] = pairs[2 * p];
            pbb[q] = pairs[2 * p + 1];
          }
        } else {
          pa[q] = 0;
          pbb[q] = 0;
        }
      }
      for (int ch = 0; ch < nchunk; ++ch) {
        const int k0 = ch * kc, kn = min(kc, K - k0);
        if (nchunk > 1 || pb == 0) {
          __syncthreads();
          for (int t = tid; t < kn * n; t += CO_THREADS) {
            const int s = t / n, e = t - (t / n) * n;
            stage[t] = base[(size_t)(SAMP ? k0 + s : list[k0 + s]) * n + e];
          }
          __syncthreads();
          if (pb == 0) {
#pragma unroll
            for (int q = 0; q < 4; ++q) {
              const int m = tid + q * CO_THREADS;
              if (m < n)
                for (int s = 0; s < kn; ++s) {
                  const float2 z = stage[s * n + m];
                  my_pw[q] += z.x * z.x + z.y * z.y;
                }
            }
          }
        }
        for (int s = 0; s < kn; ++s) {
          const float2* zs = stage + s * n;
#pragma unroll
          for (int q = 0; q < CO_QB; ++q) {
            const float2 x = zs[pa[q]], y = zs[pbb[q]];
            accr[q] = fmaf(x.x, y.x, fmaf(x.y, y.y, accr[q]));
            acci[q] = fmaf(x.y, y.x, fmaf(-x.x, y.y, acci[q]));
          }
        }
      }
      if (pb == 0) {
        __syncthreads();
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int m = tid + q * CO_THREADS;
          if (m < n) pw[m] = my_pw[q];
        }
        __syncthreads();
      }
#pragma unroll
      for (int q = 0; q < CO_QB; ++q) {
        const int p = pb + q * CO_THREADS + tid;
        if (p >= np) continue;
        const float den = sqrtf(pw[pa[q]] * pw[pbb[q]]);
        const float2 c = make_float2(accr[q] / den, acci[q] / den);
        if (RASTER) {
          const size_t o = (size_t)pt * n * n + p;
          coh[o] = c;
          if (cov) cov[o] = make_float2(accr[q] / (float)K, acci[q] / (float)K);
        } else {
          coh[(size_t)pt * npairs + p] = c;
        }
      }
    }
    __syncthreads();
  }
}

__global__ void uncompress_kernel(const float2* __restrict__ coh, long long npts, int n,
                                  const int32_t* __restrict__ pairs, int npairs, float2* __restrict__ out) {
  const long long nn = (long long)n * n;
  const long long tot = npts * nn;
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < tot; t += (long long)gridDim.x * blockDim.x) {
    const long long pt = t / nn;
    const int e = (int)(t - pt * nn);
    const int i = e / n, j = e - (e / n) * n;
    out[t] = (i == j) ? make_float2(1.f, 0.f) : make_float2(0.f, 0.f);
  }
}

__global__ void uncompress_scatter_kernel(const float2* __restrict__ coh, long long npts, int n,
                                          const int32_t* __restrict__ pairs, int npairs, float2* __restrict__ out) {
  const long long tot = npts * npairs;
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < tot; t += (long long)gridDim.x * blockDim.x) {
    const long long pt = t / npairs;
    const int p = (int)(t - pt * npairs);
    const int a = pairs[2 * p], b = pairs[2 * p + 1];
    const float2 c = coh[t];
    out[(pt * n + a) * n + b] = c;
    out[(pt * n + b) * n + a] = make_float2(c.x, -c.y);
  }
}

int co_samples_launch(const float2* samples, const long long* row_off, const int32_t* n_samp,
                      const int32_t* pix, long long pix0, long long cn, int n, const int32_t* pairs, int npairs,
                      float2* coh, hipStream_t st) {
  const int nb = (int)std::min<long long>(cn, 1 << 16);
  const size_t lds = CO_STAGE_BYTES + (size_t)n * 4;
  hipLaunchKernelGGL((co_kernel<false, true>), dim3(nb), dim3(CO_THREADS), lds, st, samples, 0, 0, n, nullptr,
                     nullptr, cn, 1, 1, pix0, pairs, npairs, nullptr, coh, row_off, n_samp, pix);
  return launch_status("ds_batch coherence");
}

static int co_check(int n, int aw, int rw) {
  if (n < 1 || 8 * n > CO_STAGE_BYTES || n > 4 * CO_THREADS) {
    set_error("emperical_co: n=%d unsupported (1..%d)", n, CO_STAGE_BYTES / 8);
    return MR_ERR_UNSUPPORTED;
  }
  if (aw < 1 || rw < 1 || aw * rw > CO_MAX_WIN) {
    set_error("emperical_co: window %dx%d unsupported (<= %d entries)", aw, rw, CO_MAX_WIN);
    return MR_ERR_UNSUPPORTED;
  }
  return MR_OK;
}

}  // namespace mr

extern "C" {

int mr_emperical_co(const float* rslc, const uint8_t* is_shp, int nl, int w, int n, int aw, int rw, int row0,
                    int nrows, float* cov, float* coh, mr_stream_t stream) {
  MR_CHECK_ARG(rslc && is_shp && coh, "emperical_co: null pointer");
  MR_CHECK_ARG(nl > 0 && w > 0 && row0 >= 0 && nrows >= 0 && row0 + nrows <= nl, "emperical_co: bad row range");
  if (int rc = mr::co_check(n, aw, rw)) return rc;
  const long long npts = (long long)nrows * w;
  if (npts == 0) return MR_OK;
  if (n <= mr::CO_MFMA_MAX_N)
    return mr::co_mfma_launch(true, reinterpret_cast<const float2*>(rslc), nl, w, n, nullptr, is_shp, npts, aw, rw,
                              (long long)row0 * w, nullptr, 0, reinterpret_cast<float2*>(cov),
                              reinterpret_cast<float2*>(coh), mr::as_stream(stream));
  const int nb = (int)std::min<long long>(npts, 1 << 16);
  const size_t lds = mr::CO_STAGE_BYTES + (size_t)n * 4;
  hipLaunchKernelGGL((mr::co_kernel<true>), dim3(nb), dim3(mr::CO_THREADS), lds, mr::as_stream(stream),
                     reinterpret_cast<const float2*>(rslc), nl, w, n, nullptr, is_shp, npts, aw, rw,
                     (long long)row0 * w, nullptr, 0, reinterpret_cast<float2*>(cov), reinterpret_cast<float2*>(coh));
  return mr::launch_status("emperical_co");
}

int mr_emperical_co_pc(const float* rslc, int nl, int w, int n, const int32_t* idx, const uint8_t* pc_is_shp, int n_pc,
                       int aw, int rw, const int32_t* pairs, int n_pairs, float* coh, mr_stream_t stream) {
  MR_CHECK_ARG(n_pc >= 0 && n_pairs >= 0, "emperical_co_pc: negative sizes");
  if (n_pc == 0 || n_pairs == 0) return MR_OK;
  MR_CHECK_ARG(rslc && idx && pc_is_shp && pairs && coh, "emperical_co_pc: null pointer");
  if (int rc = mr::co_check(n, aw, rw)) return rc;
  if (n <= mr::CO_MFMA_MAX_N)
    return mr::co_mfma_launch(false, reinterpret_cast<const float2*>(rslc), nl, w, n, idx, pc_is_shp, n_pc, aw, rw, 0,
                              pairs, n_pairs, nullptr, reinterpret_cast<float2*>(coh), mr::as_stream(stream));
  if (n >= mr::CO_MFMA_WIDE_MIN_N && n <= mr::CO_MFMA_WIDE_MAX_N)
    return mr::co_mfma_wide_launch(reinterpret_cast<const float2*>(rslc), nl, w, n, idx, pc_is_shp, n_pc, aw, rw,
                                   pairs, n_pairs, reinterpret_cast<float2*>(coh), mr::as_stream(stream));
  const int nb = std::min(n_pc, 1 << 16);
  const size_t lds = mr::CO_STAGE_BYTES + (size_t)n * 4;
  hipLaunchKernelGGL((mr::co_kernel<false>), dim3(nb), dim3(mr::CO_THREADS), lds, mr::as_stream(stream),
                     reinterpret_cast<const float2*>(rslc), nl, w, n, idx, pc_is_shp, (long long)n_pc, aw, rw, 0ll,
                     pairs, n_pairs, nullptr, reinterpret_cast<float2*>(coh));
  return mr::launch_status("emperical_co_pc");
}

int mr_uncompress_coh(const float* coh, long long n_pts, int n, const int32_t* pairs, int n_pairs, float* out,
                      mr_stream_t stream) {
  MR_CHECK_ARG(n >= 1 && n_pts >= 0 && n_pairs >= 0, "uncompress_coh: bad sizes");
  if (n_pts == 0) return MR_OK;
  MR_CHECK_ARG(out != nullptr, "uncompress_coh: null output");
  hipStream_t st = mr::as_stream(stream);
  const int nb = (int)std::min<long long>(mr::div_up((unsigned long long)n_pts * n * n, 256), 1 << 16);
  hipLaunchKernelGGL(mr::uncompress_kernel, dim3(nb), dim3(256), 0, st, reinterpret_cast<const float2*>(coh), n_pts, n,
                     pairs, n_pairs, reinterpret_cast<float2*>(out));
  if (int rc = mr::launch_status("uncompress_coh")) return rc;
  if (n_pairs == 0) return MR_OK;
  MR_CHECK_ARG(coh && pairs, "uncompress_coh: null input");
  const int nb2 = (int)std::min<long long>(mr::div_up((unsigned long long)n_pts * n_pairs, 256), 1 << 16);
  hipLaunchKernelGGL(mr::uncompress_scatter_kernel, dim3(nb2), dim3(256), 0, st, reinterpret_cast<const float2*>(coh),
                     n_pts, n, pairs, n_pairs, reinterpret_cast<float2*>(out));
  return mr::launch_status("uncompress_coh");
}

}  // extern "C"
