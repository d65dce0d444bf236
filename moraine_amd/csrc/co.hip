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

// Entry p of the row-major upper triangle (diagonal included) of an n x n
// matrix: row m holds n - m entries starting at m n - m (m - 1) / 2.
__device__ __forceinline__ void tri_index(int p, int n, int& m, int& j) {
  const float b = 2.f * n + 1.f;
  int r = (int)((b - sqrtf(b * b - 8.f * (float)p)) * 0.5f);
  r = max(0, min(r, n - 1));
  while (r > 0 && r * n - r * (r - 1) / 2 > p) --r;
  while (r + 1 < n && (r + 1) * n - (r + 1) * r / 2 <= p) ++r;
  m = r;
  j = r + (p - (r * n - r * (r - 1) / 2));
}

// SAMP: batched form, point pt is pixel pix[pix0 + pt] whose K = n_samp[id]
// stacks are the rows of its sample matrix at rslc + row_off[id] (no window).
// Z: float2 (complex64 stack, FP32 accumulators) or double2 (complex128 stack,
// FP64 accumulators and outputs, the reference's dtype rule coh.dtype = rslc.dtype).
template <bool RASTER, bool SAMP = false, typename Z = float2>
__global__ __launch_bounds__(CO_THREADS) void co_kernel(const Z* __restrict__ rslc, int nl, int w, int n,
                                                        const int32_t* __restrict__ idx,
                                                        const uint8_t* __restrict__ shp, long long npts, int aw,
                                                        int rw, long long pix0, const int32_t* __restrict__ pairs,
                                                        int npairs, Z* __restrict__ cov,
                                                        Z* __restrict__ coh,
                                                        const long long* __restrict__ row_off = nullptr,
                                                        const int32_t* __restrict__ n_samp = nullptr,
                                                        const int32_t* __restrict__ pix = nullptr) {
  __shared__ int list[CO_MAX_WIN];
  __shared__ int wtot[CO_THREADS / 64];
  __shared__ int s_k;
  using S = typename ZT<Z>::S;
  extern __shared__ __align__(16) unsigned char smem[];
  Z* stage = reinterpret_cast<Z*>(smem);
  S* pw = reinterpret_cast<S*>(smem + CO_STAGE_BYTES);
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int win = aw * rw;
  const int az = (aw - 1) / 2, r = (rw - 1) / 2;
  const int kc = CO_STAGE_BYTES / ((int)sizeof(Z) * n);  // stacks per LDS chunk
  // raster: the upper triangle m <= j (n(n+1)/2 entries) is accumulated and
  // written with its conjugate mirror, so the output is exactly Hermitian
  const int np = RASTER ? n * (n + 1) / 2 : npairs;

  for (long long pt = blockIdx.x; pt < npts; pt += gridDim.x) {
    int ia = 0, ir = 0;
    const uint8_t* ms = nullptr;
    const Z* base = rslc;
    if (SAMP) {
      const int id = pix[pix0 + pt];
      base = rslc + row_off[id];
      if (tid == 0) s_k = (int)n_samp[id];  // rows are read in place: no SHP list, no limit on K
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
    S my_pw[4] = {0, 0, 0, 0};  // power sums for images tid + 256*q
    for (int pb = 0; pb < np; pb += CO_THREADS * CO_QB) {
      S accr[CO_QB], acci[CO_QB];
      int pa[CO_QB], pbb[CO_QB];
#pragma unroll
      for (int q = 0; q < CO_QB; ++q) {
        accr[q] = 0;
        acci[q] = 0;
        const int p = pb + q * CO_THREADS + tid;
        if (p < np) {
          if (RASTER) {
            tri_index(p, n, pa[q], pbb[q]);
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
                  const Z z = stage[s * n + m];
                  my_pw[q] += z.x * z.x + z.y * z.y;
                }
            }
          }
        }
        for (int s = 0; s < kn; ++s) {
          const Z* zs = stage + s * n;
#pragma unroll
          for (int q = 0; q < CO_QB; ++q) {
            const Z x = zs[pa[q]], y = zs[pbb[q]];
            accr[q] = fma(x.x, y.x, fma(x.y, y.y, accr[q]));
            acci[q] = fma(x.y, y.x, fma(-x.x, y.y, acci[q]));
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
        const S den = sqrt(pw[pa[q]] * pw[pbb[q]]);
        const Z c = mkz<Z>(accr[q] / den, acci[q] / den);
        if (RASTER) {
          const size_t o = (size_t)pt * n * n + (size_t)pa[q] * n + pbb[q];
          const size_t o2 = (size_t)pt * n * n + (size_t)pbb[q] * n + pa[q];
          const Z cv = mkz<Z>(accr[q] / (S)K, acci[q] / (S)K);
          coh[o] = c;
          if (cov) cov[o] = cv;
          if (o2 != o) {
            coh[o2] = mkz<Z>(c.x, -c.y);
            if (cov) cov[o2] = mkz<Z>(cv.x, -cv.y);
          }
        } else {
          coh[(size_t)pt * npairs + p] = c;
        }
      }
    }
    __syncthreads();
  }
}

// uncompress_coh, each output entry written once.  A block owns a band of
// rows [row0, row0 + rows) of the N x N matrix and a run of points.  It first
// resolves, for every entry of its band, which of the reference's three
// assignments (moraine/co.py:296-298) lands there last, as a code in an LDS
// table: 0 zero, 1 the unit diagonal, 2 + p the pair assignment
// out[a_p, b_p] = coh[p], 2 + npairs + p the conjugate assignment
// out[b_p, a_p] = conj(coh[p]).  Later assignments carry larger codes (the
// conjugate pass after the direct one, a later duplicate after an earlier
// one), so an LDS atomicMax resolves collisions as NumPy's in-order assignment
// does, and the diagonal is set last.  The block then streams its points'
// bands, consecutive lanes on consecutive entries, gathering coh through the
// cache.  One band covers N <= 128.
constexpr int UNC_LUT_ENTRIES = 16384;  // 64 KiB of codes per block
constexpr int UNC_STAGE = 8192;         // rows of up to 64 KiB are staged in LDS
constexpr int UNC_STAGE_GROUP = 2048;   // G points' rows per stage: G = max(1, this / n_pairs)
constexpr int UNC_THREADS = 256;

// STAGE: the rows of G points at a time are copied to LDS (one coalesced read)
// and gathered from there; otherwise the gathers go through the cache, where
// the lower triangle's reads (a column of the upper triangle) touch one cache
// line per lane
template <bool STAGE, typename Z = float2>
__global__ __launch_bounds__(UNC_THREADS) void uncompress_kernel(const Z* __restrict__ coh, long long npts, int n,
                                                                 const int32_t* __restrict__ pairs, int npairs,
                                                                 int band_rows, long long pts_per_block, int G,
                                                                 Z* __restrict__ out) {
  extern __shared__ unsigned unc_lut[];
  const int row0 = blockIdx.y * band_rows;
  const int rows = min(band_rows, n - row0);
  const int ne = rows * n;
  for (int e = threadIdx.x; e < ne; e += UNC_THREADS) unc_lut[e] = 0u;
  __syncthreads();
  for (int p = threadIdx.x; p < npairs; p += UNC_THREADS) {
    const int a = pairs[2 * p], b = pairs[2 * p + 1];
    if ((unsigned)a >= (unsigned)n || (unsigned)b >= (unsigned)n) continue;  // the host rejects these
    if ((unsigned)(a - row0) < (unsigned)rows) atomicMax(&unc_lut[(a - row0) * n + b], (unsigned)(2 + p));
    if ((unsigned)(b - row0) < (unsigned)rows) atomicMax(&unc_lut[(b - row0) * n + a], (unsigned)(2 + npairs + p));
  }
  __syncthreads();
  for (int i = threadIdx.x; i < rows; i += UNC_THREADS) unc_lut[i * n + row0 + i] = 1u;
  __syncthreads();
  constexpr int al = (int)sizeof(Z) / 4;  // stage aligned to its element
  Z* stage = reinterpret_cast<Z*>(unc_lut + ((ne + al - 1) & ~(al - 1)));
  const long long p0 = (long long)blockIdx.x * pts_per_block;
  const long long p1 = min(npts, p0 + pts_per_block);
  const long long nn = (long long)n * n;
  for (long long g0 = p0; g0 < p1; g0 += STAGE ? G : p1 - p0) {
    const int gp = (int)min((long long)(STAGE ? G : p1 - p0), p1 - g0);
    const Z* cb = coh + g0 * npairs;
    if (STAGE) {
      const int m = gp * npairs;
      for (int t0 = threadIdx.x; t0 < m; t0 += 8 * UNC_THREADS) {
        Z v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          const int t = t0 + u * UNC_THREADS;
          v[u] = t < m ? cb[t] : mkz<Z>(0, 0);
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          const int t = t0 + u * UNC_THREADS;
          if (t < m) stage[t] = v[u];
        }
      }
      __syncthreads();
    }
    // the host keeps pts_per_block * ne below 2^31
    const unsigned tot = (unsigned)gp * (unsigned)ne;
    Z* ob = out + g0 * nn + (long long)row0 * n;
    constexpr int U = 4;
    for (unsigned t0 = threadIdx.x; t0 < tot; t0 += U * UNC_THREADS) {
      Z v[U];
      long long o[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const unsigned t = t0 + u * UNC_THREADS;
        v[u] = mkz<Z>(0, 0);
        o[u] = -1;
        if (t < tot) {
          const unsigned q = t / (unsigned)ne;
          const int e = (int)(t - q * (unsigned)ne);
          o[u] = (long long)q * nn + e;
          const unsigned code = unc_lut[e];
          if (code == 1u) {
            v[u].x = 1;
          } else if (code >= 2u) {
            const bool cj = code >= 2u + (unsigned)npairs;
            const int pi = (int)q * npairs + (int)(code - 2u - (cj ? (unsigned)npairs : 0u));
            const Z c = STAGE ? stage[pi] : cb[pi];
            v[u] = mkz<Z>(c.x, cj ? -c.y : c.y);
          }
        }
      }
#pragma unroll
      for (int u = 0; u < U; ++u)
        if (o[u] >= 0) ob[o[u]] = v[u];
    }
    if (STAGE) __syncthreads();
  }
}

int co_samples_launch(const float2* samples, const long long* row_off, const int32_t* n_samp,
                      const int32_t* pix, long long pix0, long long cn, int n, const int32_t* pairs, int npairs,
                      float2* coh, hipStream_t st) {
  const int nb = (int)std::min<long long>(cn, 1 << 16);
  const size_t lds = CO_STAGE_BYTES + (size_t)n * 4;
  hipLaunchKernelGGL((co_kernel<false, true, float2>), dim3(nb), dim3(CO_THREADS), lds, st, samples, 0, 0, n, nullptr,
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

// ad_intf_pc (moraine/co.py:302-406): adaptively multilooked interferogram of
// one image pair at each point: the SHPs of its window (inside the raster)
// give sum ref conj(sec) / sqrt(sum |ref|^2 sum |sec|^2).  One wave per point,
// lanes over the window, FP64 sums as the reference's CPU path (complex128 /
// float64 accumulators); no SHP gives 0/0 = NaN as there.
template <typename Z>
__global__ __launch_bounds__(256) void ad_intf_pc_kernel(const Z* __restrict__ ref, const Z* __restrict__ sec,
                                                         int nl, int w, const int32_t* __restrict__ idx,
                                                         const bool* __restrict__ shp, long long n_pc, int aw, int rw,
                                                         Z* __restrict__ out) {
  using S = typename ZT<Z>::S;
  const int lane = threadIdx.x & 63;
  const long long wave0 = ((long long)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const long long nw = ((long long)gridDim.x * blockDim.x) >> 6;
  const int win = aw * rw, ah = aw / 2, rh = rw / 2;
  for (long long i = wave0; i < n_pc; i += nw) {
    const int a0 = idx[2 * i] - ah, r0 = idx[2 * i + 1] - rh;
    double nr = 0.0, ni = 0.0, ra = 0.0, sa = 0.0;
    for (int e = lane; e < win; e += 64) {
      const int k = e / rw, l = e - (e / rw) * rw;
      const int az = a0 + k, rg = r0 + l;
      if (az >= 0 && az < nl && rg >= 0 && rg < w && shp[i * win + e]) {
        const Z x = ref[(long long)az * w + rg], y = sec[(long long)az * w + rg];
        const double xr = x.x, xi = x.y, yr = y.x, yi = y.y;
        ra += xr * xr + xi * xi;
        sa += yr * yr + yi * yi;
        nr += xr * yr + xi * yi;  // x conj(y)
        ni += xi * yr - xr * yi;
      }
    }
    nr = wave_sum(nr);
    ni = wave_sum(ni);
    ra = wave_sum(ra);
    sa = wave_sum(sa);
    if (lane == 0) {
      const double d = sqrt(ra * sa);
      out[i] = mkz<Z>((S)(nr / d), (S)(ni / d));
    }
  }
}

// regularize_spectral (moraine/co.py:466-477): (1 - beta) coh + beta I for a
// stack of n x n complex matrices, beta one value or one per matrix; computed
// in FP64 (product, then sum: no contraction, as NumPy rounds), rounded to
// complex64.
#pragma clang fp contract(off)
template <typename Z>
__global__ __launch_bounds__(256) void regularize_spectral_kernel(const Z* __restrict__ c, long long nmat, int n,
                                                                  const double* __restrict__ beta, int per_matrix,
                                                                  Z* __restrict__ out) {
  using S = typename ZT<Z>::S;
  const long long nn = (long long)n * n, tot = nmat * nn;
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < tot; t += (long long)gridDim.x * blockDim.x) {
    const long long m = t / nn;
    const int e = (int)(t - m * nn);
    const double b = beta[per_matrix ? m : 0];
    const Z v = c[t];
    double re = (1.0 - b) * (double)v.x, im = (1.0 - b) * (double)v.y;
    if (e / n == e % n) re += b;
    out[t] = mkz<Z>((S)re, (S)im);
  }
}
#pragma clang fp contract(on)

// complex128 forms: the VALU kernels with FP64 accumulators and outputs.
static int co_c128_launch(bool raster, const double* rslc, int nl, int w, int n, const int32_t* idx,
                          const uint8_t* shp, long long npts, int aw, int rw, long long pix0, const int32_t* pairs,
                          int npairs, double* cov, double* coh, hipStream_t st) {
  const int nb = (int)std::min<long long>(npts, 1 << 16);
  const size_t lds = CO_STAGE_BYTES + (size_t)n * 8;
  if (raster)
    hipLaunchKernelGGL((co_kernel<true, false, double2>), dim3(nb), dim3(CO_THREADS), lds, st,
                       reinterpret_cast<const double2*>(rslc), nl, w, n, nullptr, shp, npts, aw, rw, pix0, nullptr, 0,
                       reinterpret_cast<double2*>(cov), reinterpret_cast<double2*>(coh), nullptr, nullptr, nullptr);
  else
    hipLaunchKernelGGL((co_kernel<false, false, double2>), dim3(nb), dim3(CO_THREADS), lds, st,
                       reinterpret_cast<const double2*>(rslc), nl, w, n, idx, shp, npts, aw, rw, 0ll, pairs, npairs,
                       nullptr, reinterpret_cast<double2*>(coh), nullptr, nullptr, nullptr);
  return launch_status(raster ? "emperical_co_c128" : "emperical_co_pc_c128");
}

}  // namespace mr

extern "C" {

int mr_emperical_co_c128(const double* rslc, const uint8_t* is_shp, int nl, int w, int n, int aw, int rw, int row0,
                         int nrows, double* cov, double* coh, mr_stream_t stream) {
  MR_CHECK_ARG(rslc && is_shp && coh, "emperical_co_c128: null pointer");
  MR_CHECK_ARG(nl > 0 && w > 0 && row0 >= 0 && nrows >= 0 && row0 + nrows <= nl, "emperical_co_c128: bad row range");
  if (int rc = mr::co_check(n, aw, rw)) return rc;
  const long long npts = (long long)nrows * w;
  if (npts == 0) return MR_OK;
  return mr::co_c128_launch(true, rslc, nl, w, n, nullptr, is_shp, npts, aw, rw, (long long)row0 * w, nullptr, 0, cov,
                            coh, mr::as_stream(stream));
}

int mr_emperical_co_pc_c128(const double* rslc, int nl, int w, int n, const int32_t* idx, const uint8_t* pc_is_shp,
                            int n_pc, int aw, int rw, const int32_t* pairs, int n_pairs, double* coh,
                            mr_stream_t stream) {
  MR_CHECK_ARG(n_pc >= 0 && n_pairs >= 0, "emperical_co_pc_c128: negative sizes");
  if (n_pc == 0 || n_pairs == 0) return MR_OK;
  MR_CHECK_ARG(rslc && idx && pc_is_shp && pairs && coh, "emperical_co_pc_c128: null pointer");
  if (int rc = mr::co_check(n, aw, rw)) return rc;
  return mr::co_c128_launch(false, rslc, nl, w, n, idx, pc_is_shp, n_pc, aw, rw, 0, pairs, n_pairs, nullptr, coh,
                            mr::as_stream(stream));
}

int mr_uncompress_coh_c128(const double* coh, long long n_pts, int n, const int32_t* pairs, int n_pairs, double* out,
                           mr_stream_t stream) {
  MR_CHECK_ARG(n >= 1 && n_pts >= 0 && n_pairs >= 0, "uncompress_coh_c128: bad sizes");
  MR_CHECK_ARG(n <= mr::UNC_LUT_ENTRIES, "uncompress_coh_c128: n = %d above %d", n, mr::UNC_LUT_ENTRIES);
  MR_CHECK_ARG(n_pairs <= (1 << 30), "uncompress_coh_c128: %d pairs above 2^30", n_pairs);
  if (n_pts == 0) return MR_OK;
  MR_CHECK_ARG(out != nullptr, "uncompress_coh_c128: null output");
  MR_CHECK_ARG(n_pairs == 0 || (coh && pairs), "uncompress_coh_c128: null input");
  const int band = std::min(n, mr::UNC_LUT_ENTRIES / n);
  const int nbands = (int)mr::div_up(n, band);
  const long long nb = std::min<long long>(n_pts, std::max(1, 1024 / nbands));
  const long long per = std::min<long long>(mr::div_up((unsigned long long)n_pts, (unsigned long long)nb),
                                            ((1ll << 31) - 1 - 4 * mr::UNC_THREADS) / ((long long)band * n));
  const long long blocks = (n_pts + per - 1) / per;
  MR_CHECK_ARG(blocks < (1ll << 31), "uncompress_coh_c128: %lld points too many", n_pts);
  const size_t lut = ((size_t)band * n + 3) / 4 * 16;
  hipLaunchKernelGGL((mr::uncompress_kernel<false, double2>), dim3((unsigned)blocks, (unsigned)nbands),
                     dim3(mr::UNC_THREADS), lut, mr::as_stream(stream), reinterpret_cast<const double2*>(coh), n_pts,
                     n, pairs, n_pairs, band, per, 1, reinterpret_cast<double2*>(out));
  return mr::launch_status("uncompress_coh_c128");
}

int mr_ad_intf_pc_c128(const double* ref_rslc, const double* sec_rslc, int nl, int w, const int32_t* idx,
                       const uint8_t* pc_is_shp, long long n_pc, int aw, int rw, double* intf, mr_stream_t stream) {
  MR_CHECK_ARG(nl > 0 && w > 0 && n_pc >= 0, "ad_intf_pc_c128: bad sizes");
  MR_CHECK_ARG(aw >= 1 && rw >= 1 && (aw & 1) && (rw & 1), "ad_intf_pc_c128: window %d x %d must be odd", aw, rw);
  if (n_pc == 0) return MR_OK;
  MR_CHECK_ARG(ref_rslc && sec_rslc && idx && pc_is_shp && intf, "ad_intf_pc_c128: null pointer");
  const int nb = (int)std::min<long long>(mr::div_up((unsigned long long)n_pc, 4), 1 << 16);
  hipLaunchKernelGGL(mr::ad_intf_pc_kernel<double2>, dim3(nb), dim3(256), 0, mr::as_stream(stream),
                     reinterpret_cast<const double2*>(ref_rslc), reinterpret_cast<const double2*>(sec_rslc), nl, w,
                     idx, reinterpret_cast<const bool*>(pc_is_shp), n_pc, aw, rw, reinterpret_cast<double2*>(intf));
  return mr::launch_status("ad_intf_pc_c128");
}

int mr_regularize_spectral_c128(const double* coh, long long n_mat, int n, const double* beta, int beta_per_matrix,
                                double* out, mr_stream_t stream) {
  MR_CHECK_ARG(n_mat >= 0 && n >= 1, "regularize_spectral_c128: bad sizes");
  if (n_mat == 0) return MR_OK;
  MR_CHECK_ARG(coh && beta && out, "regularize_spectral_c128: null pointer");
  const int nb = (int)std::min<long long>(mr::div_up((unsigned long long)n_mat * n * n, 256), 1 << 16);
  hipLaunchKernelGGL(mr::regularize_spectral_kernel<double2>, dim3(nb), dim3(256), 0, mr::as_stream(stream),
                     reinterpret_cast<const double2*>(coh), n_mat, n, beta, beta_per_matrix,
                     reinterpret_cast<double2*>(out));
  return mr::launch_status("regularize_spectral_c128");
}

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
  hipLaunchKernelGGL((mr::co_kernel<true, false, float2>), dim3(nb), dim3(mr::CO_THREADS), lds, mr::as_stream(stream),
                     reinterpret_cast<const float2*>(rslc), nl, w, n, nullptr, is_shp, npts, aw, rw,
                     (long long)row0 * w, nullptr, 0, reinterpret_cast<float2*>(cov), reinterpret_cast<float2*>(coh));
  return mr::launch_status("emperical_co");
}

int mr_emperical_co_pc(const float* rslc, int nl, int w, int n, const int32_t* idx, const uint8_t* pc_is_shp, int n_pc,
                       int aw, int rw, const int32_t* pairs, int n_pairs, float* coh, mr_stream_t stream) {
  MR_CHECK_ARG(n_pc >= 0 && n_pairs >= 0, "emperical_co_pc: negative sizes");
  if (n_pc == 0 || n_pairs == 0) return MR_OK;
  MR_CHECK_ARG(nl > 0 && w > 0, "emperical_co_pc: empty raster (%d x %d) with %d points", nl, w, n_pc);
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
  hipLaunchKernelGGL((mr::co_kernel<false, false, float2>), dim3(nb), dim3(mr::CO_THREADS), lds, mr::as_stream(stream),
                     reinterpret_cast<const float2*>(rslc), nl, w, n, idx, pc_is_shp, (long long)n_pc, aw, rw, 0ll,
                     pairs, n_pairs, nullptr, reinterpret_cast<float2*>(coh));
  return mr::launch_status("emperical_co_pc");
}

int mr_uncompress_coh(const float* coh, long long n_pts, int n, const int32_t* pairs, int n_pairs, float* out,
                      mr_stream_t stream) {
  MR_CHECK_ARG(n >= 1 && n_pts >= 0 && n_pairs >= 0, "uncompress_coh: bad sizes");
  MR_CHECK_ARG(n <= mr::UNC_LUT_ENTRIES, "uncompress_coh: n = %d above %d", n, mr::UNC_LUT_ENTRIES);
  MR_CHECK_ARG(n_pairs <= (1 << 30), "uncompress_coh: %d pairs above 2^30", n_pairs);
  if (n_pts == 0) return MR_OK;
  MR_CHECK_ARG(out != nullptr, "uncompress_coh: null output");
  MR_CHECK_ARG(n_pairs == 0 || (coh && pairs), "uncompress_coh: null input");
  const int band = std::min(n, mr::UNC_LUT_ENTRIES / n);
  const int nbands = (int)mr::div_up(n, band);
  // one band and rows that fit: stage G points' rows at a time in LDS
  const bool stage = nbands == 1 && n_pairs >= 1 && n_pairs <= mr::UNC_STAGE;
  const int G = stage ? std::max(1, mr::UNC_STAGE_GROUP / n_pairs) : 1;
  // four blocks per CU, each over a run of whole points (more blocks build
  // more tables for no gain: 12% slower at 4096 blocks, n = 60)
  const long long nb = std::min<long long>(n_pts, std::max(1, 1024 / nbands));
  const long long per = std::min<long long>(mr::div_up((unsigned long long)n_pts, (unsigned long long)nb),
                                            ((1ll << 31) - 1 - 4 * mr::UNC_THREADS) / ((long long)band * n));
  const long long blocks = (n_pts + per - 1) / per;
  MR_CHECK_ARG(blocks < (1ll << 31), "uncompress_coh: %lld points too many", n_pts);
  const size_t lut = ((size_t)band * n + 1) / 2 * 8;
  if (stage)
    hipLaunchKernelGGL(mr::uncompress_kernel<true>, dim3((unsigned)blocks, (unsigned)nbands), dim3(mr::UNC_THREADS),
                       lut + (size_t)G * n_pairs * 8, mr::as_stream(stream), reinterpret_cast<const float2*>(coh),
                       n_pts, n, pairs, n_pairs, band, per, G, reinterpret_cast<float2*>(out));
  else
    hipLaunchKernelGGL(mr::uncompress_kernel<false>, dim3((unsigned)blocks, (unsigned)nbands), dim3(mr::UNC_THREADS),
                       lut, mr::as_stream(stream), reinterpret_cast<const float2*>(coh), n_pts, n, pairs, n_pairs,
                       band, per, 1, reinterpret_cast<float2*>(out));
  return mr::launch_status("uncompress_coh");
}

int mr_ad_intf_pc(const float* ref_rslc, const float* sec_rslc, int nl, int w, const int32_t* idx,
                  const uint8_t* pc_is_shp, long long n_pc, int aw, int rw, float* intf, mr_stream_t stream) {
  MR_CHECK_ARG(nl > 0 && w > 0 && n_pc >= 0, "ad_intf_pc: bad sizes");
  MR_CHECK_ARG(aw >= 1 && rw >= 1 && (aw & 1) && (rw & 1), "ad_intf_pc: window %d x %d must be odd", aw, rw);
  if (n_pc == 0) return MR_OK;
  MR_CHECK_ARG(ref_rslc && sec_rslc && idx && pc_is_shp && intf, "ad_intf_pc: null pointer");
  const int nb = (int)std::min<long long>(mr::div_up((unsigned long long)n_pc, 4), 1 << 16);
  hipLaunchKernelGGL(mr::ad_intf_pc_kernel<float2>, dim3(nb), dim3(256), 0, mr::as_stream(stream),
                     reinterpret_cast<const float2*>(ref_rslc), reinterpret_cast<const float2*>(sec_rslc), nl, w, idx,
                     reinterpret_cast<const bool*>(pc_is_shp), n_pc, aw, rw, reinterpret_cast<float2*>(intf));
  return mr::launch_status("ad_intf_pc");
}

int mr_regularize_spectral(const float* coh, long long n_mat, int n, const double* beta, int beta_per_matrix,
                           float* out, mr_stream_t stream) {
  MR_CHECK_ARG(n_mat >= 0 && n >= 1, "regularize_spectral: bad sizes");
  if (n_mat == 0) return MR_OK;
  MR_CHECK_ARG(coh && beta && out, "regularize_spectral: null pointer");
  const int nb = (int)std::min<long long>(mr::div_up((unsigned long long)n_mat * n * n, 256), 1 << 16);
  hipLaunchKernelGGL(mr::regularize_spectral_kernel<float2>, dim3(nb), dim3(256), 0, mr::as_stream(stream),
                     reinterpret_cast<const float2*>(coh), n_mat, n, beta, beta_per_matrix,
                     reinterpret_cast<float2*>(out));
  return mr::launch_status("regularize_spectral");
}

}  // extern "C"
