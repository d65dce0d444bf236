// Coherence of one point on MFMA (gfx950), n <= 64, one wave per point.
//
// Per point the SHP stacks z_k (k = 1..K, in the reference's window order,
// moraine/co.py:83-111) form the Hermitian rank-K update S = sum_k z_k z_k^H.
// With the real 2x2 embedding
//     Re S = X X^T,  Im S = Y X^T,   X = [Re z | Im z],  Y = [Im z | -Re z]
// both halves are real GEMMs over 2K "k-slots" (slot 2k = real part of SHP k,
// slot 2k+1 = imaginary part).  They run on v_mfma_f32_32x32x16_f16 with every
// operand split into FP16 hi + lo (x = hi + lo, |x - hi - lo| <= 2^-22 |x|); the
// three products hi*hi + hi*lo + lo*hi accumulate in FP32, which gives FP32-level
// products.  Before the split every image row is scaled by a power of two so its
// largest component lies in [32, 64): products are exact in FP32, sums stay far
// from overflow, and coherence is invariant under the per-image scaling
// (cov is unscaled exactly, the factors being powers of two).
// Output tiles: (0,0) for n <= 32; (0,0), (0,1), (1,1) for 32 < n <= 64 (the
// (1,0) tile is the conjugate transpose of (0,1)).
//
// The stacks of a point are rows base[row(k) * n + image]: row() is a window
// list (raster and point-cloud forms) or the identity (per-pixel sample
// matrices of the batched form); both run the same arithmetic.
#pragma once
#include "common.hpp"

namespace mr {

typedef _Float16 half8_t __attribute__((ext_vector_type(8)));
typedef _Float16 half2_t __attribute__((ext_vector_type(2)));
typedef float f16x_t __attribute__((ext_vector_type(16)));
typedef unsigned int u4_t __attribute__((ext_vector_type(4)));

__host__ __device__ inline int com_tile_floats(bool two) { return two ? 3 * 32 * 32 * 2 : 32 * 32 * 2; }

struct RowList {  // k -> window list entry (LDS)
  const int* l;
  __device__ __forceinline__ int operator()(int k) const { return l[k]; }
};
struct RowIdentity {  // k -> k (contiguous sample matrix)
  __device__ __forceinline__ int operator()(int k) const { return k; }
};

// In-bounds SHPs of a window in window order: list[0..K) = linear pixel index
// (row * w + col) of every flagged window position inside the raster; returns
// K (uniform).  The mask bytes of four 64-position groups are loaded before
// the first ballot, so their latencies overlap.
__device__ __forceinline__ int co_compact_shp(const uint8_t* __restrict__ ms, int win, int rw, int ia, int ir, int az,
                                              int r, int nl, int w, int lane, int* list) {
  int K = 0;
  for (int t0 = 0; t0 < win; t0 += 256) {
    uint8_t m[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int t = t0 + 64 * j + lane;
      m[j] = t < win ? ms[t] : 0;
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int t = t0 + 64 * j + lane;
      const int k = t / rw, l = t - (t / rw) * rw;
      const int sa = ia - az + k, sr = ir - r + l;
      const bool v = m[j] && sa >= 0 && sa < nl && sr >= 0 && sr < w;
      const unsigned long long b = __ballot(v);
      if (v) list[K + __popcll(b & ((1ull << lane) - 1ull))] = sa * w + sr;
      K += __popcll(b);
    }
  }
  return K;
}

// Point order inside an XCD's range (point-cloud form).  When the range is a
// run of consecutive raster pixels (its first and last points `len` - 1 pixels
// apart in raster order: the all-pixel batches of the batch driver), slot s
// walks strips of at most CO_STRIP raster rows column by column, so the points
// in flight on the XCD form a compact patch (about 16 x 16 pixels) whose SHP
// windows share its L2; otherwise slot s is point s.  Either way every point
// gets exactly one slot (a bijection on [0, len) plus empty slots), so the
// output does not depend on it.
constexpr int CO_STRIP = 16;
struct CoStrips {
  int w = 0, c0 = 0, h = 0;  // h = 0: identity order
  long long strip = 0;       // h * w slots per strip
  int nstrips = 0;
  __device__ __forceinline__ long long slots(long long len) const { return h ? strip * nstrips : len; }
  // point offset of slot s, -1 for an empty slot
  __device__ __forceinline__ long long point(long long s, long long len) const {
    if (!h) return s;
    const long long k = s / strip, rem = s - k * strip;
    const long long col = rem / h, lr = k * h + (rem - col * h);
    const long long u = lr * w + col - c0;
    return u >= 0 && u < len ? u : -1;
  }
};
// Uniform over the block: from the first and last points of the range.
__device__ __forceinline__ CoStrips co_strips(const int32_t* __restrict__ idx, long long beg, long long end, int w) {
  CoStrips st;
  if (end - beg < 2 * (long long)w) return st;  // shorter than two rows: identity order
  const int r0 = idx[2 * beg], c0 = idx[2 * beg + 1];
  const int r1 = idx[2 * (end - 1)], c1 = idx[2 * (end - 1) + 1];
  const long long len = end - beg;
  if ((long long)(r1 - r0) * w + (c1 - c0) + 1 != len || c0 < 0 || c0 >= w) return st;
  const int rows = (int)((c0 + len + w - 1) / w);  // raster rows the run touches
  st.nstrips = (rows + CO_STRIP - 1) / CO_STRIP;
  st.h = (rows + st.nstrips - 1) / st.nstrips;
  st.w = w;
  st.c0 = c0;
  st.strip = (long long)st.h * w;
  return st;
}
// Slots per XCD range of `per` points on a raster of width w, at most.
__host__ __device__ inline long long co_max_slots(long long per, int w) {
  const long long rows = (w - 1 + per + w - 1) / w;
  const long long ns = (rows + CO_STRIP - 1) / CO_STRIP;
  return std::max(per, ((rows + ns - 1) / ns) * ns * (long long)w);
}

// Split 4 complex values (8 floats: re0, im0, re1, im1, ...) into FP16 hi/lo
// halves packed as the MFMA operand: hi = RTZ(x), lo = RTZ(x - hi) (exact
// remainder, then rounded).  Also returns Y = (im, -re) pairs derived from the
// same halves (a swap of the two halves of each word and a sign flip), which is
// the operand of Im S = Y X^T.
__device__ __forceinline__ void split4c(const float (&v)[8], half8_t& xh, half8_t& xl, half8_t& yh, half8_t& yl) {
  u4_t h, l, hy, ly;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const half2_t hh = __builtin_bit_cast(half2_t, __builtin_amdgcn_cvt_pkrtz(v[2 * i], v[2 * i + 1]));
    const float r0 = v[2 * i] - (float)hh[0];
    const float r1 = v[2 * i + 1] - (float)hh[1];
    const unsigned hw = __builtin_bit_cast(unsigned, hh);
    const unsigned lw = __builtin_bit_cast(unsigned, __builtin_bit_cast(half2_t, __builtin_amdgcn_cvt_pkrtz(r0, r1)));
    h[i] = hw;
    l[i] = lw;
    // (re, im) -> (im, -re): rotate the 16-bit halves, flip the sign of the new high half
    hy[i] = ((hw >> 16) | (hw << 16)) ^ 0x80000000u;
    ly[i] = ((lw >> 16) | (lw << 16)) ^ 0x80000000u;
  }
  xh = __builtin_bit_cast(half8_t, h);
  xl = __builtin_bit_cast(half8_t, l);
  yh = __builtin_bit_cast(half8_t, hy);
  yl = __builtin_bit_cast(half8_t, ly);
}

__device__ __forceinline__ f16x_t mfma3(const half8_t& ah, const half8_t& al, const half8_t& bh, const half8_t& bl,
                                        f16x_t c) {
  c = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah, bh, c, 0, 0, 0);
  c = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah, bl, c, 0, 0, 0);
  c = __builtin_amdgcn_mfma_f32_32x32x16_f16(al, bh, c, 0, 0, 0);
  return c;
}

// Accumulators of one wave: the upper tiles of S' = sum_k (D z_k)(D z_k)^H
// (real and imaginary parts of tiles (0,0), (0,1), (1,1)), D the per-image
// power-of-two scaling; s = this lane's image scale (lane = image).
struct CoAcc {
  f16x_t r00, i00, r01, i01, r11, i11;
  float s;
};

// Per-image power-of-two scale (lane = image): 2^(6 - e) with 2^(e-1) <= the
// image's largest |component| over the K rows < 2^e, so the scaled maximum
// lies in [32, 64); 1 when that maximum is 0 or not finite.  Sixteen SHPs per
// trip: the row lookups and gathers of a trip are all in flight before the
// first maximum needs them; a short last trip re-reads SHP K-1 (a repeated
// sample leaves the maximum unchanged).
template <class Row>
__device__ __forceinline__ float co_image_scale(const float2* __restrict__ base, const Row& row, int K, int n,
                                                int lane) {
  float mx = 0.f;
  if (lane < n) {
    for (int k = 0; k < K; k += 16) {
      float2 a[16];
#pragma unroll
      for (int u = 0; u < 16; ++u) a[u] = base[(size_t)row(k + u < K ? k + u : K - 1) * n + lane];
#pragma unroll
      for (int u = 0; u < 16; ++u) mx = fmaxf(mx, fmaxf(fabsf(a[u].x), fabsf(a[u].y)));
    }
  }
  float s = 1.f;
  if (mx > 0.f && mx < INFINITY) {
    int e;
    frexpf(mx, &e);
    s = ldexpf(1.f, 6 - e);  // scaled maximum in [32, 64)
  }
  return s;
}

// PRE: the caller passes this lane's image scale (from co_image_scale) as s_pre;
// otherwise it is computed here first.
template <bool TWO, class Row, bool PRE = false>
__device__ __forceinline__ CoAcc co_mfma_accumulate(const float2* __restrict__ base, const Row& row, int K, int n,
                                                    int lane, float s_pre = 1.f) {
  const int hf = lane >> 5;  // which 8-slot half of a k-chunk this lane feeds
  const int r0 = lane & 31;  // row of block 0 (= column of block 0 as B operand)
  const int r1 = 32 + (lane & 31);
  const float s = PRE ? s_pre : co_image_scale(base, row, K, n, lane);
  const float s0 = __shfl(s, r0, 64);
  const float s1 = TWO ? __shfl(s, r1 < 64 ? r1 : 0, 64) : 1.f;
  // ---- main loop: 8 SHPs (16 k-slots) per chunk -----------------------------
  f16x_t cr00 = {}, ci00 = {}, cr01 = {}, ci01 = {}, cr11 = {}, ci11 = {};
  const int nch = (K + 7) >> 3;
  float2 za[4], zb[4];  // this lane's 4 SHPs at rows r0 / r1 (prefetched one chunk ahead)
  auto load_chunk = [&](int c, float2(&xa)[4], float2(&xb)[4]) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int k = c * 8 + hf * 4 + q;
      const bool okk = k < K;
      const size_t off = okk ? (size_t)row(k) * n : 0;
      xa[q] = (okk && r0 < n) ? base[off + r0] : make_float2(0.f, 0.f);
      if (TWO) xb[q] = (okk && r1 < n) ? base[off + r1] : make_float2(0.f, 0.f);
    }
  };
  if (nch > 0) load_chunk(0, za, zb);
  for (int c = 0; c < nch; ++c) {
    float2 na[4], nb[4];
    if (c + 1 < nch) load_chunk(c + 1, na, nb);
    float xv[8];
    half8_t x0h, x0l, y0h, y0l;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      xv[2 * q] = za[q].x * s0;
      xv[2 * q + 1] = za[q].y * s0;
    }
    split4c(xv, x0h, x0l, y0h, y0l);
    cr00 = mfma3(x0h, x0l, x0h, x0l, cr00);
    ci00 = mfma3(y0h, y0l, x0h, x0l, ci00);
    if (TWO) {
      half8_t x1h, x1l, y1h, y1l;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        xv[2 * q] = zb[q].x * s1;
        xv[2 * q + 1] = zb[q].y * s1;
      }
      split4c(xv, x1h, x1l, y1h, y1l);
      cr01 = mfma3(x0h, x0l, x1h, x1l, cr01);
      ci01 = mfma3(y0h, y0l, x1h, x1l, ci01);
      cr11 = mfma3(x1h, x1l, x1h, x1l, cr11);
      ci11 = mfma3(y1h, y1l, x1h, x1l, ci11);
    }
    if (c + 1 < nch) {
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        za[q] = na[q];
        zb[q] = nb[q];
      }
    }
  }
  return CoAcc{cr00, ci00, cr01, ci01, cr11, ci11, s};
}

// Accumulator layout of v_mfma_f32_32x32x16: lane l, register q holds row
// 8 (q / 4) + 4 (l / 32) + q % 4, column l % 32 of the tile.
__device__ __forceinline__ void co_store_tile(float2* tile, const f16x_t& re, const f16x_t& im, int lane) {
  const int hf = lane >> 5;
#pragma unroll
  for (int q = 0; q < 16; ++q) tile[(8 * (q >> 2) + 4 * hf + (q & 3)) * 32 + (lane & 31)] = make_float2(re[q], im[q]);
}

// sc[0..64) = scales, sc[64..128) = 1/sqrt(S'_mm), taken from the diagonal
// entries in the accumulators (no tile in LDS needed).
template <bool TWO>
__device__ __forceinline__ void co_store_scales(const CoAcc& acc, int n, int lane, float* sc) {
  sc[lane] = acc.s;
  const int hf = lane >> 5, c = lane & 31;
  // the lane of column c whose row group holds row c owns the diagonal entry:
  // select it first, then one reciprocal square root per lane
  bool own = false;
  float d0 = 0.f, d1 = 0.f;
#pragma unroll
  for (int q = 0; q < 16; ++q) {
    const bool hit = 8 * (q >> 2) + 4 * hf + (q & 3) == c;
    own = own || hit;
    d0 = hit ? acc.r00[q] : d0;
    if (TWO) d1 = hit ? acc.r11[q] : d1;
  }
  if (own) {
    sc[64 + c] = 1.f / sqrtf(c < n ? d0 : 0.f);
    if (TWO) sc[96 + c] = 1.f / sqrtf(32 + c < n ? d1 : 0.f);
  }
  if (!TWO && lane >= 32) sc[64 + lane] = 1.f / sqrtf(0.f);
  __builtin_amdgcn_wave_barrier();
}

// index of pair (a, b), a < b, in triu order (moraine/tnet.py:15-20)
__device__ __forceinline__ int co_pair_index(int a, int b, int n) { return a * (2 * n - a - 1) / 2 + (b - a - 1); }

// True on every thread of the block if `pairs` is the full network in triu
// order (the layout moraine uses when image_pairs is None); checked once per
// block, it lets tiles be written straight from the accumulators.  Eight pair
// loads per thread are in flight at once: a block pays about one cache
// latency per eight rounds of the list instead of one per round (with one
// point per wave and a round per load, the check took ~8% of the point-cloud
// coherence kernel at n = 60).
__device__ __forceinline__ bool co_pairs_canonical(const int32_t* __restrict__ pairs, int npairs, int n) {
  bool bad = npairs != n * (n - 1) / 2;
  if (!bad && (reinterpret_cast<uintptr_t>(pairs) & 7)) {  // not 8-byte aligned: one pair per load
    for (int p = threadIdx.x; p < npairs; p += blockDim.x) {
      const int a = pairs[2 * p], b = pairs[2 * p + 1];
      bad |= !(a >= 0 && a < b && b < n && co_pair_index(a, b, n) == p);
    }
  } else if (!bad) {
    const int2* pr = reinterpret_cast<const int2*>(pairs);
    const int st = blockDim.x;
    for (int p0 = threadIdx.x; p0 < npairs; p0 += 8 * st) {
      int2 v[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) v[u] = p0 + u * st < npairs ? pr[p0 + u * st] : make_int2(0, 1);
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int p = p0 + u * st;
        const int a = v[u].x, b = v[u].y;
        bad |= p < npairs && !(a >= 0 && a < b && b < n && co_pair_index(a, b, n) == p);
      }
    }
  }
  return !__syncthreads_or(bad);
}

// Coherence of the canonical pairs inside upper tile (ta, tb), written from the
// accumulators: register q of lane l is element (32 ta + row(q, l), 32 tb + l % 32),
// so for each q the 32 lanes of a half-wave write 32 consecutive pairs.
__device__ __forceinline__ void co_write_tile_direct(const f16x_t& re, const f16x_t& im, int ta, int tb, int n,
                                                     int lane, const float* rs, float2* __restrict__ out) {
  const int hf = lane >> 5;
  const int B = 32 * tb + (lane & 31);
  const float rb = rs[B];
#pragma unroll
  for (int q = 0; q < 16; ++q) {
    const int A = 32 * ta + 8 * (q >> 2) + 4 * hf + (q & 3);
    if (A < B && B < n) {
      const float f = rs[A] * rb;
      out[(unsigned)co_pair_index(A, B, n)] = make_float2(re[q] * f, im[q] * f);
    }
    if ((q & 3) == 3) __builtin_amdgcn_sched_barrier(0);  // bounds the addresses in flight
  }
}

// One wave: accumulate, then (after a wave barrier) `tiles` holds the upper
// tiles of S' (row-major 32x32 complex each), sc[0..64) the scales and
// sc[64..128) 1/sqrt(S'_mm), so that coh_mj = S'_mj sc[64+m] sc[64+j].
template <bool TWO, class Row>
__device__ __forceinline__ void co_mfma_point(const float2* __restrict__ base, const Row& row, int K, int n, int lane,
                                              float2* tiles, float* sc) {
  const CoAcc acc = co_mfma_accumulate<TWO>(base, row, K, n, lane);
  __builtin_amdgcn_wave_barrier();
  co_store_tile(tiles, acc.r00, acc.i00, lane);
  if (TWO) {
    co_store_tile(tiles + 1024, acc.r01, acc.i01, lane);
    co_store_tile(tiles + 2048, acc.r11, acc.i11, lane);
  }
  co_store_scales<TWO>(acc, n, lane, sc);
}

// S'(a, b) from the upper tiles (conjugated for a > b)
template <bool TWO>
__device__ __forceinline__ float2 co_tile_get(const float2* tiles, int a, int b) {
  const bool sw = a > b;
  const int lo = sw ? b : a, hi = sw ? a : b;
  int t = 0;
  if (TWO) t = (lo >= 32) ? 2 : (hi >= 32 ? 1 : 0);
  const float2 v = tiles[t * 1024 + (lo & 31) * 32 + (hi & 31)];
  return sw ? make_float2(v.x, -v.y) : v;
}

// coherence of pair (a, b): exactly the value the coherence kernels store
template <bool TWO>
struct CohTiles {
  const float2* tiles;
  const float* rs;  // 1/sqrt(S'_mm)
  __device__ __forceinline__ float2 operator()(int a, int b) const {
    const float2 v = co_tile_get<TWO>(tiles, a, b);
    const float f = rs[a] * rs[b];
    return make_float2(v.x * f, v.y * f);
  }
};

// upper tiles of the NB x NB block grid in row-major order: t -> (a, b), a <= b
template <int NB>
__device__ __forceinline__ void cow_tile_ab(int t, int& a, int& b) {
  a = 0;
#pragma unroll
  for (int i = 0; i + 1 < NB; ++i) {
    const bool past = a == i && t >= NB - i;  // beyond row i (NB - i tiles)
    t = past ? t - (NB - i) : t;
    a += past ? 1 : 0;
  }
  b = a + t;
}

// list | scales (32 NB) | 1/sqrt(diag) (32 NB) | one staged tile per wave | K
__host__ __device__ inline size_t cow_lds_bytes(int win, int nb, int waves) {
  return (((size_t)win * 4 + 15) & ~(size_t)15) + (size_t)2 * 32 * nb * 4 + (size_t)waves * 32 * 32 * 8 + 16;
}

// One point on the block's W waves: the per-image scales over its K rows,
// the upper tiles of S' accumulated on MFMA (tile t = wave + W s), 1/sqrt of
// the diagonal, and the coherence row written to out (triu order; through the
// wave's LDS tile for a pair list other than the canonical one).  Rows are
// base[(SAMP ? k : list[k]) * n + image], k < K.  LDS: sc / rs 32 NB floats
// each, stage W 32x32 complex tiles.  Every thread of the block calls it; it
// leaves LDS reusable only after the caller's next barrier.  LOW: a wave loads
// one tile's operands at a time instead of all its tiles' at once (fewer
// registers live, for callers whose register budget is set by a later phase;
// same arithmetic).  (Prefetching a chunk ahead was measured: it costs more
// registers than it hides.)
template <bool SAMP, int NB, int W, bool LOW = false>
__device__ __forceinline__ void cow_point(const float2* __restrict__ base, const int* list, int K, int n,
                                          const int32_t* __restrict__ pairs, int npairs, bool canon, float* sc,
                                          float* rs, float2* stage, float2* __restrict__ out) {
  const int wv = threadIdx.x >> 6;
  int lane = threadIdx.x & 63;
  __asm__ volatile("" : "+v"(lane));
  const int hf = lane >> 5, c32 = lane & 31;
  // ---- per-image power-of-two scale: wave w, image blocks w, w + W, ... -------
  for (int ib = wv; ib < NB; ib += W) {
    const int img = 32 * ib + c32;
    float mx = 0.f;
    if (img < n) {
      // eight independent loads in flight per lane (the two half-waves take
      // alternate SHPs); a short last trip re-reads SHP K-1, which leaves
      // the maximum unchanged
      for (int k0 = 0; k0 < K; k0 += 16) {
        float2 z[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          const int k = k0 + 2 * u + hf < K ? k0 + 2 * u + hf : K - 1;
          z[u] = base[(size_t)(SAMP ? k : list[k]) * n + img];
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) mx = fmaxf(mx, fmaxf(fabsf(z[u].x), fabsf(z[u].y)));
      }
    }
    mx = fmaxf(mx, __shfl_xor(mx, 32, 64));
    float s = 1.f;
    if (mx > 0.f && mx < INFINITY) {
      int e;
      frexpf(mx, &e);
      s = ldexpf(1.f, 6 - e);  // scaled maximum in [32, 64)
    }
    if (hf == 0) sc[img] = s;
  }
  __syncthreads();
  // ---- accumulate this wave's tiles ------------------------------------------
  constexpr int NT = NB * (NB + 1) / 2;
  constexpr int NS = (NT + W - 1) / W;  // tile slots per wave
  int ta[NS], tb[NS];
  bool on[NS];
  float sa[NS], sb[NS];
#pragma unroll
  for (int s = 0; s < NS; ++s) {
    const int t = wv + W * s;
    on[s] = t < NT;
    cow_tile_ab<NB>(on[s] ? t : 0, ta[s], tb[s]);
    sa[s] = sc[32 * ta[s] + c32];
    sb[s] = sc[32 * tb[s] + c32];
  }
  f16x_t ar[NS] = {}, ai[NS] = {};
  const int nch = (K + 7) >> 3;
  // operand rows of tile slot s for chunk c
  auto load1 = [&](int c, int s, float2(&xa)[4], float2(&xb)[4]) {
    const int ra = 32 * ta[s] + c32, rb = 32 * tb[s] + c32;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int k = c * 8 + hf * 4 + q;
      const bool okk = k < K;
      const size_t off = okk ? (size_t)(SAMP ? k : list[k]) * n : 0;
      xa[q] = (okk && ra < n) ? base[off + ra] : make_float2(0.f, 0.f);
      xb[q] = (okk && rb < n) ? base[off + rb] : make_float2(0.f, 0.f);
    }
  };
  float2 xa[NS][4], xb[NS][4];
  for (int c = 0; c < nch; ++c) {
    if (!LOW) {
#pragma unroll
      for (int s = 0; s < NS; ++s)
        if (on[s]) load1(c, s, xa[s], xb[s]);
    }
#pragma unroll
    for (int s = 0; s < NS; ++s) {
      if (!on[s]) continue;
      if (LOW) load1(c, s, xa[s], xb[s]);
      float v[8];
      half8_t x0h, x0l, y0h, y0l, x1h, x1l, y1h, y1l;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        v[2 * q] = xa[s][q].x * sa[s];
        v[2 * q + 1] = xa[s][q].y * sa[s];
      }
      split4c(v, x0h, x0l, y0h, y0l);
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        v[2 * q] = xb[s][q].x * sb[s];
        v[2 * q + 1] = xb[s][q].y * sb[s];
      }
      split4c(v, x1h, x1l, y1h, y1l);
      ar[s] = mfma3(x0h, x0l, x1h, x1l, ar[s]);
      ai[s] = mfma3(y0h, y0l, x1h, x1l, ai[s]);
      if (LOW) __builtin_amdgcn_sched_barrier(0);
    }
  }
  // ---- 1/sqrt(S'_mm) from the diagonal tiles -----------------------------------
#pragma unroll
  for (int s = 0; s < NS; ++s) {
    if (!on[s] || ta[s] != tb[s]) continue;
    // select the diagonal entry this lane owns, then one reciprocal square root
    bool own = false;
    float d = 0.f;
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      const bool hit = 8 * (q >> 2) + 4 * hf + (q & 3) == c32;
      own = own || hit;
      d = hit ? ar[s][q] : d;
    }
    const int m = 32 * ta[s] + c32;
    if (own) rs[m] = 1.f / sqrtf(m < n ? d : 0.f);
  }
  __syncthreads();
  // ---- output ------------------------------------------------------------------
  float2* tile = stage + wv * 1024;
#pragma unroll
  for (int s = 0; s < NS; ++s) {
    if (!on[s]) continue;
    if (canon) {
      co_write_tile_direct(ar[s], ai[s], ta[s], tb[s], n, lane, rs, out);
    } else {
      co_store_tile(tile, ar[s], ai[s], lane);
      __builtin_amdgcn_wave_barrier();
      for (int p = lane; p < npairs; p += 64) {
        const int a = pairs[2 * p], b = pairs[2 * p + 1];
        const int lo = min(a, b), hi = max(a, b);
        if ((lo >> 5) != ta[s] || (hi >> 5) != tb[s]) continue;
        float2 v = tile[(lo & 31) * 32 + (hi & 31)];
        if (a > b) v.y = -v.y;
        const float f = rs[a] * rs[b];
        out[p] = make_float2(v.x * f, v.y * f);
      }
      __builtin_amdgcn_wave_barrier();
    }
  }
}

}  // namespace mr
