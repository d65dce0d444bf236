// KS test + fused SHP selection for gfx950.
//
// Replaces moraine.ks_test (moraine/shp.py:269-305) and moraine.select_shp
// (moraine/shp.py:308-323).  Design:
//   * ks_sort_kernel: one wave per pixel stack; values become order-preserving
//     unsigned keys (-0 folded onto +0 so key order == IEEE '<=' order for
//     non-NaN values), bitonic-sorted in registers with cross-lane shuffles,
//     written once to a workspace with a per-pixel "has NaN" flag (the
//     reference's "last sorted element is NaN" test).
//   * ks_pairs_kernel: one 256-thread block per (azimuth line, column tile).
//     The tile's reference stacks stay in LDS (the sorted keys turned back
//     into values, a NaN after each stack); for each window row offset the
//     secondary line segment (tile + range halo) is staged in LDS and every
//     (pixel, range offset) pair runs the merge walk of moraine/shp.py:35-44
//     with the reference's float compares; at a NaN neither compare holds, so
//     a walk stops where moraine's loop does with no index bound per step.
//     The walk yields the integer step count k = N*D; p,
//     dist and the SHP bit are table lookups of k, so they are bit-exact with
//     the reference CPU path (p = dtype(f64 series), dist = dtype(k/N)).
//     Several walks are interleaved per thread (ILP) to hide LDS latency.
//     The walk is symmetric in its two stacks, so only the window offsets
//     with l > 0, or l = 0 and m >= 0, are walked (half the window plus the
//     pixel itself); each walk stores its step count k (one byte for n < 255)
//     in a workspace.
//   * ks_emit_kernel: one block per (line, column tile) writes each pixel's
//     whole p, dist and mask rows and its SHP count, reading its own walks
//     and, mirrored, those of the neighbours that walked its other offsets
//     (no atomics, full-line writes).
#include <mutex>

#include "common.hpp"

namespace mr {

template <typename T>
struct KsTables;
template <>
struct KsTables<float> {
  float p[MR_KS_MAX_N + 1];
  float d[MR_KS_MAX_N + 1];
  uint32_t shp[(MR_KS_MAX_N + 32) / 32];
};
constexpr int KS_MAX_N_F64 = 224;  // keeps the f64 table inside the 4 KiB kernel-argument limit
template <>
struct KsTables<double> {
  double p[KS_MAX_N_F64 + 1];
  double d[KS_MAX_N_F64 + 1];
  uint32_t shp[(KS_MAX_N_F64 + 32) / 32];
};

// Series of moraine/shp.py:19-32 evaluated in the same operation order (f64,
// glibc exp) so that the stored dtype(p) matches the reference bit for bit.
#pragma clang fp contract(off)
static double ks_p_series(double x) {
  double x2 = -2 * x * x;
  double p = 0.0, p2 = 0.0;
  int sign = 1;
  for (int i = 1; i <= 100; i++) {
    p += sign * 2 * std::exp(x2 * i * i);
    if (p == p2) return p;
    sign = -sign;
    p2 = p;
  }
  return p;
}

template <typename T>
static void build_tables(int n, double p_max, KsTables<T>& t) {
  const double en = std::sqrt(n / 2.0);  // moraine/shp.py:76  en = sqrt(n_imag/2)
  for (int w = 0; w < (int)(sizeof(t.shp) / 4); ++w) t.shp[w] = 0;
  for (int k = 0; k <= n; ++k) {
    const double dmax = std::fabs((double)k / (double)n);  // abs((j2-j1)/n_imag)
    const T pk = (T)ks_p_series((en + 0.12 + 0.11 / en) * dmax);
    t.p[k] = pk;
    t.d[k] = (T)dmax;
    if ((double)pk < p_max) t.shp[k >> 5] |= 1u << (k & 31);  // select_shp: p < p_max
  }
}
#pragma clang fp contract(on)

// |z| of a complex64 sample as moraine's CPU path computes it (rslc2amp,
// moraine/calamp.py:17-25: numba's abs of a complex64 calls the C library's
// hypotf, which in glibc is this): both squares and their sum in FP64 (the
// squares are exact, the sum rounds once), a correctly rounded FP64 sqrt, one
// rounding to FP32; an infinite part gives +inf even beside a NaN.
#pragma clang fp contract(off)
__device__ __forceinline__ float rslc_amp(float2 z) {
  if (__builtin_isinf(z.x) || __builtin_isinf(z.y)) return __builtin_inff();
  const double x = z.x, y = z.y;
  const double s = x * x + y * y;
  return (float)__builtin_sqrt(s);
}
#pragma clang fp contract(on)

// rslc2amp: one float per complex sample, two per thread per trip
__global__ __launch_bounds__(256) void rslc2amp_kernel(const float2* __restrict__ z, long long n,
                                                       float* __restrict__ amp) {
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += 2 * stride) {
    const float2 a = z[i];
    const bool two = i + stride < n;
    const float2 b = two ? z[i + stride] : make_float2(0.f, 0.f);
    amp[i] = rslc_amp(a);
    if (two) amp[i + stride] = rslc_amp(b);
  }
}

// amp_disp (moraine/ps.py:17-29): per pixel, the amplitudes a_k = |z_k| of its
// n images, D = std(a) / mean(a) with the population std.  One wave per pixel,
// the sums in FP64 (the reference sums in FP32; its tests hold to 1e-6).  For
// n <= 64 a wave takes AD_U pixels at a time, one amplitude per lane held in a
// register, so the AD_U pixels' wave sums overlap.
constexpr int AD_U = 4;
__global__ __launch_bounds__(256) void amp_disp_kernel(const float2* __restrict__ z, long long npx, int n,
                                                       float* __restrict__ out) {
  const int lane = threadIdx.x & 63;
  const long long wave0 = ((long long)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const long long nw = ((long long)gridDim.x * blockDim.x) >> 6;
  if (n <= 64) {
    for (long long p0 = wave0 * AD_U; p0 < npx; p0 += nw * AD_U) {
      float a[AD_U];
      double s[AD_U];
#pragma unroll
      for (int u = 0; u < AD_U; ++u) {
        const long long px = p0 + u;
        a[u] = (px < npx && lane < n) ? rslc_amp(z[px * n + lane]) : 0.f;
      }
#pragma unroll
      for (int u = 0; u < AD_U; ++u) s[u] = wave_sum((double)a[u]);
#pragma unroll
      for (int u = 0; u < AD_U; ++u) {
        const double mean = s[u] / (double)n;
        const double d = lane < n ? (double)a[u] - mean : 0.0;
        const double sd = sqrt(wave_sum(d * d) / (double)n);
        if (lane == 0 && p0 + u < npx) out[p0 + u] = (float)(sd / mean);
      }
    }
    return;
  }
  for (long long px = wave0; px < npx; px += nw) {
    const float2* zp = z + px * n;
    double s = 0.0;
    for (int i = lane; i < n; i += 64) s += (double)rslc_amp(zp[i]);
    const double mean = wave_sum(s) / (double)n;
    double q = 0.0;
    for (int i = lane; i < n; i += 64) {
      const double d = (double)rslc_amp(zp[i]) - mean;
      q += d * d;
    }
    const double sd = sqrt(wave_sum(q) / (double)n);
    if (lane == 0) out[px] = (float)(sd / mean);
  }
}

template <typename T>
struct KeyOf;
template <>
struct KeyOf<float> {
  using type = uint32_t;
  __device__ static uint32_t key(float x) {
    uint32_t b = __float_as_uint(x);
    if (b == 0x80000000u) b = 0u;  // -0 == +0 under the reference's float compares
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
  }
  // the value a key stands for (-0 comes back as +0, which compares the same)
  __device__ static float value(uint32_t k) {
    return __uint_as_float((k & 0x80000000u) ? (k & 0x7fffffffu) : ~k);
  }
};
template <>
struct KeyOf<double> {
  using type = uint64_t;
  __device__ static uint64_t key(double x) {
    uint64_t b = (uint64_t)__double_as_longlong(x);
    if (b == 0x8000000000000000ull) b = 0ull;
    return (b & 0x8000000000000000ull) ? ~b : (b | 0x8000000000000000ull);
  }
  __device__ static double value(uint64_t k) {
    return __longlong_as_double((long long)((k & 0x8000000000000000ull) ? (k & 0x7fffffffffffffffull) : ~k));
  }
};

// ---------------------------------------------------------------------------
// sort: one wave per pixel, NP = 64*E >= n keys in registers (index e*64+lane)
// ---------------------------------------------------------------------------
// CPLX (T = float): the stacks are complex64 samples, sorted by rslc_amp.
template <typename T, int E, bool CPLX = false>
__global__ __launch_bounds__(256) void ks_sort_kernel(const T* __restrict__ rmli, long long npx, int n,
                                                      typename KeyOf<T>::type* __restrict__ keys,
                                                      uint8_t* __restrict__ valid,
                                                      const float2* __restrict__ rslc = nullptr) {
  using Key = typename KeyOf<T>::type;
  constexpr int NP = 64 * E;
  const int lane = threadIdx.x & 63;
  const long long wave0 = ((long long)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const long long nwaves = ((long long)gridDim.x * blockDim.x) >> 6;
  for (long long px = wave0; px < npx; px += nwaves) {
    Key v[E];
    bool has_nan = false;
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int i = e * 64 + lane;
      if (i < n) {
        T x;
        if constexpr (CPLX)
          x = rslc_amp(rslc[px * n + i]);
        else
          x = rmli[px * n + i];
        has_nan |= (x != x);
        v[e] = KeyOf<T>::key(x);
      } else {
        v[e] = ~(Key)0;
      }
    }
#pragma unroll
    for (int k = 2; k <= NP; k <<= 1) {
#pragma unroll
      for (int j = k >> 1; j > 0; j >>= 1) {
        Key nv[E];  // all partners are read before any register is overwritten
#pragma unroll
        for (int e = 0; e < E; ++e) {
          const int i = e * 64 + lane;
          Key pv;
          if (j >= 64)
            pv = v[e ^ (j >> 6)];
          else
            pv = __shfl_xor(v[e], j, 64);
          const bool up = (i & k) == 0;
          const bool lower = (i & j) == 0;
          const Key lo = v[e] < pv ? v[e] : pv;
          const Key hi = v[e] < pv ? pv : v[e];
          nv[e] = (lower == up) ? lo : hi;
        }
#pragma unroll
        for (int e = 0; e < E; ++e) v[e] = nv[e];
      }
    }
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int i = e * 64 + lane;
      if (i < n) keys[px * n + i] = v[e];
    }
    const bool any_nan = __any(has_nan);
    if (lane == 0) valid[px] = any_nan ? 0 : 1;
  }
}

// ---------------------------------------------------------------------------
// pairs
// ---------------------------------------------------------------------------
constexpr int KS_THREADS = 256;
// walks interleaved per thread: 4 or 6, picked per launch (ks_pick_ilp)

template <typename T>
__device__ __forceinline__ void stage_stacks(T* __restrict__ dst, const typename KeyOf<T>::type* __restrict__ src,
                                             int nstk, int n, int ns) {
  using Key = typename KeyOf<T>::type;
  const int tot = nstk * n;
  // eight loads per thread are in flight before their LDS stores (a plain
  // copy loop waits for every load before issuing the next)
  for (int i0 = 0; i0 < tot; i0 += 8 * KS_THREADS) {
    Key v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int i = i0 + u * KS_THREADS + threadIdx.x;
      v[u] = i < tot ? src[i] : (Key)0;
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int i = i0 + u * KS_THREADS + threadIdx.x;
      if (i < tot) {
        const int s = i / n;
        dst[s * ns + (i - s * n)] = KeyOf<T>::value(v[u]);
      }
    }
  }
  // sentinel: NaN compares neither <= nor >= anything, so a walk that reaches
  // the end of either stack stops there (as moraine's loop does)
  for (int s = threadIdx.x; s < nstk; s += KS_THREADS) dst[s * ns + n] = (T)NAN;
}

// Walk results: the integer step count k = N*D of each walked (pixel, offset)
// pair, one element per entry of the half window the pair kernel walks, in
// the layout kbuf[line][l][column][m] (l = 0..az row offsets below the pixel,
// m = 0..rw-1 range offsets; for l = 0 only m >= r is walked).  KB is uint8
// for n <= 254 and uint16 above; its all-ones value marks a pair the
// reference writes as NaN (a NaN stack or a partner outside the raster).
template <typename KB>
__host__ __device__ constexpr KB ks_invalid() {
  return (KB)~(KB)0;
}

// XCD-aware tile order: the hardware hands workgroup i to XCD i % 8.  The
// (tile, line) units are ordered tile-column-major (every line of tile 0, then
// of tile 1, ...) and XCD x runs the x-th of eight equal contiguous chunks of
// that order, so all XCDs get the same number of full-width tiles and each
// walks down its tile columns line by line: the stacks a block stages (its
// line and the az lines below, tile plus range halo) were staged by the
// blocks of the lines just above on the same XCD and sit in its L2.  Returns
// false for the padding workgroups of the last chunk.
constexpr int KS_XCDS = 8;
__device__ __forceinline__ bool ks_tile_of(int ntx, int nl, long long chunk, int& tile, int& line) {
  const int x = blockIdx.x % KS_XCDS;
  const long long j = blockIdx.x / KS_XCDS;
  const long long f = x * chunk + j;
  if (j >= chunk || f >= (long long)ntx * nl) return false;
  tile = (int)(f / nl);
  line = (int)(f - (long long)tile * nl);
  return true;
}

static_assert(MR_KS_MAX_N < (1 << 16), "the walk's |ib - ia| is one v_sad_u16 of indices below 2^16");

template <typename T, int KS_ILP, typename KB>
__global__ __launch_bounds__(KS_THREADS) void ks_pairs_kernel(const typename KeyOf<T>::type* __restrict__ keys,
                                                              const uint8_t* __restrict__ valid, int nl, int w,
                                                              int n, int az, int r, int tw, int ntx,
                                                              long long chunk, KB* __restrict__ kbuf) {
  extern __shared__ __align__(16) unsigned char smem[];
  int tile, a;
  if (!ks_tile_of(ntx, nl, chunk, tile, a)) return;
  const int c0 = tile * tw;
  const int tcols = min(tw, w - c0);
  const int ns = n + 1;
  const int rw = 2 * r + 1;
  const int scols = tcols + 2 * r;
  T* ref = reinterpret_cast<T*>(smem);  // sorted stacks as values, NaN after each
  T* sec = ref + tw * ns;
  uint8_t* refv = reinterpret_cast<uint8_t*>(sec + (tw + 2 * r) * ns);
  uint8_t* secv = refv + tw;

  stage_stacks(ref, keys + ((size_t)a * w + c0) * n, tcols, n, ns);
  for (int c = threadIdx.x; c < tw; c += KS_THREADS) refv[c] = c < tcols ? valid[(size_t)a * w + c0 + c] : 0;
  // The KS distance of two pixels does not depend on which one is the
  // reference (the walk is symmetric), so each unordered pair is walked once:
  // offsets (l, m - r) with l > 0, or l = 0 and m >= r (m = r is the pixel
  // itself).  The emit kernel reads the result for both pixels.
  for (int l = 0; l <= az; ++l) {
    __syncthreads();
    const int sa = a + l;
    const bool row_ok = sa < nl;
    if (row_ok) {
      const int g0 = max(0, c0 - r);
      const int g1 = min(w, c0 + tcols + r);
      const int s0 = g0 - (c0 - r);
      stage_stacks(sec + s0 * ns, keys + ((size_t)sa * w + g0) * n, g1 - g0, n, ns);
      for (int s = threadIdx.x; s < scols; s += KS_THREADS) {
        const int gc = c0 - r + s;
        secv[s] = (gc >= 0 && gc < w) ? valid[(size_t)sa * w + gc] : 0;
      }
    } else {
      for (int s = threadIdx.x; s < scols; s += KS_THREADS) secv[s] = 0;
    }
    __syncthreads();
    const int m0 = l == 0 ? r : 0;
    const int mc = rw - m0;
    const int npairs = tcols * mc;
    KB* kb = kbuf + (((size_t)a * (az + 1) + l) * w + c0) * rw;
    for (int t0 = threadIdx.x; t0 < npairs; t0 += KS_THREADS * KS_ILP) {
      int ia[KS_ILP], ib[KS_ILP], ra[KS_ILP], rb[KS_ILP], km[KS_ILP], cc[KS_ILP], mm[KS_ILP];
      bool ok[KS_ILP], in[KS_ILP];
      T ka[KS_ILP], kb_[KS_ILP];
#pragma unroll
      for (int q = 0; q < KS_ILP; ++q) {
        const int t = t0 + q * KS_THREADS;
        in[q] = t < npairs;
        const int c = in[q] ? t / mc : 0;
        const int m = in[q] ? m0 + (t - c * mc) : 0;
        cc[q] = c;
        mm[q] = m;
        ok[q] = in[q] && refv[c] && secv[c + m];
        ra[q] = c * ns;
        rb[q] = ok[q] ? (tw + c + m) * ns : c * ns;  // invalid pairs walk ref vs itself (k=0, discarded)
        ia[q] = 0;
        ib[q] = 0;
        km[q] = 0;
        ka[q] = ref[ra[q]];
        kb_[q] = ref[rb[q]];
      }
      // merge walk of moraine/shp.py:35-44 on the values: '<=' advances ref,
      // '>=' advances sec.  At the NaN after either stack neither compare
      // holds, so the walk stays where moraine's loop ends (the walks of a
      // thread run until the slowest one is done; |ib - ia| can only shrink
      // after either stack is exhausted)
      for (int it = 0; it < 2 * n; it += 4) {
#pragma unroll
        for (int u = 0; u < 4; ++u) {
#pragma unroll
          for (int q = 0; q < KS_ILP; ++q) {
            const int le = ka[q] <= kb_[q];
            const int ge = ka[q] >= kb_[q];
            ia[q] += le;
            ib[q] += ge;
            // opaque: keeps the compiler from turning the index updates into
            // running address sums (more instructions per step, not fewer)
            __asm__ volatile("" : "+v"(ia[q]), "+v"(ib[q]));
            // |ib - ia| in one v_sad_u16 (both in [0, n], n < 2^16: the high halves are zero)
            km[q] = max(km[q], (int)__builtin_amdgcn_sad_u16((unsigned)ib[q], (unsigned)ia[q], 0u));
            ka[q] = ref[ra[q] + ia[q]];
            kb_[q] = ref[rb[q] + ib[q]];
          }
        }
        bool alive = false;
#pragma unroll
        for (int q = 0; q < KS_ILP; ++q) alive |= (ia[q] < n) && (ib[q] < n);
        if (!__any(alive)) break;
      }
#pragma unroll
      for (int q = 0; q < KS_ILP; ++q)
        if (in[q]) kb[cc[q] * rw + mm[q]] = ok[q] ? (KB)km[q] : ks_invalid<KB>();
    }
  }
}

// Emit: one block per (line, column tile) writes its pixels' whole p, dist and
// mask rows (W contiguous entries per pixel, consecutive threads on
// consecutive entries) and SHP counts.  Entry (lw, m) of pixel (a, c):
//   lw >= az (own walk, l = lw - az):  kbuf[a][l][c][m]  for l > 0 or m >= r,
//     kbuf[a][0][c + m - r][rw - 1 - m] for l = 0, m < r (the left neighbour's walk);
//   lw < az (l = az - lw): the pixel (a - l, c + m - r) walked it as its offset
//     (l, rw - 1 - m): kbuf[a - l][l][c + m - r][rw - 1 - m];
// a source outside the raster gives NaN, as the reference writes.
constexpr int KS_EMIT_THREADS = 256;
constexpr int KS_EMIT_TILE = 32;
constexpr int KS_EMIT_U = 4;
constexpr int KS_EMIT_MAX_W = 4096;  // window entries the per-block offset table holds
template <typename T, typename KB>
__global__ __launch_bounds__(KS_EMIT_THREADS) void ks_emit_kernel(const KB* __restrict__ kbuf, int nl, int w, int az,
                                                                  int r, KsTables<T> tab, T* __restrict__ p_out,
                                                                  T* __restrict__ d_out, uint8_t* __restrict__ m_out,
                                                                  int32_t* __restrict__ shp_num) {
  __shared__ int cnt[KS_EMIT_TILE];
  __shared__ int e_src[KS_EMIT_MAX_W];         // window entry -> ((column shift + 2048) << 12) | sm
  __shared__ uint16_t e_row[KS_EMIT_MAX_W];    // window entry -> l
  __shared__ T t_p[MR_KS_MAX_N + 1], t_d[MR_KS_MAX_N + 1];
  __shared__ uint32_t t_s[(MR_KS_MAX_N + 32) / 32];
  const int a = blockIdx.y;
  const int c0 = blockIdx.x * KS_EMIT_TILE;
  const int tcols = min(KS_EMIT_TILE, w - c0);
  const int rw = 2 * r + 1, aw = 2 * az + 1;
  const int W = aw * rw;
  if (threadIdx.x < KS_EMIT_TILE) cnt[threadIdx.x] = 0;
  constexpr int NT = (int)(sizeof(tab.p) / sizeof(T));
  for (int i = threadIdx.x; i < NT; i += KS_EMIT_THREADS) {
    t_p[i] = tab.p[i];
    t_d[i] = tab.d[i];
  }
  for (int i = threadIdx.x; i < (int)(sizeof(tab.shp) / 4); i += KS_EMIT_THREADS) t_s[i] = tab.shp[i];
  // per window entry (lw, m): the row offset l of the walk that holds it, the
  // column shift of the walking pixel and the walk's range index sm; the line
  // of the walk is a for l = lw - az >= 0, a - l for entries above the pixel
  for (int e = threadIdx.x; e < W; e += KS_EMIT_THREADS) {
    const int lw = e / rw, m = e - lw * rw;
    int l, dc, sm;
    if (lw > az || (lw == az && m >= r)) {  // the pixel's own walk
      l = lw - az;
      dc = 0;
      sm = m;
    } else {  // the left neighbour (same line) or the pixel above walked it, mirrored
      l = az - lw;
      dc = m - r;
      sm = rw - 1 - m;
    }
    e_src[e] = ((dc + 2048) << 12) | sm;
    e_row[e] = (uint16_t)l;
  }
  __syncthreads();
  const size_t plane = (size_t)w * rw;  // one (line, l) plane of kbuf
  const size_t obase = ((size_t)a * w + c0) * W;
  // consecutive threads on consecutive entries of the tile's contiguous rows,
  // KS_EMIT_U entries per thread per trip with their gathers in flight together;
  // (c, e) advance by the block size without divisions
  const int qd = KS_EMIT_THREADS / W, qr = KS_EMIT_THREADS - (KS_EMIT_THREADS / W) * W;
  int c = threadIdx.x / W, e = threadIdx.x - (threadIdx.x / W) * W;
  const int tot = tcols * W;
  for (int t0 = threadIdx.x; t0 < tot; t0 += KS_EMIT_U * KS_EMIT_THREADS) {
    int cc[KS_EMIT_U], kk[KS_EMIT_U];
    KB v[KS_EMIT_U];
#pragma unroll
    for (int u = 0; u < KS_EMIT_U; ++u) {
      const int t = t0 + u * KS_EMIT_THREADS;
      cc[u] = c;
      kk[u] = -1;  // -1: no source (outside the raster or beyond the tile)
      v[u] = ks_invalid<KB>();
      if (t < tot) {
        const int code = e_src[e];
        const int l = e_row[e], dc = (code >> 12) - 2048, sm = code & 4095;
        const bool own = (e >= az * rw + r);  // lw > az, or lw == az and m >= r
        const int line = own || l == 0 ? a : a - l;
        const int src = c0 + c + dc;
        if (line >= 0 && src >= 0 && src < w) {
          kk[u] = 0;
          v[u] = kbuf[((size_t)line * (az + 1) + l) * plane + (size_t)src * rw + sm];
        }
      }
      c += qd;
      e += qr;
      if (e >= W) {
        e -= W;
        ++c;
      }
    }
#pragma unroll
    for (int u = 0; u < KS_EMIT_U; ++u) {
      const int t = t0 + u * KS_EMIT_THREADS;
      if (t >= tot) break;
      const int k = (kk[u] == 0 && v[u] != ks_invalid<KB>()) ? (int)v[u] : -1;
      const bool s = k >= 0 && ((t_s[k >> 5] >> (k & 31)) & 1u);
      if (p_out) p_out[obase + t] = k >= 0 ? t_p[k] : (T)NAN;
      if (d_out) d_out[obase + t] = k >= 0 ? t_d[k] : (T)NAN;
      if (m_out) m_out[obase + t] = s ? 1 : 0;
      if (s && shp_num) atomicAdd(&cnt[cc[u]], 1);
    }
  }
  __syncthreads();
  if (shp_num && threadIdx.x < tcols) shp_num[(size_t)a * w + c0 + threadIdx.x] = cnt[threadIdx.x];
}

// select_shp on an existing p array: one wave per pixel.
template <typename T>
__global__ __launch_bounds__(256) void select_shp_kernel(const T* __restrict__ p, long long npix, int win,
                                                         double p_max, uint8_t* __restrict__ is_shp,
                                                         int32_t* __restrict__ shp_num) {
  const int lane = threadIdx.x & 63;
  const long long wave0 = ((long long)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const long long nw = ((long long)gridDim.x * blockDim.x) >> 6;
  for (long long px = wave0; px < npix; px += nw) {
    int c = 0;
    for (int i = lane; i < win; i += 64) {
      const T v = p[px * win + i];
      const bool s = (double)v < p_max;  // NaN -> false; float64 p compares in float64
      if (is_shp) is_shp[px * win + i] = s;
      c += s;
    }
    c = wave_sum(c);
    if (lane == 0 && shp_num) shp_num[px] = c;
  }
}

static size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

// Walk-result element of n images: one byte up to n = 254 (k <= n, all ones = NaN), two above.
static size_t ks_kb_bytes(int n) { return n <= 254 ? 1 : 2; }

// Workspace of ks_test_impl: sorted keys (key_bytes per sample), one NaN flag
// per pixel, and the walk results ((az + 1) x rw elements per pixel).
static size_t ks_workspace_need(size_t npx, int n, int az, int r, size_t key_bytes) {
  return align256(npx * (size_t)n * key_bytes) + align256(npx) +
         align256(npx * (size_t)(az + 1) * (2 * (size_t)r + 1) * ks_kb_bytes(n));
}

// Walks per thread for a window row of `pairs` (pixel, offset) walks per block:
// a trip runs KS_THREADS x ilp walks and takes about ilp walk-times, so the
// cost of a row is ceil(pairs / (KS_THREADS ilp)) * ilp; ties go to 6 (more
// loads in flight).  E.g. 64 columns x 21 offsets: 4 -> two trips (8), 6 -> one
// trip (6); 64 x 15: 4 -> one trip (4), 6 -> one trip (6).
static int ks_pick_ilp(int pairs) {
  auto cost = [&](int ilp) { return (int)div_up((unsigned long long)pairs, (unsigned long long)KS_THREADS * ilp) * ilp; };
  return cost(6) <= cost(4) ? 6 : 4;
}

// rslc (T = float only): KS on rslc_amp of a complex64 stack instead of rmli
template <typename T>
static int ks_test_impl(const T* rmli, int nl, int w, int n, int az, int r, double p_max, T* p_out, T* d_out,
                        uint8_t* m_out, int32_t* shp_num, void* ws, size_t ws_bytes, mr_stream_t stream_,
                        const float2* rslc = nullptr) {
  using Key = typename KeyOf<T>::type;
  const int maxn = sizeof(T) == 4 ? MR_KS_MAX_N : KS_MAX_N_F64;
  MR_CHECK_ARG((rmli != nullptr || rslc != nullptr) && ws != nullptr, "ks_test: null rmli or workspace");
  MR_CHECK_ARG(nl > 0 && w > 0, "ks_test: empty raster (%d x %d)", nl, w);
  if (n < 1 || n > maxn) {
    set_error("ks_test: n=%d outside 1..%d", n, maxn);
    return MR_ERR_UNSUPPORTED;
  }
  MR_CHECK_ARG(az >= 0 && r >= 0, "ks_test: negative half window");
  const size_t npx = (size_t)nl * w;
  const size_t need = ks_workspace_need(npx, n, az, r, sizeof(Key));
  MR_CHECK_ARG(ws_bytes >= need, "ks_test: workspace %zu < %zu bytes", ws_bytes, need);
  hipStream_t st = as_stream(stream_);
  Key* keys = reinterpret_cast<Key*>(ws);
  uint8_t* valid = reinterpret_cast<uint8_t*>(ws) + align256(npx * n * sizeof(Key));
  void* kbuf = valid + align256(npx);

  const int nb = (int)std::min<size_t>(div_up(npx, 4), 65536);
  if constexpr (sizeof(T) == 4) {
    if (rslc) {
      if (n <= 64)
        hipLaunchKernelGGL((ks_sort_kernel<T, 1, true>), dim3(nb), dim3(256), 0, st, nullptr, (long long)npx, n, keys,
                           valid, rslc);
      else if (n <= 128)
        hipLaunchKernelGGL((ks_sort_kernel<T, 2, true>), dim3(nb), dim3(256), 0, st, nullptr, (long long)npx, n, keys,
                           valid, rslc);
      else
        hipLaunchKernelGGL((ks_sort_kernel<T, 4, true>), dim3(nb), dim3(256), 0, st, nullptr, (long long)npx, n, keys,
                           valid, rslc);
    }
  }
  if (!rslc) {
    if (n <= 64)
      hipLaunchKernelGGL((ks_sort_kernel<T, 1>), dim3(nb), dim3(256), 0, st, rmli, (long long)npx, n, keys, valid);
    else if (n <= 128)
      hipLaunchKernelGGL((ks_sort_kernel<T, 2>), dim3(nb), dim3(256), 0, st, rmli, (long long)npx, n, keys, valid);
    else
      hipLaunchKernelGGL((ks_sort_kernel<T, 4>), dim3(nb), dim3(256), 0, st, rmli, (long long)npx, n, keys, valid);
  }
  int rc = launch_status("ks_sort");
  if (rc) return rc;

  // column tile: as wide as 64 pixels while LDS stays <= 48 KiB (>= 3 blocks/CU)
  const size_t stk = (size_t)(n + 1) * sizeof(Key);
  auto lds_for = [&](int tw) { return (2 * (size_t)tw + 2 * r) * stk + 2 * (size_t)tw + 2 * r + 16; };
  // one trip of four walks per thread covers a window row when tw * rw <= 4 * KS_THREADS: a narrower tile
  // than 64 columns then also leaves LDS for more blocks per CU
  int tw = std::max(1, std::min(64, 4 * KS_THREADS / (2 * r + 1)));
  while (tw > 1 && lds_for(tw) > 48 * 1024) --tw;
  tw = std::min(tw, w);
  const size_t lds = lds_for(tw);
  if (lds > 160 * 1024) {
    set_error("ks_test: window too large for LDS (n=%d, r_half=%d)", n, r);
    return MR_ERR_UNSUPPORTED;
  }
  // pass 1: the walks, XCD-chunked tile order (ks_tile_of); pass 2: whole output rows
  const int ntx = (int)div_up(w, tw);
  const long long chunk = ((long long)ntx * nl + KS_XCDS - 1) / KS_XCDS;  // (tile, line) units per XCD
  const long long nwg = KS_XCDS * chunk;
  if ((2 * az + 1) * (2 * r + 1) > KS_EMIT_MAX_W) {
    set_error("ks_test: window %d x %d above %d entries", 2 * az + 1, 2 * r + 1, KS_EMIT_MAX_W);
    return MR_ERR_UNSUPPORTED;
  }
  if (nwg >= (1ll << 31)) {
    set_error("ks_test: raster too large (%d x %d)", nl, w);
    return MR_ERR_UNSUPPORTED;
  }
  const bool ilp6 = ks_pick_ilp(tw * (2 * r + 1)) == 6;
  const dim3 g1((unsigned)nwg);
  if (ks_kb_bytes(n) == 1) {
    uint8_t* kb = reinterpret_cast<uint8_t*>(kbuf);
    if (ilp6)
      hipLaunchKernelGGL((ks_pairs_kernel<T, 6, uint8_t>), g1, dim3(KS_THREADS), lds, st, keys, valid, nl, w, n, az,
                         r, tw, ntx, chunk, kb);
    else
      hipLaunchKernelGGL((ks_pairs_kernel<T, 4, uint8_t>), g1, dim3(KS_THREADS), lds, st, keys, valid, nl, w, n, az,
                         r, tw, ntx, chunk, kb);
  } else {
    uint16_t* kb = reinterpret_cast<uint16_t*>(kbuf);
    if (ilp6)
      hipLaunchKernelGGL((ks_pairs_kernel<T, 6, uint16_t>), g1, dim3(KS_THREADS), lds, st, keys, valid, nl, w, n,
                         az, r, tw, ntx, chunk, kb);
    else
      hipLaunchKernelGGL((ks_pairs_kernel<T, 4, uint16_t>), g1, dim3(KS_THREADS), lds, st, keys, valid, nl, w, n,
                         az, r, tw, ntx, chunk, kb);
  }
  if (int rc2 = launch_status("ks_pairs")) return rc2;
  KsTables<T> tab;
  build_tables<T>(n, p_max, tab);
  const dim3 g2(div_up(w, KS_EMIT_TILE), nl);
  if (ks_kb_bytes(n) == 1)
    hipLaunchKernelGGL((ks_emit_kernel<T, uint8_t>), g2, dim3(KS_EMIT_THREADS), 0, st,
                       reinterpret_cast<const uint8_t*>(kbuf), nl, w, az, r, tab, p_out, d_out, m_out, shp_num);
  else
    hipLaunchKernelGGL((ks_emit_kernel<T, uint16_t>), g2, dim3(KS_EMIT_THREADS), 0, st,
                       reinterpret_cast<const uint16_t*>(kbuf), nl, w, az, r, tab, p_out, d_out, m_out, shp_num);
  return launch_status("ks_emit");
}

// ---------------------------------------------------------------------------
// Library-owned scratch for mr_ks_test_f32_noalloc: the sorted-key workspace
// of ks_test_impl in a file-scope device array, for callers that cannot
// allocate device memory.  In host code `ks_scratch` names a host-side stand-in,
// not the device array: its device address is looked up once per device with
// hipGetSymbolAddress (with that device current) and only that address is ever
// passed to kernels.  Every call on one GPU shares that GPU's copy.
// ---------------------------------------------------------------------------
static __device__ __align__(256) unsigned char ks_scratch[MR_KS_NOALLOC_SCRATCH_BYTES];

constexpr int KS_SCRATCH_MAX_DEVICES = 64;
static std::mutex ks_scratch_mutex;
static void* ks_scratch_dev[KS_SCRATCH_MAX_DEVICES] = {};

// Device address of ks_scratch on the current device (looked up on the first call per device).
static int ks_scratch_address(void** out) {
  int dev = -1;
  MR_CHECK_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= KS_SCRATCH_MAX_DEVICES) {
    set_error("ks_test_noalloc: device %d outside 0..%d", dev, KS_SCRATCH_MAX_DEVICES - 1);
    return MR_ERR_UNSUPPORTED;
  }
  std::lock_guard<std::mutex> lock(ks_scratch_mutex);
  if (ks_scratch_dev[dev] == nullptr) {
    void* p = nullptr;
    MR_CHECK_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(ks_scratch)));
    if (p == nullptr) {
      set_error("ks_test_noalloc: hipGetSymbolAddress returned a null address on device %d", dev);
      return MR_ERR_HIP;
    }
    ks_scratch_dev[dev] = p;
  }
  *out = ks_scratch_dev[dev];
  return MR_OK;
}

}  // namespace mr

extern "C" {

int mr_ks_test_f32_noalloc(const float* rmli, int nl, int w, int n, int az, int r, double p_max, float* p_out,
                           float* dist_out, uint8_t* mask_out, int32_t* shp_num, mr_stream_t stream) {
  MR_CHECK_ARG(nl > 0 && w > 0 && n >= 1, "ks_test_noalloc: bad sizes (%d x %d x %d)", nl, w, n);
  MR_CHECK_ARG(az >= 0 && r >= 0, "ks_test_noalloc: negative half window");
  const size_t need = mr::ks_workspace_need((size_t)nl * w, n, az, r, sizeof(uint32_t));
  if (need > (size_t)MR_KS_NOALLOC_SCRATCH_BYTES) {
    mr::set_error("ks_test_noalloc: a %d x %d x %d stack needs %zu bytes of scratch, more than the %zu bytes "
                  "of the library's static scratch (use mr_ks_test_f32 with a workspace)",
                  nl, w, n, need, (size_t)MR_KS_NOALLOC_SCRATCH_BYTES);
    return MR_ERR_UNSUPPORTED;
  }
  void* ws = nullptr;
  if (int rc = mr::ks_scratch_address(&ws)) return rc;
  return mr::ks_test_impl<float>(rmli, nl, w, n, az, r, p_max, p_out, dist_out, mask_out, shp_num, ws,
                                 (size_t)MR_KS_NOALLOC_SCRATCH_BYTES, stream);
}

size_t mr_ks_workspace_bytes(int nl, int w, int n, int az, int r, int elem_bytes) {
  if (nl <= 0 || w <= 0 || n < 1 || az < 0 || r < 0) return 0;
  return mr::ks_workspace_need((size_t)nl * w, n, az, r, elem_bytes == 8 ? sizeof(uint64_t) : sizeof(uint32_t));
}

int mr_ks_test_f32(const float* rmli, int nl, int w, int n, int az, int r, double p_max, float* p_out,
                   float* dist_out, uint8_t* mask_out, int32_t* shp_num, void* workspace, size_t workspace_bytes,
                   mr_stream_t stream) {
  return mr::ks_test_impl<float>(rmli, nl, w, n, az, r, p_max, p_out, dist_out, mask_out, shp_num, workspace,
                                 workspace_bytes, stream);
}

int mr_ks_test_f64(const double* rmli, int nl, int w, int n, int az, int r, double p_max, double* p_out,
                   double* dist_out, uint8_t* mask_out, int32_t* shp_num, void* workspace, size_t workspace_bytes,
                   mr_stream_t stream) {
  return mr::ks_test_impl<double>(rmli, nl, w, n, az, r, p_max, p_out, dist_out, mask_out, shp_num, workspace,
                                  workspace_bytes, stream);
}

int mr_ks_test_rslc_f32(const float* rslc, int nl, int w, int n, int az, int r, double p_max, float* p_out,
                        float* dist_out, uint8_t* mask_out, int32_t* shp_num, void* workspace, size_t workspace_bytes,
                        mr_stream_t stream) {
  MR_CHECK_ARG(rslc != nullptr, "ks_test_rslc: null rslc");
  return mr::ks_test_impl<float>(nullptr, nl, w, n, az, r, p_max, p_out, dist_out, mask_out, shp_num, workspace,
                                 workspace_bytes, stream, reinterpret_cast<const float2*>(rslc));
}

int mr_rslc2amp(const float* rslc, long long n, float* amp, mr_stream_t stream) {
  MR_CHECK_ARG(n >= 0, "rslc2amp: negative size");
  if (n == 0) return MR_OK;
  MR_CHECK_ARG(rslc && amp, "rslc2amp: null pointer");
  const int nb = (int)std::min<long long>(mr::div_up((unsigned long long)n, 512), 1 << 16);
  hipLaunchKernelGGL(mr::rslc2amp_kernel, dim3(nb), dim3(256), 0, mr::as_stream(stream),
                     reinterpret_cast<const float2*>(rslc), n, amp);
  return mr::launch_status("rslc2amp");
}

int mr_amp_disp(const float* rslc, int nl, int w, int n, float* out, mr_stream_t stream) {
  MR_CHECK_ARG(nl >= 0 && w >= 0 && n >= 1, "amp_disp: bad sizes (%d x %d x %d)", nl, w, n);
  const long long npx = (long long)nl * w;
  if (npx == 0) return MR_OK;
  MR_CHECK_ARG(rslc && out, "amp_disp: null pointer");
  const int nb = (int)std::min<long long>(mr::div_up((unsigned long long)npx, 4 * mr::AD_U), 1 << 16);
  hipLaunchKernelGGL(mr::amp_disp_kernel, dim3(nb), dim3(256), 0, mr::as_stream(stream),
                     reinterpret_cast<const float2*>(rslc), npx, n, out);
  return mr::launch_status("amp_disp");
}

int mr_ks_tables_f32(int n, double p_max, float* p, float* dist, uint8_t* shp) {
  MR_CHECK_ARG(n >= 1 && n <= MR_KS_MAX_N, "ks_tables: n=%d outside 1..%d", n, MR_KS_MAX_N);
  mr::KsTables<float> t;
  mr::build_tables<float>(n, p_max, t);
  for (int k = 0; k <= n; ++k) {
    if (p) p[k] = t.p[k];
    if (dist) dist[k] = t.d[k];
    if (shp) shp[k] = (t.shp[k >> 5] >> (k & 31)) & 1u;
  }
  return MR_OK;
}

int mr_select_shp_f32(const float* p, long long npix, int win, double p_max, uint8_t* is_shp, int32_t* shp_num,
                      mr_stream_t stream) {
  MR_CHECK_ARG(p != nullptr && npix >= 0 && win > 0, "select_shp: bad arguments");
  if (npix == 0) return MR_OK;
  const int nb = (int)std::min<long long>(mr::div_up(npix, 4), 65536);
  hipLaunchKernelGGL(mr::select_shp_kernel<float>, dim3(nb), dim3(256), 0, mr::as_stream(stream), p, npix, win, p_max,
                     is_shp, shp_num);
  return mr::launch_status("select_shp");
}

int mr_select_shp_f64(const double* p, long long npix, int win, double p_max, uint8_t* is_shp, int32_t* shp_num,
                      mr_stream_t stream) {
  MR_CHECK_ARG(p != nullptr && npix >= 0 && win > 0, "select_shp_f64: bad arguments");
  if (npix == 0) return MR_OK;
  const int nb = (int)std::min<long long>(mr::div_up(npix, 4), 65536);
  hipLaunchKernelGGL(mr::select_shp_kernel<double>, dim3(nb), dim3(256), 0, mr::as_stream(stream), p, npix, win,
                     p_max, is_shp, shp_num);
  return mr::launch_status("select_shp_f64");
}

}  // extern "C"
