// isPD and nearestPD for gfx950.
//
// Replaces moraine.isPD (moraine/co.py:409-414: Cholesky, finite factor ->
// PD) and moraine.nearestPD (moraine/co.py:429-463: Higham 1988, with the
// reference's undefined `wherePD` read as isPD).  Per matrix A (real n x n):
//   B = (A + A^T)/2;  A = U S V^T;  H = V S V^T;  A3 = sym((B + H)/2);
//   while A3 is not PD and k < 100:  k += 1;  A3 += (-lambda_min(A3) k^2 + spacing) I
// with spacing the gap from ||A||_F to the next number of the input's dtype.
// One 256-thread block per matrix, everything FP64 in a per-block workspace
// slice (the caller's):
//   * singular values and right singular vectors by one-sided (Hestenes)
//     Jacobi: columns of A are rotated pairwise until orthogonal, the
//     rotations accumulated in V; the n/2 disjoint pairs of a round-robin round
//     are spread over the block's four waves (one pair per wave at a time,
//     lanes over rows, FP64 wave sums);
//   * lambda_min of the symmetric A3 as sigma_min(A3 + c I) - c with
//     c = ||A3||_F (A3 + c I is PSD, so its singular values are its
//     eigenvalues), by the same Jacobi;
//   * the PD test is a right-looking Cholesky of the matrix as it will be
//     returned (rounded to the output dtype), every pivot positive and finite.
#include "common.hpp"

namespace mr {

constexpr int PD_THREADS = 256;
constexpr int PD_WAVES = PD_THREADS / 64;
constexpr int PD_MAX_N = 1024;
constexpr int PD_BLOCKS = 1024;

// workspace per block: AT, V, W (n x n each), s (n)
__host__ __device__ inline size_t pd_slice_doubles(int n) { return 3 * (size_t)n * n + (size_t)n; }

// One-sided Jacobi on the columns of a matrix stored transposed (row j of CT is
// column j), V^T accumulated in VT (initialised by the caller).  On return the
// columns are mutually orthogonal: C V = U S with S = column norms.
__device__ void pd_jacobi(double* CT, double* VT, int n, int* sflag) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int m = n + (n & 1);  // players of the round-robin (one dummy when n is odd)
  for (int sweep = 0; sweep < 60; ++sweep) {
    if (threadIdx.x == 0) *sflag = 0;
    __syncthreads();
    for (int rd = 0; rd < m - 1; ++rd) {
      for (int pi = wv; pi < m / 2; pi += PD_WAVES) {
        // circle method: player m-1 fixed, pair 0 is (rd, m-1), pair j is (rd + j, rd - j) mod (m-1)
        int p, q;
        if (pi == 0) {
          p = rd;
          q = m - 1;
        } else {
          p = (rd + pi) % (m - 1);
          q = (rd - pi + (m - 1)) % (m - 1);
        }
        if (p > q) {
          const int t = p;
          p = q;
          q = t;
        }
        if (q >= n) continue;  // the dummy player
        double* cp = CT + (size_t)p * n;
        double* cq = CT + (size_t)q * n;
        double al = 0.0, be = 0.0, ga = 0.0;
        for (int i = lane; i < n; i += 64) {
          const double x = cp[i], y = cq[i];
          al = fma(x, x, al);
          be = fma(y, y, be);
          ga = fma(x, y, ga);
        }
        al = wave_sum(al);
        be = wave_sum(be);
        ga = wave_sum(ga);
        // converged pair: |cos| of the angle between the columns below n ulp
        if (ga == 0.0 || fabs(ga) <= 2.2e-16 * (double)n * sqrt(al * be)) continue;
        const double ze = (be - al) / (2.0 * ga);
        const double t = (ze >= 0.0 ? 1.0 : -1.0) / (fabs(ze) + sqrt(1.0 + ze * ze));
        const double c = 1.0 / sqrt(1.0 + t * t), s = c * t;
        double* vp = VT + (size_t)p * n;
        double* vq = VT + (size_t)q * n;
        for (int i = lane; i < n; i += 64) {
          const double x = cp[i], y = cq[i];
          cp[i] = c * x - s * y;
          cq[i] = s * x + c * y;
          const double u = vp[i], w = vq[i];
          vp[i] = c * u - s * w;
          vq[i] = s * u + c * w;
        }
        if (lane == 0) *sflag = 1;
      }
      __syncthreads();
    }
    if (*sflag == 0) break;
    __syncthreads();
  }
}

__device__ double pd_block_sum(double v, double* red) {
  v = wave_sum(v);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  double s = 0.0;
  for (int i = 0; i < PD_WAVES; ++i) s += red[i];
  __syncthreads();
  return s;
}

// Cholesky test of X (row-major, symmetric; lower triangle used) in W:
// every pivot positive and finite.  F32: on the float32-rounded values.
template <bool F32>
__device__ bool pd_cholesky_ok(const double* X, double* W, int n, int* sflag) {
  for (int t = threadIdx.x; t < n * n; t += PD_THREADS) W[t] = F32 ? (double)(float)X[t] : X[t];
  if (threadIdx.x == 0) *sflag = 1;
  __syncthreads();
  for (int k = 0; k < n; ++k) {
    const double d = W[(size_t)k * n + k];
    const bool ok = d > 0.0 && isfinite(d);
    __syncthreads();
    if (!ok) {
      if (threadIdx.x == 0) *sflag = 0;
      __syncthreads();
      return false;
    }
    const double rs = 1.0 / sqrt(d);
    // column k of L, then the trailing lower triangle
    for (int i = k + 1 + threadIdx.x; i < n; i += PD_THREADS) W[(size_t)i * n + k] *= rs;
    __syncthreads();
    const int mm = n - k - 1;
    for (int t = threadIdx.x; t < mm * mm; t += PD_THREADS) {
      const int i = k + 1 + t / mm, j = k + 1 + t % mm;
      if (j <= i) W[(size_t)i * n + j] -= W[(size_t)i * n + k] * W[(size_t)j * n + k];
    }
    __syncthreads();
  }
  return true;
}

// lambda_min of the symmetric X = sigma_min(X + c I) - c, c = ||X||_F (singular
// values of the PSD X + c I are its eigenvalues).
__device__ double pd_lambda_min(const double* X, double* CT, double* VT, double* sv, int n, double* red, int* sflag) {
  double f = 0.0;
  for (int t = threadIdx.x; t < n * n; t += PD_THREADS) f = fma(X[t], X[t], f);
  const double c = sqrt(pd_block_sum(f, red));
  for (int t = threadIdx.x; t < n * n; t += PD_THREADS) {
    const int i = t / n, j = t % n;
    CT[t] = X[(size_t)j * n + i] + (i == j ? c : 0.0);
    VT[t] = i == j ? 1.0 : 0.0;
  }
  __syncthreads();
  pd_jacobi(CT, VT, n, sflag);
  double mn = 1e308;
  for (int j = threadIdx.x; j < n; j += PD_THREADS) {
    double s2 = 0.0;
    for (int i = 0; i < n; ++i) s2 = fma(CT[(size_t)j * n + i], CT[(size_t)j * n + i], s2);
    mn = fmin(mn, sqrt(s2));
  }
  // block min
  for (int o = 32; o > 0; o >>= 1) mn = fmin(mn, __shfl_xor(mn, o, 64));
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = mn;
  __syncthreads();
  double r = red[0];
  for (int i = 1; i < PD_WAVES; ++i) r = fmin(r, red[i]);
  __syncthreads();
  return r - c;
}

template <typename T, bool NEAREST>
__global__ __launch_bounds__(PD_THREADS) void pd_kernel(const T* __restrict__ a, long long nmat, int n,
                                                       double* __restrict__ ws, T* __restrict__ out,
                                                       uint8_t* __restrict__ is_pd) {
  __shared__ double red[PD_WAVES];
  __shared__ int sflag;
  const size_t nn = (size_t)n * n;
  double* CT = ws + (size_t)blockIdx.x * pd_slice_doubles(n);
  double* VT = CT + nn;
  double* W = VT + nn;
  double* sv = W + nn;
  constexpr bool F32 = sizeof(T) == 4;
  for (long long mi = blockIdx.x; mi < nmat; mi += gridDim.x) {
    const T* A = a + (size_t)mi * nn;
    if (!NEAREST) {
      for (int t = threadIdx.x; t < (int)nn; t += PD_THREADS) CT[t] = (double)A[t];
      __syncthreads();
      const bool ok = pd_cholesky_ok<false>(CT, W, n, &sflag);
      if (threadIdx.x == 0) is_pd[mi] = ok ? 1 : 0;
      __syncthreads();
      continue;
    }
    // ||A||_F in the input's dtype and its spacing
    double f = 0.0;
    for (int t = threadIdx.x; t < (int)nn; t += PD_THREADS) f = fma((double)A[t], (double)A[t], f);
    const double nrm64 = sqrt(pd_block_sum(f, red));
    double spacing;
    if (F32) {
      const float nf = (float)nrm64;
      spacing = (double)(nextafterf(nf, __builtin_inff()) - nf);
    } else {
      spacing = nextafter(nrm64, __builtin_inf()) - nrm64;
    }
    // SVD of A: columns of A are rows of CT
    for (int t = threadIdx.x; t < (int)nn; t += PD_THREADS) {
      const int i = t / n, j = t % n;
      CT[t] = (double)A[(size_t)j * n + i];
      VT[t] = i == j ? 1.0 : 0.0;
    }
    __syncthreads();
    pd_jacobi(CT, VT, n, &sflag);
    for (int j = threadIdx.x; j < n; j += PD_THREADS) {
      double s2 = 0.0;
      for (int i = 0; i < n; ++i) s2 = fma(CT[(size_t)j * n + i], CT[(size_t)j * n + i], s2);
      sv[j] = sqrt(s2);
    }
    __syncthreads();
    // W = A3 = sym((B + H) / 2), H = V S V^T, B = (A + A^T) / 2
    for (int t = threadIdx.x; t < (int)nn; t += PD_THREADS) {
      const int i = t / n, k = t % n;
      double h = 0.0;
      for (int j = 0; j < n; ++j) h = fma(VT[(size_t)j * n + i] * sv[j], VT[(size_t)j * n + k], h);
      const double b = ((double)A[(size_t)i * n + k] + (double)A[(size_t)k * n + i]) / 2.0;
      CT[t] = (b + h) / 2.0;  // A2
    }
    __syncthreads();
    for (int t = threadIdx.x; t < (int)nn; t += PD_THREADS) {
      const int i = t / n, k = t % n;
      W[t] = (CT[t] + CT[(size_t)k * n + i]) / 2.0;
    }
    __syncthreads();
    // A3 moves to CT's place only through the loop below: W holds it, the
    // Cholesky test uses VT, lambda_min uses CT and VT (sv reused as scratch)
    for (int k = 0;; ++k) {
      const bool ok = pd_cholesky_ok<F32>(W, VT, n, &sflag);
      if (ok || k >= 100) break;
      const double lmin = pd_lambda_min(W, CT, VT, sv, n, red, &sflag);
      const double sh = -lmin * (double)((k + 1) * (k + 1)) + spacing;
      for (int i = threadIdx.x; i < n; i += PD_THREADS) W[(size_t)i * n + i] += sh;
      __syncthreads();
    }
    for (int t = threadIdx.x; t < (int)nn; t += PD_THREADS) out[(size_t)mi * nn + t] = (T)W[t];
    __syncthreads();
  }
}

// blocks of a launch: at most PD_BLOCKS, and workspace slices within 1 GiB
static long long pd_blocks(long long nmat, int n) {
  const long long by_mem = std::max<long long>(1, (1ll << 30) / (long long)(pd_slice_doubles(n) * 8));
  return std::max<long long>(1, std::min<long long>(std::min<long long>(nmat, PD_BLOCKS), by_mem));
}

template <typename T, bool NEAREST>
static int pd_launch(const T* a, long long nmat, int n, void* ws, size_t ws_bytes, T* out, uint8_t* is_pd,
                     mr_stream_t stream, const char* name) {
  MR_CHECK_ARG(nmat >= 0 && n >= 1 && n <= PD_MAX_N, "%s: bad sizes (n = %d, 1..%d)", name, n, PD_MAX_N);
  if (nmat == 0) return MR_OK;
  MR_CHECK_ARG(a && (NEAREST ? out != nullptr : is_pd != nullptr), "%s: null pointer", name);
  const int nb = (int)pd_blocks(nmat, n);
  const size_t need = (size_t)nb * pd_slice_doubles(n) * 8;
  MR_CHECK_ARG(ws != nullptr && ws_bytes >= need, "%s: workspace %zu < %zu bytes", name, ws_bytes, need);
  hipLaunchKernelGGL((pd_kernel<T, NEAREST>), dim3(nb), dim3(PD_THREADS), 0, as_stream(stream), a, nmat, n,
                     reinterpret_cast<double*>(ws), out, is_pd);
  return launch_status(name);
}

}  // namespace mr

extern "C" {

size_t mr_pd_workspace_bytes(long long n_mat, int n) {
  if (n_mat <= 0 || n < 1) return 0;
  return (size_t)mr::pd_blocks(n_mat, n) * mr::pd_slice_doubles(n) * 8;
}

int mr_is_pd_f32(const float* a, long long n_mat, int n, uint8_t* is_pd, void* workspace, size_t workspace_bytes,
                 mr_stream_t stream) {
  return mr::pd_launch<float, false>(a, n_mat, n, workspace, workspace_bytes, nullptr, is_pd, stream, "isPD");
}
int mr_is_pd_f64(const double* a, long long n_mat, int n, uint8_t* is_pd, void* workspace, size_t workspace_bytes,
                 mr_stream_t stream) {
  return mr::pd_launch<double, false>(a, n_mat, n, workspace, workspace_bytes, nullptr, is_pd, stream, "isPD");
}
int mr_nearest_pd_f32(const float* a, long long n_mat, int n, float* out, void* workspace, size_t workspace_bytes,
                      mr_stream_t stream) {
  return mr::pd_launch<float, true>(a, n_mat, n, workspace, workspace_bytes, out, nullptr, stream, "nearestPD");
}
int mr_nearest_pd_f64(const double* a, long long n_mat, int n, double* out, void* workspace, size_t workspace_bytes,
                      mr_stream_t stream) {
  return mr::pd_launch<double, true>(a, n_mat, n, workspace, workspace_bytes, out, nullptr, stream, "nearestPD");
}

}  // extern "C"
