/*
 * C/OpenMP restatement of moraine's numba CPU kernels -- TEST INFRASTRUCTURE ONLY.
 *
 * It is the parity checker for large cases and the timed CPU baseline ("port"):
 * the reference CPU path itself is numba, which is not available.  Each
 * function follows the loop order and accumulator widths of the numba code it
 * cites.  Built with -ffp-contract=off so float products round like numba's
 * (which never contracts into FMA).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

/* _ks_p_numba (moraine/shp.py:19-32), literal. */
static double ks_p(double x) {
  double x2 = -2 * x * x;
  double p = 0.0, p2 = 0.0;
  int sign = 1;
  for (int i = 1; i <= 100; i++) {
    p += sign * 2 * exp(x2 * i * i);
    if (p == p2) return p;
    sign = -sign;
    p2 = p;
  }
  return p;
}

/* np.sort order (NaN last), insertion sort: stacks are short. */
static void sort_nan_last(float* v, int n) {
  for (int i = 1; i < n; i++) {
    float x = v[i];
    int j = i - 1;
    if (isnan(x)) continue; /* NaN stays behind every number moved after it */
    while (j >= 0 && (isnan(v[j]) || v[j] > x)) {
      v[j + 1] = v[j];
      j--;
    }
    v[j + 1] = x;
  }
}

/* _ks_d_numba (moraine/shp.py:35-44) returning the integer step count k = N*D. */
static int ks_k(const float* ref, const float* sec, int n) {
  int j1 = 0, j2 = 0, kmax = 0;
  while (j1 < n && j2 < n) {
    float f1 = ref[j1], f2 = sec[j2];
    if (f1 <= f2) j1++;
    if (f1 >= f2) j2++;
    int d = j2 - j1;
    if (d < 0) d = -d;
    if (d > kmax) kmax = d;
  }
  return kmax;
}

/* _sort_numba + _ks_test_numba (moraine/shp.py:80-125).  dist may be NULL.
 * k_out (int32, may be NULL) receives k or -1 where NaN is written.
 * per_pair = 0: p depends only on k, so the series is evaluated once per k;
 * per_pair = 1: once per pixel pair, as the reference does (moraine/shp.py:118,
 * up to 100 exp per pair) -- the same values at the reference's cost, for the
 * timed CPU baseline. */
void port_ks_test_f32(const float* rmli, int nl, int w, int n, int az, int r,
                      float* p, float* dist, int32_t* k_out, int nthreads, int per_pair) {
  size_t npx = (size_t)nl * w;
  float* s = (float*)malloc(npx * n * sizeof(float));
  const int aw = 2 * az + 1, rw = 2 * r + 1, win = aw * rw;
  const double en = sqrt(n / 2.0);
  /* p depends only on k: evaluate the series once per k, same operation order */
  double* pk = (double*)malloc((n + 1) * sizeof(double));
  for (int k = 0; k <= n; k++) pk[k] = ks_p((en + 0.12 + 0.11 / en) * fabs((double)k / n));
#pragma omp parallel for schedule(static) num_threads(nthreads)
  for (long long i = 0; i < (long long)npx; i++) {
    memcpy(s + i * n, rmli + i * n, n * sizeof(float));
    sort_nan_last(s + i * n, n);
  }
#pragma omp parallel for schedule(dynamic, 4) num_threads(nthreads)
  for (int i = 0; i < nl; i++) {
    for (int j = 0; j < w; j++) {
      const float* ref = s + ((size_t)i * w + j) * n;
      size_t o0 = ((size_t)i * w + j) * win;
      int ref_nan = isnan(ref[n - 1]);
      for (int l = -az; l <= az; l++) {
        for (int m = -r; m <= r; m++) {
          size_t o = o0 + (size_t)(l + az) * rw + (m + r);
          int si = i + l, sj = j + m;
          if (ref_nan || si < 0 || si >= nl || sj < 0 || sj >= w ||
              isnan(s[((size_t)si * w + sj) * n + n - 1])) {
            p[o] = NAN;
            if (dist) dist[o] = NAN;
            if (k_out) k_out[o] = -1;
            continue;
          }
          int k = ks_k(ref, s + ((size_t)si * w + sj) * n, n);
          p[o] = per_pair ? (float)ks_p((en + 0.12 + 0.11 / en) * fabs((double)k / n)) : (float)pk[k];
          if (dist) dist[o] = (float)fabs((double)k / n);
          if (k_out) k_out[o] = k;
        }
      }
    }
  }
  free(pk);
  free(s);
}

/* _emperical_co_pc_numba (moraine/co.py:193-227): complex64 products into a
 * complex128 accumulator, float32 squares into float64 power sums. */
void port_emperical_co_pc(const float* rslc /* (nl,w,n) complex64 interleaved */, int nl, int w,
                          int n, const int32_t* idx /* (n_pc,2) */, const uint8_t* pc_is_shp,
                          int n_pc, int aw, int rw, const int32_t* pairs, int npairs,
                          float* coh /* (n_pc,npairs) complex64 */, int nthreads) {
  const int az = aw / 2, r = rw / 2;
#pragma omp parallel for schedule(dynamic, 16) num_threads(nthreads)
  for (int i = 0; i < n_pc; i++) {
    for (int pi = 0; pi < npairs; pi++) {
      int m = pairs[2 * pi], jj = pairs[2 * pi + 1];
      double nr = 0.0, ni = 0.0, am = 0.0, aj = 0.0;
      for (int k = 0; k < aw; k++) {
        for (int l = 0; l < rw; l++) {
          int a = idx[2 * i] - az + k, b = idx[2 * i + 1] - r + l;
          if (a >= 0 && a < nl && b >= 0 && b < w && pc_is_shp[(size_t)i * aw * rw + k * rw + l]) {
            const float* zm = rslc + (((size_t)a * w + b) * n + m) * 2;
            const float* zj = rslc + (((size_t)a * w + b) * n + jj) * 2;
            float sm = zm[0] * zm[0] + zm[1] * zm[1];
            float sj = zj[0] * zj[0] + zj[1] * zj[1];
            am += sm;
            aj += sj;
            float pr = zm[0] * zj[0] + zm[1] * zj[1];
            float pim = zm[1] * zj[0] - zm[0] * zj[1];
            nr += pr;
            ni += pim;
          }
        }
      }
      double den = sqrt(am * aj);
      coh[((size_t)i * npairs + pi) * 2] = (float)(nr / den);
      coh[((size_t)i * npairs + pi) * 2 + 1] = (float)(ni / den);
    }
  }
}

/* _ds_temp_coh_numba (moraine/pl.py:80-100): complex64 accumulation of
 * coh*(conj(ph_n)*ph_k)/|coh|. */
void port_ds_temp_coh(const float* coh, const float* ph, int n_pts, int n, const int32_t* pairs,
                      int npairs, float* out, int nthreads) {
#pragma omp parallel for schedule(static) num_threads(nthreads)
  for (int i = 0; i < n_pts; i++) {
    float tr = 0.f, ti = 0.f;
    for (int j = 0; j < npairs; j++) {
      int a = pairs[2 * j], b = pairs[2 * j + 1];
      float pnr = ph[((size_t)i * n + a) * 2], pni = -ph[((size_t)i * n + a) * 2 + 1];
      float pkr = ph[((size_t)i * n + b) * 2], pki = ph[((size_t)i * n + b) * 2 + 1];
      float ir = pnr * pkr - pni * pki, ii = pnr * pki + pni * pkr;
      float cr = coh[((size_t)i * npairs + j) * 2], ci = coh[((size_t)i * npairs + j) * 2 + 1];
      float dr = cr * ir - ci * ii, di = cr * ii + ci * ir;
      float ac = hypotf(cr, ci);
      tr += dr / ac;
      ti += di / ac;
    }
    out[i] = (float)((double)hypotf(tr, ti) / npairs);
  }
}

/* _temp_coh_pc_numba (moraine/pqm.py:46-71) without mutating the inputs. */
void port_temp_coh(const float* intf, const float* rslc, int n_pts, int n, const int32_t* pairs,
                   int npairs, float* out, int nthreads) {
#pragma omp parallel for schedule(static) num_threads(nthreads)
  for (int i = 0; i < n_pts; i++) {
    float tr = 0.f, ti = 0.f;
    for (int j = 0; j < npairs; j++) {
      int a = pairs[2 * j], b = pairs[2 * j + 1];
      const float* ra = rslc + ((size_t)i * n + a) * 2;
      const float* rb = rslc + ((size_t)i * n + b) * 2;
      const float* f = intf + ((size_t)i * npairs + j) * 2;
      float na = hypotf(ra[0], ra[1]), nb = hypotf(rb[0], rb[1]), nf = hypotf(f[0], f[1]);
      float ar = ra[0] / na, ai = -ra[1] / na, br = rb[0] / nb, bi = rb[1] / nb;
      float fr = f[0] / nf, fi = f[1] / nf;
      float xr = ar * br - ai * bi, xi = ar * bi + ai * br;
      tr += fr * xr - fi * xi;
      ti += fr * xi + fi * xr;
    }
    out[i] = (float)((double)hypotf(tr, ti) / npairs);
  }
}
