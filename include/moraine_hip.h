/*
 * moraine_hip.h -- C ABI of the MI355X (gfx950) implementation of moraine's
 * distributed-scatterer hot path:  ks_test -> select_shp -> emperical_co[_pc]
 * -> emi -> temporal coherence.
 *
 * Conventions (all entry points):
 *   - every array pointer is DEVICE memory, C-contiguous, caller-owned; the
 *     library never allocates device memory except where a workspace query
 *     says so (the caller passes that workspace in);
 *   - complex64 arrays are interleaved (re, im) float pairs, passed as float*;
 *   - `stream` is a hipStream_t (NULL = legacy default stream); every call is
 *     asynchronous on that stream;
 *   - return value: MR_OK (0) or an MR_ERR_* code; mr_last_error_string()
 *     gives a message (thread-local);
 *   - kernels never fail silently: invalid cells are written as NaN exactly
 *     where the reference writes NaN.
 * Reference citations are paths in moraine v0.8.5.
 */
#ifndef MORAINE_HIP_H
#define MORAINE_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MR_OK 0
#define MR_ERR_ARG 1         /* bad shape / argument */
#define MR_ERR_HIP 2         /* HIP runtime error */
#define MR_ERR_UNSUPPORTED 3 /* size outside what this build supports */

typedef void* mr_stream_t; /* hipStream_t */

const char* mr_last_error_string(void);
int mr_abi_version(void);

/* ---------------------------------------------------------------------------
 * KS test + fused SHP selection.
 * Replaces moraine.ks_test (moraine/shp.py:269-305: _sort_numba/cp.sort :117-125,
 * :290; _ks_test[_no_dist]_kernel :130-266) and moraine.select_shp
 * (moraine/shp.py:308-323).
 *   rmli      (nl, w, n) float32 | float64 amplitude/intensity stack (unsorted)
 *   p_out     (nl, w, 2az+1, 2r+1) same dtype, or NULL        (ks_test p)
 *   dist_out  same shape/dtype, or NULL                          (return_dist)
 *   mask_out  same shape, uint8 (0/1), or NULL                   (is_shp)
 *   shp_num   (nl, w) int32, or NULL                             (shp_num)
 *   p_max     SHP threshold used for mask_out/shp_num (is_shp = p < p_max)
 *   workspace >= mr_ks_workspace_bytes(nl, w, n, az, r, sizeof(element)) bytes
 *             of device memory (sorted keys, NaN flags, walk results)
 * n must be 1..MR_KS_MAX_N.  p and dist are bit-exact with the reference CPU
 * path: p = dtype(ks_p_f64(lambda(k))) with k = N*D the integer merge-walk step
 * count.
 * ------------------------------------------------------------------------- */
#define MR_KS_MAX_N 256
size_t mr_ks_workspace_bytes(int nl, int w, int n, int az, int r, int elem_bytes);
int mr_ks_test_f32(const float* rmli, int nl, int w, int n, int az, int r, double p_max,
                   float* p_out, float* dist_out, uint8_t* mask_out, int32_t* shp_num,
                   void* workspace, size_t workspace_bytes, mr_stream_t stream);
int mr_ks_test_f64(const double* rmli, int nl, int w, int n, int az, int r, double p_max,
                   double* p_out, double* dist_out, uint8_t* mask_out, int32_t* shp_num,
                   void* workspace, size_t workspace_bytes, mr_stream_t stream);

/* ks_test + select_shp with no device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same arguments, outputs and bit-exact results as mr_ks_test_f32, but the
 * sorted-key workspace is the library's static device scratch of
 * MR_KS_NOALLOC_SCRATCH_BYTES per GPU (allocated when the library's code is
 * loaded on the GPU, held until the process ends): a stack needs
 * mr_ks_workspace_bytes(nl, w, n, az, r, 4) bytes of it, and a larger one is
 * rejected with MR_ERR_UNSUPPORTED before any device work.  Nothing is
 * allocated and nothing needs clearing: every call writes the scratch before
 * reading it.  All calls on one GPU share its scratch, so calls that overlap on
 * one GPU (two streams, two host threads) race; serialise them. */
#define MR_KS_NOALLOC_SCRATCH_BYTES ((size_t)128 << 20)
int mr_ks_test_f32_noalloc(const float* rmli, int nl, int w, int n, int az, int r, double p_max,
                           float* p_out, float* dist_out, uint8_t* mask_out, int32_t* shp_num,
                           mr_stream_t stream);

/* rslc2amp (moraine/calamp.py:33-40): amp[i] = |rslc[i]|, complex64 (n
 * interleaved re/im float pairs) -> float32, rounded as the C library's hypotf
 * (glibc: the FP64 square sum's correctly rounded sqrt, rounded to FP32),
 * which the reference's CPU path calls (moraine/calamp.py:17-25). */
int mr_rslc2amp(const float* rslc, long long n, float* amp, mr_stream_t stream);

/* amp_disp (moraine/ps.py:17-72): amplitude dispersion index per pixel of a
 * complex64 stack (nl, w, n): std(|z|) / mean(|z|) over the n images
 * (population std), float32 (nl, w).  |z| as mr_rslc2amp. */
int mr_amp_disp(const float* rslc, int nl, int w, int n, float* out, mr_stream_t stream);

/* ks_test on the amplitudes of a complex64 stack (nl, w, n): the CLI's
 * rslc2amp then ks_test (moraine/cli/shp.py:98-101; moraine/shp.py:269-305)
 * with the amplitude computed while the stacks are staged, so the amplitude
 * array is never written.  Equals mr_ks_test_f32 on mr_rslc2amp's output bit
 * for bit; outputs, workspace and limits as mr_ks_test_f32. */
int mr_ks_test_rslc_f32(const float* rslc, int nl, int w, int n, int az, int r, double p_max,
                        float* p_out, float* dist_out, uint8_t* mask_out, int32_t* shp_num,
                        void* workspace, size_t workspace_bytes, mr_stream_t stream);

/* Host-only (no GPU): the k -> (p, dist, is_shp) tables ks_test uses, for
 * k = 0..n (n+1 entries each; shp as 0/1 bytes).  Exposed so the tables can be
 * checked against the reference series without a device. */
int mr_ks_tables_f32(int n, double p_max, float* p, float* dist, uint8_t* shp);

/* select_shp on an existing p array (moraine/shp.py:308-323): is_shp = p < p_max
 * (NaN -> 0), shp_num = count per pixel.  p: (npix, win) float32. */
int mr_select_shp_f32(const float* p, long long npix, int win, double p_max, uint8_t* is_shp,
                      int32_t* shp_num, mr_stream_t stream);

/* select_shp on a float64 p array (moraine/shp.py:308-323): the comparison
 * p < p_max is made in float64. */
int mr_select_shp_f64(const double* p, long long npix, int win, double p_max, uint8_t* is_shp,
                      int32_t* shp_num, mr_stream_t stream);

/* ---------------------------------------------------------------------------
 * Coherence estimation.
 * mr_emperical_co replaces moraine.emperical_co (moraine/co.py:117-137,
 * kernel :61-114): per pixel of rows [row0, row0+nrows) and all (m, j):
 *   cov[m,j] = S[m,j] / n_shp,  coh[m,j] = S[m,j] / sqrt(P[m] P[j]),
 *   S = sum over in-bounds SHPs of z_m conj(z_j), P[m] = sum |z_m|^2.
 *   rslc (nl, w, n) complex64;  is_shp (nl, w, aw, rw) uint8 (full raster);
 *   cov, coh: (nrows, w, n, n) complex64 (cov may be NULL).  Half windows are
 *   (aw-1)/2, (rw-1)/2 as in the reference.
 * mr_emperical_co_pc replaces moraine.emperical_co_pc (moraine/co.py:230-266,
 * kernel :141-190, numba :193-227):
 *   idx (n_pc, 2) int32 (az, r); pc_is_shp (n_pc, aw, rw) uint8;
 *   pairs (n_pairs, 2) int32 (ref < sec); coh (n_pc, n_pairs) complex64.
 * ------------------------------------------------------------------------- */
int mr_emperical_co(const float* rslc, const uint8_t* is_shp, int nl, int w, int n, int aw, int rw,
                    int row0, int nrows, float* cov, float* coh, mr_stream_t stream);
int mr_emperical_co_pc(const float* rslc, int nl, int w, int n, const int32_t* idx,
                       const uint8_t* pc_is_shp, int n_pc, int aw, int rw, const int32_t* pairs,
                       int n_pairs, float* coh, mr_stream_t stream);

/* complex128 forms of the two (moraine/co.py:204: coh has rslc's dtype):
 * rslc, cov, coh complex128 (interleaved double pairs); FP64 accumulation
 * on the VALU kernel; n <= 1024; other arguments as above. */
int mr_emperical_co_c128(const double* rslc, const uint8_t* is_shp, int nl, int w, int n, int aw, int rw,
                         int row0, int nrows, double* cov, double* coh, mr_stream_t stream);
int mr_emperical_co_pc_c128(const double* rslc, int nl, int w, int n, const int32_t* idx,
                            const uint8_t* pc_is_shp, int n_pc, int aw, int rw, const int32_t* pairs,
                            int n_pairs, double* coh, mr_stream_t stream);

/* ---------------------------------------------------------------------------
 * EMI phase linking.  Replaces moraine.emi (moraine/pl.py:68-77; _emi_cp
 * :18-42; _emi_numba :45-65).
 *   coh (n_pts, n(n-1)/2) complex64 in triu order (all pairs);
 *   ph (n_pts, n) complex64; quality (n_pts,) float32 (smallest eigenvalue).
 * Gamma = uncompress(coh); M = inv(|Gamma|) o Gamma; v = eigenvector of the
 * smallest eigenvalue of M; ph = v*conj(v[ref]) / |v*conj(v[ref])|.
 * inv(|Gamma|) is formed in FP64; the eigen-solve precision is stated in the
 * implementation notes of emi.hip.  Singular |Gamma| gives NaN for that point.
 * 32 < n <= 64: a shift-invert pass, a Householder pass over the points it
 * declines, and an FP64 re-solve of those whose estimated phase error exceeds
 * 1e-3 rad; once the stream has passed the call, the workspace's first two
 * 32-bit words hold the number of points of the Householder pass and of the
 * re-solve.  workspace >= mr_emi_workspace_bytes(n_pts, n).
 * ------------------------------------------------------------------------- */
size_t mr_emi_workspace_bytes(long long n_pts, int n);
int mr_emi(const float* coh, long long n_pts, int n, int ref, float* ph, float* quality,
           void* workspace, size_t workspace_bytes, mr_stream_t stream);

/* complex128 EMI (moraine/pl.py:45-65 with coh complex128: ph has coh's dtype,
 * quality float32): every stage in FP64 (|Gamma| inverse by the sweep
 * operator, Householder tridiagonalisation, eigenpair of T to FP64
 * resolution), one 256-thread block per point; n >= 2.
 * workspace >= mr_emi_c128_workspace_bytes(n_pts, n) (0 for n <= 64). */
size_t mr_emi_c128_workspace_bytes(long long n_pts, int n);
int mr_emi_c128(const double* coh, long long n_pts, int n, int ref, double* ph, float* quality,
                void* workspace, size_t workspace_bytes, mr_stream_t stream);

/* ad_intf_pc (moraine/co.py:302-406): adaptively multilooked interferogram of
 * one image pair per point: over the SHPs of the point's aw x rw window that
 * lie inside the nl x w raster, sum ref conj(sec) / sqrt(sum|ref|^2 sum|sec|^2)
 * (FP64 sums), complex64 (n_pc,); NaN without SHPs.  ref_rslc, sec_rslc:
 * complex64 (nl, w); idx (n_pc, 2) int32; pc_is_shp (n_pc, aw, rw) bytes. */
int mr_ad_intf_pc(const float* ref_rslc, const float* sec_rslc, int nl, int w, const int32_t* idx,
                  const uint8_t* pc_is_shp, long long n_pc, int aw, int rw, float* intf, mr_stream_t stream);

int mr_ad_intf_pc_c128(const double* ref_rslc, const double* sec_rslc, int nl, int w, const int32_t* idx,
                       const uint8_t* pc_is_shp, long long n_pc, int aw, int rw, double* intf,
                       mr_stream_t stream);

/* regularize_spectral (moraine/co.py:466-477): out = (1 - beta) coh + beta I
 * for n_mat complex64 n x n matrices; beta (device, FP64) holds one value, or
 * one per matrix when beta_per_matrix != 0; FP64 arithmetic, complex64 out
 * (out may equal coh). */
int mr_regularize_spectral(const float* coh, long long n_mat, int n, const double* beta, int beta_per_matrix,
                           float* out, mr_stream_t stream);

int mr_regularize_spectral_c128(const double* coh, long long n_mat, int n, const double* beta,
                                int beta_per_matrix, double* out, mr_stream_t stream);

/* isPD (moraine/co.py:409-414) and nearestPD (moraine/co.py:429-463) for a
 * stack of n_mat real n x n matrices (float32 or float64, row-major),
 * 1 <= n <= 1024.  isPD: is_pd[i] = 1 when the Cholesky factorisation of
 * matrix i succeeds (FP64, every pivot positive and finite), as the
 * reference's CuPy path reports it.  nearestPD: Higham's method as the
 * reference writes it, its undefined `wherePD` read as isPD: B = (A + A^T)/2,
 * H = V S V^T from the SVD of A, A3 = sym((B + H)/2), then while A3 (rounded
 * to the output dtype) is not PD and k < 100: A3 += (-lambda_min(A3) k^2 +
 * spacing) I, spacing = the gap from ||A||_F to the next number of the dtype.
 * FP64 throughout (one-sided Jacobi SVD), output in the input's dtype (out may
 * not alias a).  workspace >= mr_pd_workspace_bytes(n_mat, n). */
size_t mr_pd_workspace_bytes(long long n_mat, int n);
int mr_is_pd_f32(const float* a, long long n_mat, int n, uint8_t* is_pd, void* workspace,
                 size_t workspace_bytes, mr_stream_t stream);
int mr_is_pd_f64(const double* a, long long n_mat, int n, uint8_t* is_pd, void* workspace,
                 size_t workspace_bytes, mr_stream_t stream);
int mr_nearest_pd_f32(const float* a, long long n_mat, int n, float* out, void* workspace,
                      size_t workspace_bytes, mr_stream_t stream);
int mr_nearest_pd_f64(const double* a, long long n_mat, int n, double* out, void* workspace,
                      size_t workspace_bytes, mr_stream_t stream);

/* uncompress_coh (moraine/co.py:281-299): (n_pts, n_pairs) -> (n_pts, n, n),
 * [a,b] = c, [b,a] = conj(c), diagonal 1, other entries 0, resolved in the
 * reference's order when entries collide: all pairs, then all conjugates, a
 * later pair over an earlier one, the diagonal last.  Pair indices must lie in
 * [0, n) (others are skipped); n <= 16384, n_pairs <= 2^30. */
int mr_uncompress_coh(const float* coh, long long n_pts, int n, const int32_t* pairs, int n_pairs,
                      float* out, mr_stream_t stream);

int mr_uncompress_coh_c128(const double* coh, long long n_pts, int n, const int32_t* pairs, int n_pairs,
                           double* out, mr_stream_t stream);

/* ---------------------------------------------------------------------------
 * Temporal coherence.
 * mr_ds_temp_coh replaces moraine.ds_temp_coh (moraine/pl.py:126-145, numba
 * :80-100, kernel :103-123): t = |sum_p coh_p/|coh_p| conj(ph_a) ph_b| / n_pairs.
 * mr_temp_coh replaces moraine.temp_coh (moraine/pqm.py:124-159, numba :46-71,
 * kernel :99-121): t = |sum_p intf_p conj(r_a) r_b / (|intf_p||r_a||r_b|)| / n_pairs.
 * Inputs are never modified.
 * ------------------------------------------------------------------------- */
int mr_ds_temp_coh(const float* coh, const float* ph, long long n_pts, int n, const int32_t* pairs,
                   int n_pairs, float* t_coh, mr_stream_t stream);
int mr_temp_coh(const float* intf, const float* rslc, long long n_pts, int n, const int32_t* pairs,
                int n_pairs, float* t_coh, mr_stream_t stream);

/* complex128 inputs (moraine/pl.py:103-123: the accumulator has the input's
 * complex type): sums in FP64, t_coh float32; n <= 2048. */
int mr_ds_temp_coh_c128(const double* coh, const double* ph, long long n_pts, int n, const int32_t* pairs,
                        int n_pairs, float* t_coh, mr_stream_t stream);
int mr_temp_coh_c128(const double* intf, const double* rslc, long long n_pts, int n, const int32_t* pairs,
                     int n_pairs, float* t_coh, mr_stream_t stream);

/* EMI followed by the DS temporal coherence of its own phases: the inner two
 * stages of moraine.emperical_co_emi_temp_coh_pc (moraine/pl.py:148-168),
 * emi (moraine/pl.py:68-77) then ds_temp_coh (moraine/pl.py:126-145) on the
 * same coherence rows.  Results equal mr_emi followed by mr_ds_temp_coh bit
 * for bit.  n_pairs must be n(n-1)/2 (EMI reads full rows in triu order;
 * pairs names the image pair of each row entry, as for mr_ds_temp_coh).  For
 * 32 < n <= 64 one kernel does both, reading each row from HBM once.
 * workspace >= mr_emi_workspace_bytes(n_pts, n). */
int mr_emi_ds_temp_coh(const float* coh, long long n_pts, int n, int ref, const int32_t* pairs, int n_pairs,
                       float* ph, float* quality, float* t_coh, void* workspace, size_t workspace_bytes,
                       mr_stream_t stream);

/* One batch of moraine.emperical_co_emi_temp_coh_pc (moraine/pl.py:148-168):
 * emperical_co_pc on the full pair network (moraine/co.py:230-266), emi
 * (moraine/pl.py:68-77) and ds_temp_coh (moraine/pl.py:126-145) for n_pc
 * points.  Arguments as mr_emperical_co_pc (complex64 rslc (nl, w, n), idx
 * (n_pc, 2) int32, pc_is_shp (n_pc, aw, rw) bytes) and mr_emi_ds_temp_coh
 * (ph (n_pc, n) complex64, quality and t_coh (n_pc,) float32).  Results equal
 * mr_emperical_co_pc -> mr_emi_ds_temp_coh bit for bit.  For 32 < n <= 64 one
 * kernel computes each point's coherence into on-chip memory and runs the
 * shift-invert EMI pass and the temporal coherence from there; only the rows
 * of the points that pass leaves to the Householder pass are written to the
 * workspace.  Other n run the two calls through the workspace.
 * workspace >= mr_co_emi_ds_temp_coh_pc_workspace_bytes(n_pc, n): the EMI
 * lists (as mr_emi_workspace_bytes), the pair network, and one coherence row
 * per point. */
size_t mr_co_emi_ds_temp_coh_pc_workspace_bytes(long long n_pc, int n);
int mr_co_emi_ds_temp_coh_pc(const float* rslc, int nl, int w, int n, const int32_t* idx, const uint8_t* pc_is_shp,
                             long long n_pc, int aw, int rw, int ref, float* ph, float* quality, float* t_coh,
                             void* workspace, size_t workspace_bytes, mr_stream_t stream);

/* Precision form of emi / emi + ds_temp_coh for 2 <= n <= 200 (above, mr_emi
 * is FP64 throughout), no reference counterpart (moraine/pl.py:18-77 computes
 * in the array library's precision): the same solve as mr_emi with the checks
 * at tol (radians).  Householder solves (every point for n <= 32 and
 * n > 64; those the shift-invert pass declines for 32 < n <= 64) are flagged
 * when their estimated phase error eps32 ||T|| n / (gap min_i |v_i| 60 sqrt 60)
 * exceeds tol (for 32 < n <= 64: min(tol, 1e-3); T the tridiagonal form, gap
 * its two smallest eigenvalues' distance, v the unit eigenvector).
 * Shift-invert solves (32 < n <= 64) form the FP64 residual of the iterate
 * against inv(|Gamma|) o Gamma (inv in FP64) and estimate the phase error from
 * the FP32 inverse applied to it; above tol the iterate is corrected in FP64
 * (iterative refinement), and points that do not converge are flagged.
 * Flagged points are solved again in FP64 throughout.  With t_coh non-null
 * the DS temporal coherence is written as mr_emi_ds_temp_coh does (pairs: the
 * full triu network).  Points neither corrected nor flagged get exactly
 * mr_emi's outputs.
 * workspace >= mr_emi_refined_workspace_bytes(n_pts, n); once the stream has
 * passed the call, its first 4 bytes hold the number of re-solved points. */
size_t mr_emi_refined_workspace_bytes(long long n_pts, int n);
int mr_emi_refined(const float* coh, long long n_pts, int n, int ref, double tol, const int32_t* pairs,
                   int n_pairs, float* ph, float* quality, float* t_coh, void* workspace, size_t workspace_bytes,
                   mr_stream_t stream);

/* ---------------------------------------------------------------------------
 * Batched DS form.  No single reference function: it is the batch driver
 * moraine.emperical_co_emi_temp_coh_pc (moraine/pl.py:148-168, i.e.
 * emperical_co_pc -> emi -> ds_temp_coh) applied to independent pixels that
 * each carry their own sample matrix instead of a window of a raster.
 *   samples    complex64 rows; pixel id's K = n_samples[id] rows of n images
 *              start at complex element row_off[id] (row-major K x n);
 *   pix        (n_pix,) pixel ids of this call, all with the same n;
 *   pairs      the full triu network (n_pairs = n(n-1)/2);
 *   ph         complex64, pixel id's n phases at complex element ph_off[id];
 *   quality, t_coh  float32 indexed by pixel id.
 * blocks >= 0 runs the fused launches for n <= 200 (coherence never goes
 * through an HBM buffer of the chunk; `blocks` sets the grid, 0 = default):
 * 32 < n <= 64 one wave per pixel with the coherence in LDS tiles, a
 * shift-invert pass, a Householder pass over the pixels it declines and an
 * FP64 re-solve of those the conditioning test flags, as in mr_emi; n <= 32
 * one wave per pixel, coherence in an LDS tile; 64 < n <= 200 one workgroup
 * per pixel whose coherence row goes to that workgroup's slot of the
 * workspace (re-used for each of its pixels).  Workspace >=
 * mr_ds_batch_fused_workspace_bytes(n_pix, n, blocks): the pixel lists for
 * 32 < n <= 64 (256 + 8 bytes per pixel), min(n_pix, blocks or 2048) slots of
 * n(n-1)/2 complex64 for 64 < n <= 200, nothing for n <= 32.
 * n > 200, or blocks < 0 (per-stage mode: coherence, EMI and temporal
 * coherence as separate launches through HBM, the drop-in sequence), runs per
 * chunk of `chunk` pixels with workspace >= mr_ds_batch_workspace_bytes(chunk, n).
 * coh_waves: the tile layout of the per-stage coherence kernel for
 * 64 < n <= 224 (its grid of 32 x 32 tiles spread over that many waves per
 * workgroup: 0 = the drop-in's layout, 3 waves up to n = 96, 4 up to n = 128 and
 * 8 above; 2 or 3 for n <= 96, 4 or 8); the outputs do not depend on it, nor on
 * the mode or grid.
 * For a sample matrix made of a pixel's SHP stacks in window order the outputs
 * equal the drop-in composition exactly.
 * ------------------------------------------------------------------------- */
size_t mr_ds_batch_workspace_bytes(long long chunk, int n);
size_t mr_ds_batch_fused_workspace_bytes(long long n_pix, int n, int blocks);
int mr_ds_batch(const float* samples, const long long* row_off, const int32_t* n_samples,
                const long long* ph_off, const int32_t* pix, long long n_pix, int n, int ref,
                const int32_t* pairs, int n_pairs, float* ph, float* quality, float* t_coh, int blocks,
                long long chunk, int coh_waves, void* workspace, size_t workspace_bytes, mr_stream_t stream);

#ifdef __cplusplus
}
#endif
#endif /* MORAINE_HIP_H */
