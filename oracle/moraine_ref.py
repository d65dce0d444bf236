"""numpy restatement of moraine's DS hot path -- TEST INFRASTRUCTURE ONLY (the oracle).

Every function cites the reference lines it restates (paths relative to the
moraine v0.8.5 source tree).  Two flavours exist where precision matters:

* ``*_ref`` / plain names: the reference's CPU (numba) semantics -- integer KS
  step count with the f64 p series cast to the input dtype, c128 coherence
  accumulators over c64 products, c64 temporal-coherence accumulation, FP32
  LAPACK EMI.
* ``*64``: FP64 "truth" used to state errors of FP32/mixed-precision results.

How each function is pinned (tests/test_oracle.py):

==========================  ===================================================
function                    pinned by
==========================  ===================================================
ks_p                        printed p(N=20,k=7)=0.13494715 and p(0)=0
                            (nbs/API/shp.ipynb:495); scipy.special.kolmogorov
                            (the same alternating series, closed-form library
                            routine) for lambda >= 0.3
ks_k_walk / ks_k_sorted     literal merge walk vs. CDF-difference formulation
                            vs. scipy.stats.ks_2samp statistic*N on random and
                            tie-heavy inputs (nbs/API/shp.ipynb:594)
ks_test                     printed 1x2x20 NaN/OOB pattern and dist=0.35
                            (nbs/API/shp.ipynb:460-499); literal loop on tiny
                            stacks; symmetry p[i,j,l,m] = p[i+l,j+m,-l,-m]
select_shp                  closed form p < p_max and count (shp.ipynb:664)
emperical_co[_pc]           brute-force python loop (co.ipynb:444-470);
                            upper triangle of emperical_co == emperical_co_pc
                            (co.ipynb:807); Cauchy-Schwarz |coh| <= 1, diag = 1
uncompress_coh / pairs      exact arrays printed in nbs/API/tnet.ipynb:245,
                            :299-365 and nimage_from_npair(6) == 4
emi64                       closure-free Gamma: ph = theta*conj(theta_ref)/|.|,
                            quality = 1 (Fiedler); eigen-residual; numpy eigh
emi_ref32                   same closure-free case at FP32 tolerance
ds_temp_coh / temp_coh      t = 1 for consistent inputs; hand-computed 3-image
                            case; FP64 variant agreement
isPD / nearestPD            Cholesky success on PD / indefinite matrices;
                            Higham's theorem: symmetric input -> Q max(L, 0) Q^T
                            (+ a shift of the order of the spacing), PD input
                            unchanged, output symmetric PD, no random PSD matrix
                            closer to a symmetric input (the reference's call is
                            undefined: parity pinned by mathematics only)
ad_intf_pc64 / temp_coh64   closed forms (sec = ref e^{i phi} -> e^{-i phi};
                            consistent phases -> t = 1) and agreement with the
                            complex64 restatements on complex64 data
==========================  ===================================================
"""
from __future__ import annotations

import math

import numpy as np

# ---------------------------------------------------------------------------
# temporal network (moraine/tnet.py)
# ---------------------------------------------------------------------------


def image_pairs_from_bandwidth(nimages: int, bandwidth: int | None = None) -> np.ndarray:
    """All i<j pairs in ``np.triu_indices`` order with sec-ref <= bandwidth.

    Restates ``_imagepair_from_bandwidth`` (moraine/tnet.py:15-20).
    """
    if bandwidth is None:
        bandwidth = nimages
    assert nimages >= bandwidth
    out = [(i, j) for i in range(nimages) for j in range(i + 1, nimages) if j - i <= bandwidth]
    return np.asarray(out, dtype=np.int32).reshape(-1, 2)


def nimage_from_npair(npair: int) -> int:
    """Restates ``nimage_from_npair`` (moraine/tnet.py:123-130) by direct search."""
    for n in range(1, 100000):
        if n * (n - 1) == 2 * npair:
            return n
        if n * (n - 1) > 2 * npair:
            break
    raise ValueError("input is not correct value of number of all pairs")


def ad_intf_pc(ref_rslc, sec_rslc, idx, pc_is_shp) -> np.ndarray:
    """Restates ``_ad_intf_pc_numba`` (moraine/co.py:302-333) with FP64 sums:
    over window entries inside the raster and marked SHP, sum ref*conj(sec) /
    sqrt(sum|ref|^2 * sum|sec|^2)."""
    ref = np.asarray(ref_rslc, np.complex64).astype(np.complex128)
    sec = np.asarray(sec_rslc, np.complex64).astype(np.complex128)
    nl, w = ref.shape
    aw, rw = pc_is_shp.shape[1:]
    out = np.empty(idx.shape[0], np.complex64)
    for i, (a, r) in enumerate(idx):
        num, ra, sa = 0j, 0.0, 0.0
        for k in range(aw):
            for l in range(rw):
                y, x = a - aw // 2 + k, r - rw // 2 + l
                if 0 <= y < nl and 0 <= x < w and pc_is_shp[i, k, l]:
                    num += ref[y, x] * np.conj(sec[y, x])
                    ra += abs(ref[y, x]) ** 2
                    sa += abs(sec[y, x]) ** 2
        with np.errstate(invalid="ignore", divide="ignore"):
            out[i] = num / np.sqrt(ra * sa) if ra * sa > 0 else complex(np.nan, np.nan)
    return out


def ad_intf_pc64(ref_rslc, sec_rslc, idx, pc_is_shp) -> np.ndarray:
    """FP64 truth of ``ad_intf_pc`` (moraine/co.py:302-406) for complex128 input:
    the same window sums as ``ad_intf_pc`` on the unrounded complex128 samples,
    complex128 out."""
    ref = np.asarray(ref_rslc).astype(np.complex128)
    sec = np.asarray(sec_rslc).astype(np.complex128)
    nl, w = ref.shape
    aw, rw = pc_is_shp.shape[1:]
    out = np.empty(idx.shape[0], np.complex128)
    for i, (a, r) in enumerate(idx):
        sa, sr, ok = _window_valid(nl, w, aw // 2, rw // 2, int(a), int(r), aw, rw)
        sel = ok & pc_is_shp[i].astype(bool)
        x, y = ref[sa[sel], sr[sel]], sec[sa[sel], sr[sel]]
        with np.errstate(invalid="ignore", divide="ignore"):
            out[i] = np.sum(x * np.conj(y)) / np.sqrt(np.sum(np.abs(x) ** 2) * np.sum(np.abs(y) ** 2))
    return out


def isPD(co) -> np.ndarray:
    """Restates ``isPD`` (moraine/co.py:409-414) with the CuPy path's semantics
    (a failed factorisation gives a non-finite factor, not an exception):
    True where the Cholesky factorisation of the matrix succeeds (FP64)."""
    co = np.asarray(co, np.float64)
    out = np.empty(co.shape[:-2], bool)
    for i in np.ndindex(out.shape):
        try:
            out[i] = bool(np.isfinite(np.linalg.cholesky(co[i])).all())
        except np.linalg.LinAlgError:
            out[i] = False
    return out


def nearestPD(co) -> np.ndarray:
    """Restates ``nearestPD`` (moraine/co.py:429-463), Higham's method, in FP64
    with the undefined ``wherePD`` read as ``isPD``: B = (A + A^T)/2; A = U S Vh
    (SVD of the input, as the reference takes it); H = Vh^T S Vh; A3 = the
    symmetric part of (B + H)/2; while some matrix is not PD (k < 100): k += 1,
    A3 += (-lambda_min(A3) k^2 + spacing) I on those matrices, spacing the
    distance from ||A||_F to the next number of the input's dtype.  Returns the
    input's dtype."""
    a = np.asarray(co)
    dt = a.dtype if a.dtype in (np.float32, np.float64) else np.float64
    A = a.astype(np.float64)
    B = (A + np.swapaxes(A, -1, -2)) / 2
    _, s, Vh = np.linalg.svd(A)
    H = np.swapaxes(Vh, -1, -2) @ (s[..., :, None] * Vh)
    A2 = (B + H) / 2
    A3 = (A2 + np.swapaxes(A2, -1, -2)) / 2
    I = np.eye(A.shape[-1])
    nrm = np.linalg.norm(a.astype(dt), axis=(-2, -1)).astype(dt)
    spacing = (np.nextafter(nrm, np.asarray(np.inf, dt)) - nrm).astype(np.float64)
    k = 0
    while True:
        pd = isPD(A3.astype(dt))
        if pd.all() or k >= 100:
            break
        k += 1
        mineig = np.min(np.linalg.eigvalsh(A3), axis=-1)
        A3 = A3 + (~pd)[..., None, None] * I * (-mineig * k ** 2 + spacing)[..., None, None]
    return A3.astype(dt)


def regularize_spectral(coh, beta) -> np.ndarray:
    """Restates ``regularize_spectral`` (moraine/co.py:466-477) in complex128:
    (1 - beta) coh + beta I, beta broadcast over the stack's leading shape."""
    c = np.asarray(coh).astype(np.complex128)
    b = np.asarray(beta, np.float64)[..., None, None]
    return (1 - b) * c + b * np.eye(c.shape[-1])


def uncompress_coh(coh: np.ndarray, image_pairs: np.ndarray | None = None) -> np.ndarray:
    """Compressed (..., P) -> Hermitian (..., N, N), diag = 1.

    Restates ``uncompress_coh`` (moraine/co.py:281-299) with explicit loops.
    """
    if image_pairs is None:
        n = nimage_from_npair(coh.shape[-1])
        image_pairs = image_pairs_from_bandwidth(n)
    else:
        n = int(image_pairs[-1, -1]) + 1
    out = np.zeros(coh.shape[:-1] + (n, n), dtype=coh.dtype)
    # three passes in the reference's order (co.py:296-298): every pair, then
    # every conjugate, then the diagonal; a later write to an entry wins
    for p, (a, b) in enumerate(image_pairs):
        out[..., a, b] = coh[..., p]
    for p, (a, b) in enumerate(image_pairs):
        out[..., b, a] = np.conj(coh[..., p])
    for i in range(n):
        out[..., i, i] = 1
    return out


# ---------------------------------------------------------------------------
# SLC amplitude (moraine/calamp.py)
# ---------------------------------------------------------------------------


def rslc2amp(rslc: np.ndarray) -> np.ndarray:
    """Restates ``_rslc2amp_numba`` (moraine/calamp.py:17-25): ``np.abs`` of each
    complex64 sample under numba, i.e. the C library's ``hypotf``; glibc computes
    it as the FP64 sum of the exact squares, a correctly rounded sqrt and one
    rounding to float32, with +inf for an infinite part.  (NumPy's own complex64
    ``abs`` rounds differently in the last bit.)"""
    z = np.asarray(rslc, dtype=np.complex64)
    x = z.real.astype(np.float64)
    y = z.imag.astype(np.float64)
    with np.errstate(invalid="ignore", over="ignore"):
        a = np.sqrt(x * x + y * y).astype(np.float32)
    a[np.isinf(z.real) | np.isinf(z.imag)] = np.inf
    return a


def amp_disp(rslc: np.ndarray) -> np.ndarray:
    """Restates ``_amp_disp_numba`` (moraine/ps.py:17-29): per pixel amp = |z|
    (as rslc2amp), std(amp) / mean(amp) with the population std, in FP64."""
    a = rslc2amp(rslc).astype(np.float64)
    return (a.std(axis=-1) / a.mean(axis=-1)).astype(np.float32)


# ---------------------------------------------------------------------------
# KS test (moraine/shp.py)
# ---------------------------------------------------------------------------


def ks_p(x: float) -> float:
    """Literal restatement of ``_ks_p_numba`` (moraine/shp.py:19-32), f64."""
    x2 = -2 * x * x
    p = 0.0
    p2 = 0.0
    sign = 1
    for i in range(1, 101):
        p += sign * 2 * math.exp(x2 * i * i)
        if p == p2:
            return p
        sign = -sign
        p2 = p
    return p


def ks_lambda(n: int, k: int) -> float:
    """lambda = (en + 0.12 + 0.11/en) * dmax with en = sqrt(N/2), dmax = |k|/N.

    Same operation order as moraine/shp.py:42 (``abs((j2-j1)/n_imag)``) and
    :76-77 (``en = math.sqrt(n_imag/2)``; ``(en+0.12+0.11/en)*dmax``).
    """
    en = math.sqrt(n / 2)
    dmax = abs(k / n)
    return (en + 0.12 + 0.11 / en) * dmax


def ks_k_walk(ref: np.ndarray, sec: np.ndarray) -> int:
    """Literal merge walk of ``_ks_d_numba`` (moraine/shp.py:35-44) returning k = N*D.

    ``ref``/``sec`` are sorted 1-D stacks of equal length without NaN.
    """
    n = ref.shape[0]
    j1 = j2 = 0
    kmax = 0
    while j1 < n and j2 < n:
        f1 = ref[j1]
        f2 = sec[j2]
        if f1 <= f2:
            j1 += 1
        if f1 >= f2:
            j2 += 1
        d = abs(j2 - j1)
        if d > kmax:
            kmax = d
    return kmax


def ks_k_sorted(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    """Vectorised k for rows of sorted stacks ``a``, ``b`` of shape (n, N).

    k = max over distinct sample values x of |#{a<=x} - #{b<=x}|, which equals
    the merge-walk maximum of moraine/shp.py:35-44 (ties advance both
    pointers, so the walk visits exactly the end of every tie group).
    """
    n, N = a.shape
    v = np.concatenate([a, b], axis=1)
    lab = np.concatenate([np.ones((n, N), np.int32), -np.ones((n, N), np.int32)], axis=1)
    order = np.argsort(v, axis=1, kind="stable")
    vs = np.take_along_axis(v, order, axis=1)
    cs = np.cumsum(np.take_along_axis(lab, order, axis=1), axis=1)
    last = np.ones_like(vs, dtype=bool)
    last[:, :-1] = vs[:, :-1] != vs[:, 1:]
    return np.max(np.abs(cs) * last, axis=1).astype(np.int32)


def ks_tables(n: int, dtype=np.float32):
    """p(k) and dist(k) for k = 0..N in the output dtype (f64 series, then cast)."""
    p = np.array([ks_p(ks_lambda(n, k)) for k in range(n + 1)], dtype=np.float64).astype(dtype)
    d = np.array([abs(k / n) for k in range(n + 1)], dtype=np.float64).astype(dtype)
    return p, d


def ks_k_field(rmli: np.ndarray, az_half: int, r_half: int, chunk: int = 1 << 18):
    """k for every (pixel, window offset); -1 where the reference writes NaN.

    NaN rule (moraine/shp.py:57-75): ref stack with NaN (sorted last) -> whole
    window NaN; secondary out of bounds or with NaN -> NaN.
    """
    nl, w, n = rmli.shape
    s = np.sort(rmli, axis=-1)  # NaN last, as np.sort / _sort_numba (moraine/shp.py:117-125)
    nan = np.isnan(s[..., -1])
    aw, rw = 2 * az_half + 1, 2 * r_half + 1
    k = np.full((nl, w, aw, rw), -1, dtype=np.int32)
    ii, jj = np.meshgrid(np.arange(nl), np.arange(w), indexing="ij")
    for l in range(-az_half, az_half + 1):
        for m in range(-r_half, r_half + 1):
            si, sj = ii + l, jj + m
            ok = (si >= 0) & (si < nl) & (sj >= 0) & (sj < w) & ~nan
            si_c, sj_c = np.clip(si, 0, nl - 1), np.clip(sj, 0, w - 1)
            ok &= ~nan[si_c, sj_c]
            idx = np.nonzero(ok.ravel())[0]
            if idx.size == 0:
                continue
            a_all = s.reshape(-1, n)
            sec_lin = (si_c * w + sj_c).ravel()
            out = k[:, :, l + az_half, m + r_half].reshape(-1)
            for c0 in range(0, idx.size, chunk):
                sel = idx[c0:c0 + chunk]
                out[sel] = ks_k_sorted(a_all[sel], a_all[sec_lin[sel]])
            k[:, :, l + az_half, m + r_half] = out.reshape(nl, w)
    return k


def ks_test(rmli: np.ndarray, az_half_win: int, r_half_win: int, return_dist: bool = False):
    """Exact-semantics restatement of ``ks_test`` CPU path (moraine/shp.py:269-288).

    Output dtype = rmli.dtype; p = dtype(ks_p_f64(lambda(k))); dist = dtype(k/N).
    Where the secondary stack has NaN the numba dist is uninitialised
    (moraine/shp.py:106-107); this oracle writes NaN there, as the CUDA kernel
    does (moraine/shp.py:158-162).
    """
    n = rmli.shape[-1]
    k = ks_k_field(rmli, az_half_win, r_half_win)
    pt, dt = ks_tables(n, rmli.dtype)
    bad = k < 0
    kk = np.where(bad, 0, k)
    p = pt[kk]
    p[bad] = np.nan
    if not return_dist:
        return p
    d = dt[kk]
    d[bad] = np.nan
    return d, p


def ks_test_loop(rmli: np.ndarray, az_half_win: int, r_half_win: int):
    """Literal pure-python restatement of ``_ks_test_numba`` (moraine/shp.py:80-113).

    For tiny inputs only.  Returns (dist, p) with NaN where the CUDA path
    writes NaN.
    """
    nl, w, n = rmli.shape
    s = np.sort(rmli, axis=-1)
    aw, rw = 2 * az_half_win + 1, 2 * r_half_win + 1
    dist = np.empty((nl, w, aw, rw), dtype=rmli.dtype)
    p = np.empty((nl, w, aw, rw), dtype=rmli.dtype)
    for i in range(nl):
        for j in range(w):
            ref = s[i, j]
            for l in range(-az_half_win, az_half_win + 1):
                for m in range(-r_half_win, r_half_win + 1):
                    si, sj = i + l, j + m
                    o = (i, j, l + az_half_win, m + r_half_win)
                    if math.isnan(ref[-1]) or si < 0 or si >= nl or sj < 0 or sj >= w:
                        dist[o] = np.nan
                        p[o] = np.nan
                        continue
                    sec = s[si, sj]
                    if math.isnan(sec[-1]):
                        dist[o] = np.nan
                        p[o] = np.nan
                        continue
                    k = ks_k_walk(ref, sec)
                    en = math.sqrt(n / 2)
                    dmax = abs(k / n)
                    dist[o] = dmax
                    p[o] = ks_p((en + 0.12 + 0.11 / en) * dmax)
    return dist, p


def select_shp(p: np.ndarray, p_max: float):
    """Restates ``select_shp`` (moraine/shp.py:308-323): is_shp = p < p_max, count."""
    is_shp = p < p_max
    return is_shp, np.count_nonzero(is_shp, axis=(-2, -1)).astype(np.int32)


# ---------------------------------------------------------------------------
# coherence (moraine/co.py)
# ---------------------------------------------------------------------------


def _window_valid(nl, w, az, r, ia, ir, aw, rw):
    k = np.arange(aw)[:, None]
    l = np.arange(rw)[None, :]
    sa, sr = np.broadcast_arrays(ia - az + k, ir - r + l)
    return sa, sr, (sa >= 0) & (sa < nl) & (sr >= 0) & (sr < w)


def emperical_co_pc(rslc, idx, pc_is_shp, image_pairs=None):
    """Restates ``_emperical_co_pc_numba`` (moraine/co.py:193-227).

    Accumulators: ``_co_nume`` complex128 fed with complex64 products
    (co.py:213, :224); power sums float64 fed with float32 squares (co.py:214-223);
    result cast to complex64.  Half window = shape//2 (co.py:201).
    """
    nl, w, n = rslc.shape
    if image_pairs is None:
        image_pairs = image_pairs_from_bandwidth(n)
    image_pairs = np.asarray(image_pairs, dtype=np.int32)
    aw, rw = pc_is_shp.shape[1:]
    az, r = aw // 2, rw // 2
    m_idx = image_pairs[:, 0]
    j_idx = image_pairs[:, 1]
    out = np.empty((idx.shape[0], image_pairs.shape[0]), dtype=np.complex64)
    for i in range(idx.shape[0]):
        sa, sr, ok = _window_valid(nl, w, az, r, int(idx[i, 0]), int(idx[i, 1]), aw, rw)
        sel = ok & pc_is_shp[i].astype(bool)
        z = rslc[sa[sel], sr[sel]].astype(np.complex64)  # (K, N) row-major window order
        zm = z[:, m_idx]
        zj = z[:, j_idx]
        prod = (zm * np.conj(zj)).astype(np.complex64)  # c64 products
        nume = np.zeros(image_pairs.shape[0], np.complex128)
        am = np.zeros(image_pairs.shape[0], np.float64)
        aj = np.zeros(image_pairs.shape[0], np.float64)
        for kk in range(z.shape[0]):  # sequential accumulation in window order
            nume += prod[kk]
            am += (zm[kk].real ** 2 + zm[kk].imag ** 2).astype(np.float32)
            aj += (zj[kk].real ** 2 + zj[kk].imag ** 2).astype(np.float32)
        out[i] = (nume / np.sqrt(am * aj)).astype(np.complex64)
    return out


def emperical_co64(rslc, is_shp):
    """FP64 truth of ``emperical_co`` (moraine/co.py:61-137): (cov, coh) complex128.

    cov = sum/n over in-bounds SHPs, coh = sum/sqrt(sum|z_m|^2 sum|z_j|^2), all
    (m, j) including the diagonal (the CUDA kernel computes it, co.py:83-111).
    """
    nl, w, n = rslc.shape
    aw, rw = is_shp.shape[-2:]
    az, r = (aw - 1) // 2, (rw - 1) // 2
    z64 = rslc.astype(np.complex128)
    cov = np.empty((nl, w, n, n), np.complex128)
    coh = np.empty((nl, w, n, n), np.complex128)
    for i in range(nl):
        for j in range(w):
            sa, sr, ok = _window_valid(nl, w, az, r, i, j, aw, rw)
            sel = ok & is_shp[i, j].astype(bool)
            z = z64[sa[sel], sr[sel]]
            s = z.T @ z.conj()
            pw = np.real(np.einsum("kn,kn->n", z, z.conj()))
            cnt = z.shape[0]
            with np.errstate(invalid="ignore", divide="ignore"):
                cov[i, j] = s / cnt
                coh[i, j] = s / np.sqrt(np.outer(pw, pw))
    return cov, coh


def emperical_co_pc64(rslc, idx, pc_is_shp, image_pairs=None):
    """FP64 truth of ``emperical_co_pc`` (moraine/co.py:141-266)."""
    nl, w, n = rslc.shape
    if image_pairs is None:
        image_pairs = image_pairs_from_bandwidth(n)
    image_pairs = np.asarray(image_pairs)
    aw, rw = pc_is_shp.shape[1:]
    az, r = aw // 2, rw // 2
    z64 = rslc.astype(np.complex128)
    out = np.empty((idx.shape[0], image_pairs.shape[0]), np.complex128)
    for i in range(idx.shape[0]):
        sa, sr, ok = _window_valid(nl, w, az, r, int(idx[i, 0]), int(idx[i, 1]), aw, rw)
        sel = ok & pc_is_shp[i].astype(bool)
        z = z64[sa[sel], sr[sel]]
        zm = z[:, image_pairs[:, 0]]
        zj = z[:, image_pairs[:, 1]]
        with np.errstate(invalid="ignore", divide="ignore"):
            out[i] = np.sum(zm * zj.conj(), 0) / np.sqrt(np.sum(np.abs(zm) ** 2, 0) * np.sum(np.abs(zj) ** 2, 0))
    return out


# ---------------------------------------------------------------------------
# EMI phase linking (moraine/pl.py)
# ---------------------------------------------------------------------------


def _emi_from_gamma(G, ref, inv, eigh, cdtype, rdtype):
    A = np.abs(G).astype(rdtype)
    Ainv = inv(A)
    M = (Ainv * G).astype(cdtype)
    lam, V = eigh(M)
    v = V[..., :, 0] * np.conj(V[..., ref:ref + 1, 0])
    with np.errstate(invalid="ignore", divide="ignore"):
        ph = v / np.abs(v)
    return ph, lam[..., 0]


def emi64(coh: np.ndarray, ref: int = 0):
    """FP64 truth of ``emi`` (moraine/pl.py:18-77): complex128 throughout.

    Gamma = uncompress(coh); M = inv(|Gamma|) o Gamma; (lam, V) = eigh(M);
    ph = V[:,0]*conj(V[ref,0]) normalised elementwise; quality = lam[0].
    """
    G = uncompress_coh(coh.astype(np.complex128))
    return _emi_from_gamma(G, ref, np.linalg.inv, np.linalg.eigh, np.complex128, np.float64)


def emi_ref32(coh: np.ndarray, ref: int = 0):
    """Restates ``_emi_numba`` (moraine/pl.py:45-65): float32 ``inv`` of |Gamma|
    and complex64 ``eigh`` (LAPACK), ph complex64, quality float32."""
    G = uncompress_coh(coh.astype(np.complex64))
    ph, q = _emi_from_gamma(G, ref, np.linalg.inv, np.linalg.eigh, np.complex64, np.float32)
    return ph.astype(np.complex64), q.astype(np.float32)


# ---------------------------------------------------------------------------
# temporal coherence (moraine/pl.py, moraine/pqm.py)
# ---------------------------------------------------------------------------


def ds_temp_coh(coh: np.ndarray, ph: np.ndarray, image_pairs=None):
    """Restates ``_ds_temp_coh_numba`` (moraine/pl.py:80-100): c64 accumulator,
    term = coh*(conj(ph_n)*ph_k)/|coh|; t = |sum| / n_pairs (float32)."""
    n = ph.shape[-1]
    if image_pairs is None:
        image_pairs = image_pairs_from_bandwidth(n)
    image_pairs = np.asarray(image_pairs)
    coh = coh.astype(np.complex64)
    ph = ph.astype(np.complex64)
    a, b = image_pairs[:, 0], image_pairs[:, 1]
    icp = (np.conj(ph[:, a]) * ph[:, b]).astype(np.complex64)
    with np.errstate(invalid="ignore", divide="ignore"):
        term = ((coh * icp).astype(np.complex64) / np.abs(coh)).astype(np.complex64)
    acc = np.zeros(ph.shape[0], np.complex64)
    for j in range(image_pairs.shape[0]):
        acc = (acc + term[:, j]).astype(np.complex64)
    return (np.abs(acc) / image_pairs.shape[0]).astype(np.float32)


def ds_temp_coh64(coh, ph, image_pairs=None):
    """FP64 truth of the temporal coherence (moraine/pl.py:80-123)."""
    n = ph.shape[-1]
    if image_pairs is None:
        image_pairs = image_pairs_from_bandwidth(n)
    image_pairs = np.asarray(image_pairs)
    c = coh.astype(np.complex128)
    p = ph.astype(np.complex128)
    a, b = image_pairs[:, 0], image_pairs[:, 1]
    with np.errstate(invalid="ignore", divide="ignore"):
        t = np.sum(c / np.abs(c) * np.conj(p[:, a]) * p[:, b], axis=1)
    return np.abs(t) / image_pairs.shape[0]


def temp_coh(intf: np.ndarray, rslc: np.ndarray, image_pairs=None):
    """Restates ``temp_coh`` (moraine/pqm.py:46-71, :124-159) WITHOUT mutating inputs.

    rslc and intf are normalised to unit modulus (c64), then
    t = |sum intf_j * conj(r_n) * r_k| / n_pairs.  3-D raster input is treated as
    a point cloud and reshaped back.
    """
    is_ras = intf.ndim == 3
    shape = intf.shape[:2]
    intf = intf.reshape(-1, intf.shape[-1]).astype(np.complex64)
    rslc = rslc.reshape(-1, rslc.shape[-1]).astype(np.complex64)
    n = rslc.shape[-1]
    if image_pairs is None:
        image_pairs = image_pairs_from_bandwidth(n)
    image_pairs = np.asarray(image_pairs)
    with np.errstate(invalid="ignore", divide="ignore"):
        r = (rslc / np.abs(rslc)).astype(np.complex64)
        f = (intf / np.abs(intf)).astype(np.complex64)
    a, b = image_pairs[:, 0], image_pairs[:, 1]
    term = (f * (np.conj(r[:, a]) * r[:, b]).astype(np.complex64)).astype(np.complex64)
    acc = np.zeros(r.shape[0], np.complex64)
    for j in range(image_pairs.shape[0]):
        acc = (acc + term[:, j]).astype(np.complex64)
    out = (np.abs(acc) / image_pairs.shape[0]).astype(np.float32)
    return out.reshape(shape) if is_ras else out


def temp_coh64(intf, rslc, image_pairs=None):
    """FP64 truth of ``temp_coh`` (moraine/pqm.py:46-121) on point-cloud input:
    t = |sum_p intf_p conj(r_a) r_b / (|intf_p| |r_a| |r_b|)| / n_pairs."""
    n = rslc.shape[-1]
    if image_pairs is None:
        image_pairs = image_pairs_from_bandwidth(n)
    image_pairs = np.asarray(image_pairs)
    f = np.asarray(intf).astype(np.complex128)
    r = np.asarray(rslc).astype(np.complex128)
    a, b = image_pairs[:, 0], image_pairs[:, 1]
    with np.errstate(invalid="ignore", divide="ignore"):
        t = np.sum(f * np.conj(r[:, a]) * r[:, b] / (np.abs(f) * np.abs(r[:, a]) * np.abs(r[:, b])), axis=1)
    return np.abs(t) / image_pairs.shape[0]


def emperical_co_emi_temp_coh_pc(rslc, idx, pc_is_shp, batch_size=1000):
    """Restates the batch driver (moraine/pl.py:148-168) over the oracle parts."""
    n_pc = idx.shape[0]
    n = rslc.shape[2]
    ph = np.empty((n_pc, n), np.complex64)
    q = np.empty(n_pc, np.float32)
    t = np.empty(n_pc, np.float32)
    for s in range(0, n_pc, batch_size):
        e = min(s + batch_size, n_pc)
        c = emperical_co_pc(rslc, idx[s:e], pc_is_shp[s:e])
        ph[s:e], q[s:e] = emi_ref32(c)
        t[s:e] = ds_temp_coh(c, ph[s:e])
    return ph, q, t
