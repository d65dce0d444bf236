"""Pins the oracle (CPU restatement of the reference) before it is trusted as the checker.

Each test targets a plausible restatement mistake: the printed reference values,
closed forms / library routines that compute the same mathematics, literal
loops, and the C port against the numpy restatement.
"""
import json
import math
import os

import numpy as np
import pytest
import scipy.special as sps
import scipy.stats as st
from hypothesis import given, settings
from hypothesis import strategies as hs

from oracle import moraine_ref as R

GOLD = os.path.join(os.path.dirname(__file__), "golden")
PRINTED = json.load(open(os.path.join(GOLD, "printed.json")))


# ---------------------------------------------------------------- KS p series
def test_ks_p_printed_value():
    p, d = R.ks_tables(20)
    assert format(p[7], ".8g") == PRINTED["ks_p_N20_k7_f32"]
    assert np.float32(PRINTED["ks_p_N20_k7_f32"]) == p[7]
    assert d[7] == np.float32(PRINTED["ks_dist_N20_k7"])
    assert p[0] == PRINTED["ks_self_p"]  # the D=0 artifact: the 100-term series sums to 0


def test_ks_p_is_kolmogorov_survival():
    # sum (-1)^(i-1) 2 exp(-2 i^2 x^2) is scipy.special.kolmogorov(x) (converged region)
    xs = np.linspace(0.3, 4.0, 300)
    err = max(abs(R.ks_p(x) - sps.kolmogorov(x)) for x in xs)
    assert err < 1e-13


def test_ks_p_wrong_sign_or_term_is_caught():
    # a dropped factor 2 in the exponent or a flipped sign gives a different p(20,7)
    x = R.ks_lambda(20, 7)
    bad1 = sum((-1) ** (i - 1) * 2 * math.exp(-1 * i * i * x * x) for i in range(1, 101))
    bad2 = sum(2 * math.exp(-2 * i * i * x * x) for i in range(1, 101))
    printed = np.float32(PRINTED["ks_p_N20_k7_f32"])
    assert np.float32(R.ks_p(x)) == printed
    assert np.float32(bad1) != printed and np.float32(bad2) != printed


def test_shp_threshold_table():
    # k_c(N) at p_max = 0.05: SHP iff k == 0 or k >= k_c
    expect = {17: 8, 20: 9, 30: 11, 60: 15, 90: 18, 120: 21, 150: 24, 200: 27}
    for n, kc in expect.items():
        p, _ = R.ks_tables(n)
        shp = [k for k in range(n + 1) if float(p[k]) < 0.05]
        assert shp == [0] + list(range(kc, n + 1))


# ---------------------------------------------------------------- KS merge walk
def _walk_vs_scipy(a, b):
    a, b = np.sort(a), np.sort(b)
    n = a.size
    k1 = R.ks_k_walk(a, b)
    k2 = int(R.ks_k_sorted(a[None], b[None])[0])
    k3 = int(round(st.ks_2samp(a, b, method="asymp").statistic * n))
    return k1, k2, k3


@settings(max_examples=300, deadline=None)
@given(n=hs.integers(1, 30), lo=hs.integers(0, 6), seed=hs.integers(0, 2**31 - 1))
def test_walk_matches_cdf_and_scipy_tie_heavy(n, lo, seed):
    rng = np.random.default_rng(seed)
    a = rng.integers(0, lo + 2, n).astype(np.float32)
    b = rng.integers(0, lo + 2, n).astype(np.float32)
    k1, k2, k3 = _walk_vs_scipy(a, b)
    assert k1 == k2 == k3


def test_walk_continuous_and_signed_zero():
    rng = np.random.default_rng(3)
    for _ in range(200):
        n = int(rng.integers(2, 64))
        a = rng.standard_normal(n).astype(np.float32)
        b = (rng.standard_normal(n) * 1.5 + 0.3).astype(np.float32)
        k1, k2, k3 = _walk_vs_scipy(a, b)
        assert k1 == k2 == k3
    # -0.0 and +0.0 compare equal in the reference's float compares
    assert R.ks_k_walk(np.array([-0.0, 1.0], np.float32), np.array([0.0, 1.0], np.float32)) == 0


def test_ks_test_printed_1x2_pattern():
    # find a 20-image pair with k = 7 (the printed example's D = 0.35), check NaN/OOB layout
    rng = np.random.default_rng(0)
    while True:
        s1 = rng.uniform(size=20).astype(np.float32)
        s2 = rng.standard_normal(20).astype(np.float32)
        if R.ks_k_walk(np.sort(s1), np.sort(s2)) == 7:
            break
    rmli = np.stack((s1, s2)).reshape(1, 2, 20)
    d, p = R.ks_test(rmli, 1, 1, return_dist=True)
    assert list(d.shape) == PRINTED["ks_1x2_shape"]
    pat = np.array(PRINTED["ks_1x2_dist_pattern"], dtype=object)
    for idx in np.ndindex(d.shape):
        v = pat[idx]
        if v is None:
            assert np.isnan(d[idx]) and np.isnan(p[idx])
        else:
            assert d[idx] == np.float32(v)
    assert format(p[0, 0, 1, 2], ".8g") == PRINTED["ks_p_N20_k7_f32"]
    assert p[0, 0, 1, 1] == 0.0


def test_ks_test_literal_loop_matches_vectorised():
    rng = np.random.default_rng(9)
    rmli = rng.integers(0, 4, (5, 6, 7)).astype(np.float32)  # heavy ties
    rmli[1, 2, 3] = np.nan
    d1, p1 = R.ks_test_loop(rmli, 2, 3)
    d2, p2 = R.ks_test(rmli, 2, 3, return_dist=True)
    np.testing.assert_array_equal(p1, p2)
    np.testing.assert_array_equal(d1, d2)


def test_ks_test_symmetry():
    rng = np.random.default_rng(4)
    rmli = rng.random((9, 11, 13), dtype=np.float32)
    az, r = 2, 3
    p = R.ks_test(rmli, az, r)
    for i in range(9):
        for j in range(11):
            for l in range(2 * az + 1):
                for m in range(2 * r + 1):
                    si, sj = i + l - az, j + m - r
                    if 0 <= si < 9 and 0 <= sj < 11:
                        assert p[i, j, l, m] == p[si, sj, 2 * az - l, 2 * r - m]


def test_port_matches_numpy_restatement(port):
    g = np.load(os.path.join(GOLD, "ks_cfg1.npz"))
    rmli, k = g["rmli"], g["k"].astype(np.int32)
    d, p, kp = port.ks_test(rmli, int(g["az"]), int(g["r"]), return_dist=True, return_k=True)
    np.testing.assert_array_equal(kp, k)
    pt, dt = R.ks_tables(rmli.shape[-1])
    ok = k >= 0
    np.testing.assert_array_equal(p[ok], pt[k[ok]])
    np.testing.assert_array_equal(d[ok], dt[k[ok]])
    assert np.all(np.isnan(p[~ok]))


def test_port_per_pair_series_equals_per_k_table(port):
    """The timed baseline evaluates the p series for every pixel pair, as the
    reference does; the values equal the per-k table's exactly."""
    g = np.load(os.path.join(GOLD, "ks_cfg1.npz"))
    rmli = g["rmli"]
    p_k = port.ks_test(rmli, int(g["az"]), int(g["r"]))
    p_pair = port.ks_test(rmli, int(g["az"]), int(g["r"]), per_pair=True)
    np.testing.assert_array_equal(p_pair.view(np.uint32), p_k.view(np.uint32))


def test_golden_ks_regression():
    g = np.load(os.path.join(GOLD, "ks_cfg1.npz"))
    k = R.ks_k_field(g["rmli"], int(g["az"]), int(g["r"]))
    np.testing.assert_array_equal(k, g["k"].astype(np.int32))


def test_select_shp_closed_form():
    rng = np.random.default_rng(1)
    p = rng.random(20 * 20 * 5 * 5).astype(np.float32)
    p[rng.choice(p.size, 100, replace=False)] = np.nan
    p = p.reshape(20, 20, 5, 5)
    m, c = R.select_shp(p, 0.05)
    np.testing.assert_array_equal(m, p < 0.05)
    np.testing.assert_array_equal(c, np.count_nonzero(p < 0.05, axis=(-2, -1)).astype(np.int32))


# ---------------------------------------------------------------- coherence
def test_co_brute_force_and_consistency():
    g = np.load(os.path.join(GOLD, "co_small.npz"))
    rslc, is_shp = g["rslc"], g["is_shp"]
    cov, coh = R.emperical_co64(rslc, is_shp)
    az, r = 1, 2
    # python-loop brute force as in the reference test (nbs/API/co.ipynb:444-470)
    for i, j in [(0, 0), (2, 3), (4, 9)]:
        for a, b in [(0, 0), (3, 1), (16, 5)]:
            s, pa, pb, ns = 0j, 0.0, 0.0, 0
            for m in range(3):
                for n in range(5):
                    if is_shp[i, j, m, n]:
                        za = complex(rslc[i + m - az, j + n - r, a])
                        zb = complex(rslc[i + m - az, j + n - r, b])
                        s += za * zb.conjugate()
                        pa += abs(za) ** 2
                        pb += abs(zb) ** 2
                        ns += 1
            assert abs(s / ns - cov[i, j, a, b]) < 1e-12
            assert abs(s / math.sqrt(pa * pb) - coh[i, j, a, b]) < 1e-12
    np.testing.assert_allclose(np.abs(np.diagonal(coh, axis1=-2, axis2=-1)), 1.0, atol=1e-12)
    assert np.all(np.abs(coh) <= 1 + 1e-12)
    iu = np.triu_indices(17, 1)
    pc64 = R.emperical_co_pc64(rslc, g["idx"], is_shp.reshape(-1, 3, 5))
    np.testing.assert_allclose(coh.reshape(-1, 17, 17)[:, iu[0], iu[1]], pc64, atol=1e-13)
    # c128-accumulator restatement agrees with FP64 truth to c64 rounding
    np.testing.assert_allclose(g["coh_pc"], pc64, atol=2e-7)
    np.testing.assert_allclose(R.emperical_co_pc(rslc, g["idx"], is_shp.reshape(-1, 3, 5)), g["coh_pc"], atol=0)


def test_co_pc_port_matches_numpy(port):
    g = np.load(os.path.join(GOLD, "co_small.npz"))
    pairs = R.image_pairs_from_bandwidth(17)
    c = port.emperical_co_pc(g["rslc"], g["idx"], g["is_shp"].reshape(-1, 3, 5), pairs)
    np.testing.assert_allclose(c, g["coh_pc"], atol=1e-7)


# ---------------------------------------------------------------- tnet / uncompress
def test_pairs_and_uncompress_printed():
    np.testing.assert_array_equal(R.image_pairs_from_bandwidth(5, 2), PRINTED["tnet_from_bandwidth_5_2"])
    assert R.image_pairs_from_bandwidth(17).shape[0] == PRINTED["co_pc_npairs_17"]
    assert R.image_pairs_from_bandwidth(5, 3).shape[0] == PRINTED["tnet_from_bandwidth_5_3_npairs"]
    assert R.nimage_from_npair(6) == PRINTED["nimage_from_npair_6"]
    pairs = R.image_pairs_from_bandwidth(5, 2)
    c = (np.arange(7) + 1j * np.arange(7)).astype(np.complex64)
    u = R.uncompress_coh(c, pairs)
    assert u[0, 3] == 0 and u[4, 0] == 0 and u[1, 3] == c[3] and u[3, 1] == np.conj(c[3])
    np.testing.assert_array_equal(np.diag(u), 1)
    iu = np.triu_indices(6, 1)
    c2 = np.random.default_rng(0).random(15).astype(np.complex64)
    u2 = R.uncompress_coh(c2)
    np.testing.assert_array_equal(u2[iu], c2)
    np.testing.assert_array_equal(u2, u2.conj().T)


def test_uncompress_colliding_pairs_hand_derived():
    # pairs (0,1), (1,0), (0,0), (0,1) with coh c0..c3, n = 2 from the last
    # pair; the reference's statements (moraine/co.py:296-298) give, in order:
    #   pass 1: [0,1]=c0, [1,0]=c1, [0,0]=c2, [0,1]=c3
    #   pass 2: [1,0]=c0*, [0,1]=c1*, [0,0]=c2*, [1,0]=c3*
    #   diag  : [0,0]=[1,1]=1
    c = np.array([1 + 2j, 3 + 4j, 5 + 6j, 7 + 8j], np.complex64)
    pairs = np.array([[0, 1], [1, 0], [0, 0], [0, 1]], np.int32)
    u = R.uncompress_coh(c, pairs)
    want = np.array([[1, np.conj(c[1])], [np.conj(c[3]), 1]], np.complex64)
    np.testing.assert_array_equal(u, want)


# ---------------------------------------------------------------- EMI
@pytest.mark.parametrize("n", [17, 30, 60])
def test_emi64_closure_free_closed_form(n):
    g = np.load(os.path.join(GOLD, "emi_small.npz"))
    ph, q = R.emi64(g[f"cf_coh_{n}"])
    exp = g[f"cf_ph_{n}"]
    err = np.abs(np.angle(ph * np.conj(exp)))
    assert err.max() < 5e-5  # coherence stored as complex64
    np.testing.assert_allclose(q, 1.0, atol=1e-4)
    ph32, q32 = R.emi_ref32(g[f"cf_coh_{n}"])
    assert np.abs(np.angle(ph32 * np.conj(exp))).max() < 2e-3
    np.testing.assert_allclose(ph[:, 0], 1.0, atol=1e-12)


def test_emi64_eigen_residual_and_minimality():
    g = np.load(os.path.join(GOLD, "emi_small.npz"))
    c = g["ds_coh_30"].astype(np.complex128)
    G = R.uncompress_coh(c)
    M = np.linalg.inv(np.abs(G)) * G
    ph, q = R.emi64(g["ds_coh_30"])
    lam, V = np.linalg.eigh(M)
    np.testing.assert_allclose(q, lam[:, 0], rtol=1e-10, atol=1e-10)
    v = V[:, :, 0]
    r = np.einsum("pij,pj->pi", M, v) - lam[:, :1] * v
    assert np.max(np.linalg.norm(r, axis=1)) < 1e-9 * np.max(np.abs(lam))
    np.testing.assert_allclose(np.abs(ph), 1.0, atol=1e-12)


def test_emi_golden_regression():
    g = np.load(os.path.join(GOLD, "emi_small.npz"))
    for n in (17, 30, 60):
        ph, q = R.emi64(g[f"ds_coh_{n}"])
        assert np.abs(np.angle(ph * np.conj(g[f"ds_ph64_{n}"]))).max() < 1e-6
        np.testing.assert_allclose(q, g[f"ds_q64_{n}"], rtol=1e-12)


# ---------------------------------------------------------------- temporal coherence
def test_tcoh_consistent_is_one_and_hand_case():
    rng = np.random.default_rng(2)
    n = 6
    th = np.exp(1j * rng.uniform(-np.pi, np.pi, (4, n))).astype(np.complex64)
    pairs = R.image_pairs_from_bandwidth(n)
    coh = (0.7 * th[:, pairs[:, 0]] * np.conj(th[:, pairs[:, 1]])).astype(np.complex64)
    np.testing.assert_allclose(R.ds_temp_coh(coh, th), 1.0, atol=1e-6)
    np.testing.assert_allclose(R.ds_temp_coh64(coh, th), 1.0, atol=1e-12)
    np.testing.assert_allclose(R.temp_coh(coh, th), 1.0, atol=1e-6)
    # 3 images: pairs (0,1),(0,2),(1,2); phases 0; coh phases a, b, c -> |e^ia + e^ib + e^ic|/3
    a, b, c = 0.3, -1.1, 2.0
    coh3 = np.array([[np.exp(1j * a), np.exp(1j * b), np.exp(1j * c)]], np.complex64)
    ph3 = np.ones((1, 3), np.complex64)
    exp = abs(np.exp(1j * a) + np.exp(1j * b) + np.exp(1j * c)) / 3
    assert abs(R.ds_temp_coh(coh3, ph3)[0] - exp) < 1e-6
    assert abs(R.temp_coh(coh3 * 5, ph3 * 2)[0] - exp) < 1e-6  # normalises magnitudes


def test_tcoh_port_and_golden(port):
    g = np.load(os.path.join(GOLD, "tcoh_small.npz"))
    pairs = R.image_pairs_from_bandwidth(17)
    np.testing.assert_allclose(R.ds_temp_coh(g["coh"], g["ph"]), g["t_ds"], atol=0)
    np.testing.assert_allclose(R.temp_coh(g["coh"], g["ph"]), g["t_pq"], atol=0)
    np.testing.assert_allclose(port.ds_temp_coh(g["coh"], g["ph"], pairs), g["t_ds"], atol=2e-6)
    np.testing.assert_allclose(port.temp_coh(g["coh"], g["ph"], pairs), g["t_pq"], atol=2e-6)
    np.testing.assert_allclose(R.ds_temp_coh64(g["coh"], g["ph"]), g["t_ds"], atol=2e-6)


# ---------------------------------------------------------------- amplitude
def test_rslc2amp_matches_libm_hypotf():
    """The oracle's amplitude equals the C library's hypotf (what numba's abs of a
    complex64 calls) on random bit patterns, specials and denormals."""
    import ctypes

    libm = ctypes.CDLL("libm.so.6")
    hypotf = libm.hypotf
    hypotf.restype = ctypes.c_float
    hypotf.argtypes = [ctypes.c_float, ctypes.c_float]
    rng = np.random.default_rng(5)
    bits = rng.integers(0, 2 ** 32, size=(20000, 2), dtype=np.uint64).astype(np.uint32)
    xy = bits.view(np.float32).reshape(-1, 2)
    sp = np.array([0.0, -0.0, np.inf, -np.inf, np.nan, 1e-45, -1e-45, 1.17e-38, 3.4e38, -3.4e38, 1.0, 3.0, 4.0],
                  np.float32)
    X, Y = np.meshgrid(sp, sp)
    xy = np.concatenate([xy, np.stack([X.ravel(), Y.ravel()], -1)])
    z = (xy[:, 0].astype(np.complex64) + 1j * xy[:, 1].astype(np.complex64)).astype(np.complex64)
    z.real[:] = xy[:, 0]
    z.imag[:] = xy[:, 1]
    got = R.rslc2amp(z)
    want = np.array([hypotf(float(a), float(b)) for a, b in xy], np.float32)
    same = (got.view(np.uint32) == want.view(np.uint32)) | (np.isnan(got) & np.isnan(want))
    assert same.all(), xy[~same][:5]
    assert R.rslc2amp(np.array([3 + 4j], np.complex64))[0] == 5.0


def test_amp_disp_closed_forms():
    """Amplitude dispersion: constant amplitude -> 0; amplitudes {1, 3} -> std 1 /
    mean 2 = 0.5; phases do not matter; scale invariant."""
    ph = np.exp(1j * np.linspace(0, 6, 8)).astype(np.complex64)
    z = np.stack([2 * ph, np.where(np.arange(8) % 2 == 0, 1, 3) * ph]).reshape(1, 2, 8).astype(np.complex64)
    d = R.amp_disp(z)
    assert d.shape == (1, 2)
    assert abs(d[0, 0]) < 1e-7 and abs(d[0, 1] - 0.5) < 1e-6
    np.testing.assert_allclose(R.amp_disp(z * np.float32(7.5)), d, rtol=1e-6, atol=1e-7)


def test_ad_intf_pc_closed_forms():
    """sec = ref * exp(i phi) gives exp(-i phi) at every point with an SHP
    inside the raster; no SHP gives NaN; windows clipped at the raster edge."""
    rng = np.random.default_rng(3)
    ref = (rng.standard_normal((9, 11)) + 1j * rng.standard_normal((9, 11))).astype(np.complex64)
    sec = (ref * np.exp(1j * 0.7)).astype(np.complex64)
    idx = np.array([[0, 0], [4, 5], [8, 10], [2, 9]], np.int32)
    shp = np.ones((4, 3, 5), bool)
    shp[3] = False
    out = R.ad_intf_pc(ref, sec, idx, shp)
    np.testing.assert_allclose(out[:3], np.exp(-1j * 0.7), atol=1e-6)
    assert np.isnan(out[3].real)
    # a different secondary: |intf| <= 1 (Cauchy-Schwarz), equals 1 only for proportional stacks
    sec2 = (rng.standard_normal((9, 11)) + 1j * rng.standard_normal((9, 11))).astype(np.complex64)
    assert np.all(np.abs(R.ad_intf_pc(ref, sec2, idx[:3], shp[:3])) < 1)


def test_regularize_spectral_closed_forms():
    """beta = 0 keeps coh, beta = 1 gives I, the diagonal of a unit-diagonal
    coherence stays 1, per-matrix beta broadcasts over the stack."""
    rng = np.random.default_rng(2)
    c = (rng.standard_normal((3, 4, 5, 5)) + 1j * rng.standard_normal((3, 4, 5, 5))).astype(np.complex64)
    np.testing.assert_array_equal(R.regularize_spectral(c, 0.0), c.astype(np.complex128))
    np.testing.assert_array_equal(R.regularize_spectral(c, 1.0), np.broadcast_to(np.eye(5), c.shape))
    u = c.copy()
    u[..., range(5), range(5)] = 1
    np.testing.assert_allclose(np.diagonal(R.regularize_spectral(u, 0.3), axis1=-2, axis2=-1), 1, atol=1e-15)
    b = rng.random((3, 4))
    r = R.regularize_spectral(c, b)
    np.testing.assert_allclose(r[1, 2], (1 - b[1, 2]) * c[1, 2] + b[1, 2] * np.eye(5), atol=1e-12)


def test_fp64_variants_closed_forms_and_agreement():
    """ad_intf_pc64: sec = ref e^{i phi} -> e^{-i phi} on complex128 data, equals
    ad_intf_pc on complex64 data to complex64 rounding.  temp_coh64: consistent
    phases -> 1, equals temp_coh to float32 rounding."""
    rng = np.random.default_rng(9)
    ref = rng.standard_normal((7, 9)) + 1j * rng.standard_normal((7, 9))
    sec = ref * np.exp(1j * 1.3)
    idx = np.array([[0, 0], [3, 4], [6, 8]], np.int32)
    shp = rng.random((3, 3, 5)) < 0.7
    shp[:, 1, 2] = True
    np.testing.assert_allclose(R.ad_intf_pc64(ref, sec, idx, shp), np.exp(-1j * 1.3), atol=1e-14)
    r64, s64 = ref.astype(np.complex64), (rng.standard_normal((7, 9)) + 1j).astype(np.complex64)
    np.testing.assert_allclose(R.ad_intf_pc64(r64, s64, idx, shp), R.ad_intf_pc(r64, s64, idx, shp), atol=2e-7)
    n = 7
    th = np.exp(1j * rng.uniform(-np.pi, np.pi, (5, n)))
    pairs = R.image_pairs_from_bandwidth(n)
    f = 2.5 * th[:, pairs[:, 0]] * np.conj(th[:, pairs[:, 1]])
    np.testing.assert_allclose(R.temp_coh64(f, 3 * th), 1.0, atol=1e-14)
    f2 = (rng.standard_normal(f.shape) + 1j * rng.standard_normal(f.shape)).astype(np.complex64)
    th2 = th.astype(np.complex64)
    np.testing.assert_allclose(R.temp_coh64(f2, th2), R.temp_coh(f2, th2), atol=2e-6)


def test_nearest_pd_higham_properties():
    """nearestPD (moraine/co.py:429-463; its wherePD is undefined, read as isPD):
    pinned by Higham's theorem.  For symmetric A = Q L Q^T the nearest PSD matrix
    in the Frobenius norm is Q max(L, 0) Q^T; the result adds at most a shift of
    the order of the spacing and is PD; PD input comes back unchanged; no random
    PSD matrix is closer; non-symmetric input gives a symmetric PD result."""
    rng = np.random.default_rng(4)
    n = 12
    X = rng.standard_normal((5, n, n))
    S = (X + np.swapaxes(X, -1, -2)) / 2  # symmetric indefinite
    assert not R.isPD(S).any()
    out = R.nearestPD(S)
    assert R.isPD(out).all()
    np.testing.assert_array_equal(out, np.swapaxes(out, -1, -2))
    lam, Q = np.linalg.eigh(S)
    clip = Q @ (np.maximum(lam, 0)[..., :, None] * np.swapaxes(Q, -1, -2))
    nrm = np.linalg.norm(S, axis=(-2, -1))
    assert (np.linalg.norm(out - clip, axis=(-2, -1)) < 1e-12 * nrm).all()
    for _ in range(20):  # optimality against random PSD matrices
        Y = rng.standard_normal((5, n, n))
        P = Y @ np.swapaxes(Y, -1, -2) * rng.uniform(0.01, 1.0)
        assert (np.linalg.norm(S - out, axis=(-2, -1)) <= np.linalg.norm(S - P, axis=(-2, -1)) + 1e-12).all()
    # PD input: unchanged to rounding
    Apd = X @ np.swapaxes(X, -1, -2) + n * np.eye(n)
    assert R.isPD(Apd).all()
    np.testing.assert_allclose(R.nearestPD(Apd), Apd, rtol=1e-13, atol=1e-13 * np.abs(Apd).max())
    # non-symmetric: symmetric and PD; float32 keeps its dtype
    o2 = R.nearestPD(X.astype(np.float32))
    assert o2.dtype == np.float32 and R.isPD(o2).all()
    np.testing.assert_array_equal(o2, np.swapaxes(o2, -1, -2))
    # a planted mistake (the polar factor of B instead of A) is seen on non-symmetric input
    B = (X + np.swapaxes(X, -1, -2)) / 2
    _, s, Vh = np.linalg.svd(B)
    Hb = np.swapaxes(Vh, -1, -2) @ (s[..., :, None] * Vh)
    assert np.abs(R.nearestPD(X) - (B + Hb) / 2).max() > 1e-3
