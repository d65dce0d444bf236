"""GPU parity: HIP kernels (through the C ABI) against the oracle.

Bit-exact for the KS test (p, dist, SHP mask, SHP count); stated tolerances for
floating-point stages, written next to each assertion.
"""
import os
import sys

import numpy as np
import pytest

from moraine_amd import synth
from oracle import moraine_ref as R

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")


def _case(msg):
    print(f"[gpu case] {msg}", flush=True)
    sys.stdout.flush()


def _mr():
    import moraine_amd as mr

    return mr


def _bits_equal(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    assert a.shape == b.shape
    if a.dtype == np.float32:
        np.testing.assert_array_equal(a.view(np.uint32), b.view(np.uint32))
    elif a.dtype == np.float64:
        np.testing.assert_array_equal(a.view(np.uint64), b.view(np.uint64))
    else:
        np.testing.assert_array_equal(a, b)


# ---------------------------------------------------------------------------- KS
def test_ks_200x200x17_11x15_bit_exact(port):
    """200x200x17 DS stack with 4 brightness regions, 11x15 window, p=0.05."""
    mr = _mr()
    z = synth.ds_stack(200, 200, 17, seed=1)
    rmli = (np.abs(z) ** 2).astype(np.float32)
    _case("ks 200x200x17 az=5 r=7")
    d, p = mr.ks_test(rmli, 5, 7, return_dist=True)
    d_ref, p_ref = port.ks_test(rmli, 5, 7, return_dist=True)
    _bits_equal(p, p_ref)
    _bits_equal(d, d_ref)
    p2 = mr.ks_test(rmli, 5, 7)
    _bits_equal(p2, p_ref)
    pf, m, c = mr.ks_test_select_shp(rmli, 5, 7, 0.05)
    m_ref, c_ref = R.select_shp(p_ref, 0.05)
    _bits_equal(pf, p_ref)
    np.testing.assert_array_equal(m, m_ref)
    np.testing.assert_array_equal(c, c_ref)
    m2, c2 = mr.select_shp(p_ref, 0.05)
    np.testing.assert_array_equal(m2, m_ref)
    np.testing.assert_array_equal(c2, c_ref)


def test_ks_n60_15x21_model_subtile_bit_exact(port):
    """N=60 with 10% PS-like pixels (constant amplitude -> tie-heavy), 15x21 window."""
    mr = _mr()
    z = synth.ds_stack(48, 96, 60, seed=2, ps_frac=0.1)
    rmli = np.abs(z).astype(np.float32)
    _case("ks N=60 48x96 az=7 r=10")
    p = mr.ks_test(rmli, 7, 10)
    p_ref = port.ks_test(rmli, 7, 10)
    _bits_equal(p, p_ref)


@pytest.mark.parametrize("n", [1, 2, 20, 63, 64, 65, 128, 129, 200, 256])
def test_ks_sizes_and_ties(port, n):
    mr = _mr()
    rng = np.random.default_rng(n)
    rmli = rng.integers(0, 6, (7, 9, n)).astype(np.float32)  # heavy ties
    rmli[0, 0] = 0.0
    rmli[0, 1] = -0.0  # signed zero equals zero
    rmli[2, 3, n // 2] = np.nan
    rmli[5, 8, -1] = np.inf
    rmli[4, 4, 0] = -np.inf
    _case(f"ks ties n={n}")
    d, p = mr.ks_test(rmli, 2, 3, return_dist=True)
    d_ref, p_ref = R.ks_test(rmli, 2, 3, return_dist=True)
    _bits_equal(p, p_ref)
    _bits_equal(d, d_ref)


@pytest.mark.parametrize("dt", [np.float32, np.float64])
def test_ks_denormal_and_signed_values(port, dt):
    """The walk compares the values as the reference does: denormals stay
    distinct (no flush to zero), -0 ties +0, negatives order below them."""
    mr = _mr()
    rng = np.random.default_rng(11)
    tiny = np.finfo(dt).smallest_subnormal
    vals = np.array([0.0, -0.0, tiny, -tiny, 2 * tiny, -3 * tiny, 1e3 * tiny, np.finfo(dt).tiny, -1.0, 1.0], dt)
    rmli = vals[rng.integers(0, vals.size, (6, 8, 40))]
    _case(f"ks denormal and signed values, {np.dtype(dt).name}")
    d, p = mr.ks_test(rmli, 2, 3, return_dist=True)
    d_ref, p_ref = R.ks_test(rmli, 2, 3, return_dist=True)
    _bits_equal(p, p_ref)
    _bits_equal(d, d_ref)


def test_ks_edge_windows_and_f64(port):
    mr = _mr()
    rng = np.random.default_rng(7)
    rmli = rng.random((3, 4, 20)).astype(np.float32)
    for az, r in [(0, 0), (0, 3), (2, 0), (5, 9)]:
        _case(f"ks window az={az} r={r} (window exceeds image)")
        _bits_equal(mr.ks_test(rmli, az, r), R.ks_test(rmli, az, r))
    r64 = rng.random((6, 7, 33))
    _case("ks float64")
    d64, p64 = mr.ks_test(r64, 2, 2, return_dist=True)
    d_ref, p_ref = R.ks_test(r64, 2, 2, return_dist=True)
    assert p64.dtype == np.float64
    _bits_equal(p64, p_ref)
    _bits_equal(d64, d_ref)


def test_ks_printed_example():
    mr = _mr()
    rng = np.random.default_rng(0)
    while True:
        s1 = rng.uniform(size=20).astype(np.float32)
        s2 = rng.standard_normal(20).astype(np.float32)
        if R.ks_k_walk(np.sort(s1), np.sort(s2)) == 7:
            break
    rmli = np.stack((s1, s2)).reshape(1, 2, 20)
    _case("ks printed 1x2x20 example")
    d, p = mr.ks_test(rmli, 1, 1, return_dist=True)
    assert d[0, 0, 1, 2] == np.float32(0.35) and format(p[0, 0, 1, 2], ".8g") == "0.13494715"
    assert p[0, 0, 1, 1] == 0 and np.isnan(p[0, 0, 0, 0]) and np.isnan(p[0, 1, 1, 2])


def test_ks_torch_io():
    import torch

    mr = _mr()
    rmli = np.random.default_rng(5).random((10, 12, 30)).astype(np.float32)
    _case("ks torch in/out")
    pt = mr.ks_test(torch.from_numpy(rmli).cuda(), 3, 4)
    assert isinstance(pt, torch.Tensor) and pt.is_cuda
    _bits_equal(pt.cpu().numpy(), R.ks_test(rmli, 3, 4))


# ---------------------------------------------------------------------------- coherence
def test_co_small_vs_fp64():
    mr = _mr()
    g = np.load(os.path.join(GOLD, "co_small.npz"))
    _case("emperical_co 5x10x17 golden")
    cov, coh = mr.emperical_co(g["rslc"], g["is_shp"])
    # tolerance of the reference's own test (nbs/API/co.ipynb:462): 1e-6 absolute
    np.testing.assert_allclose(coh, g["coh64"], rtol=0, atol=1e-6)
    np.testing.assert_allclose(cov, g["cov64"], rtol=1e-5, atol=1e-6)
    c_pc = mr.emperical_co_pc(g["rslc"], g["idx"], g["is_shp"].reshape(-1, 3, 5))
    iu = np.triu_indices(17, 1)
    # same accumulation order in both kernels: the upper triangle is identical
    np.testing.assert_array_equal(coh.reshape(-1, 17, 17)[:, iu[0], iu[1]], c_pc)
    np.testing.assert_allclose(c_pc, g["coh_pc"], rtol=0, atol=1e-6)


@pytest.mark.parametrize("n,az,r", [(60, 7, 10), (30, 5, 7), (90, 5, 7), (120, 4, 6), (128, 3, 4), (129, 3, 4),
                                    (150, 5, 7), (192, 3, 4), (200, 3, 4), (224, 2, 3), (225, 2, 3)])
def test_co_pc_large_k_vs_fp64(n, az, r):
    mr = _mr()
    z = synth.ds_stack(20, 30, n, seed=n, ps_frac=0.1)
    rng = np.random.default_rng(n)
    shp = rng.random((40, 2 * az + 1, 2 * r + 1)) < 0.6
    idx = np.stack([rng.integers(0, 20, 40), rng.integers(0, 30, 40)], -1).astype(np.int32)
    _case(f"emperical_co_pc n={n} window {2*az+1}x{2*r+1}")
    c = mr.emperical_co_pc(z, idx, shp)
    c64 = R.emperical_co_pc64(z, idx, shp)
    # FP32 accumulation over up to 315 SHPs: 1e-5 relative to the largest |coh| per point
    err = np.max(np.abs(c - c64), axis=1) / np.max(np.abs(c64), axis=1)
    assert err.max() < 1e-5
    bw = R.image_pairs_from_bandwidth(n, 3)
    cb = mr.emperical_co_pc(z, idx, shp, image_pairs=bw)
    c64b = R.emperical_co_pc64(z, idx, shp, bw)
    np.testing.assert_allclose(cb, c64b, rtol=0, atol=1e-5)


def test_ds_candidates_bandwidth_network_gpu():
    """DS-candidate point cloud (shp_num >= 50) with a bandwidth-limited pair
    network: coherence and DS temporal coherence on that network against the
    oracle (phases from EMI on the full network)."""
    mr = _mr()
    n, az, r = 30, 7, 10
    z = synth.ds_stack(36, 44, n, seed=21, ps_frac=0.1)
    rmli = (np.abs(z) ** 2).astype(np.float32)
    _, shp, cnt = mr.ks_test_select_shp(rmli, az, r, 0.05, return_p=False)
    idx = np.argwhere(cnt >= 50).astype(np.int32)
    assert 50 < idx.shape[0] < cnt.size
    pc = shp[idx[:, 0], idx[:, 1]]
    _case(f"DS candidates {idx.shape[0]} points, bandwidth-5 network, n={n}")
    ph, _ = mr.emi(mr.emperical_co_pc(z, idx, pc))
    bw = R.image_pairs_from_bandwidth(n, 5)
    cb = mr.emperical_co_pc(z, idx, pc, image_pairs=bw)
    np.testing.assert_allclose(cb, R.emperical_co_pc64(z, idx, pc, bw), rtol=0, atol=1e-5)
    t = mr.ds_temp_coh(cb, ph, image_pairs=bw)
    t_ref = R.ds_temp_coh(cb, ph, bw)
    # 1e-5 relative (float32 accumulation order differs)
    np.testing.assert_allclose(t, t_ref, rtol=1e-5, atol=1e-6)


def test_co_zero_shp_is_nan():
    mr = _mr()
    z = synth.ds_stack(4, 4, 17, seed=3)
    shp = np.zeros((2, 3, 3), bool)
    idx = np.array([[1, 1], [2, 2]], np.int32)
    _case("emperical_co_pc zero SHPs")
    c = mr.emperical_co_pc(z, idx, shp)
    assert np.all(np.isnan(c))


def test_rslc2amp_gpu():
    """Amplitude of random bit patterns, specials and denormals, bit for bit with
    the oracle (the C library's hypotf)."""
    mr = _mr()
    rng = np.random.default_rng(11)
    xy = rng.integers(0, 2 ** 32, size=(200000, 2), dtype=np.uint64).astype(np.uint32).view(np.float32)
    sp = np.array([0.0, -0.0, np.inf, -np.inf, np.nan, 1e-45, 1.17e-38, 3.4e38, 1.0, 3.0, 4.0], np.float32)
    X, Y = np.meshgrid(sp, sp)
    xy = np.concatenate([xy, np.stack([X.ravel(), Y.ravel()], -1), rng.standard_normal((9999, 2)).astype(np.float32)])
    z = np.empty(xy.shape[0], np.complex64)
    z.real[:] = xy[:, 0]
    z.imag[:] = xy[:, 1]
    _case("rslc2amp 210120 samples")
    got = mr.rslc2amp(z.reshape(-1, 5, 4))
    want = R.rslc2amp(z).reshape(-1, 5, 4)
    assert got.shape == want.shape and got.dtype == np.float32
    same = (got.view(np.uint32) == want.view(np.uint32)) | (np.isnan(got) & np.isnan(want))
    assert same.all()


@pytest.mark.parametrize("n", [17, 60, 130])
def test_ks_from_complex_stack_equals_composition_gpu(n):
    """ks_test / ks_test_select_shp on a complex64 stack equal the same calls on
    rslc2amp of it, bit for bit (constant-amplitude PS pixels give ties; NaN samples)."""
    import torch

    mr = _mr()
    z = synth.ds_stack_torch(40, 56, n, seed=n, ps_frac=0.1, device="cuda")
    z[3, 5, 0] = complex(float("nan"), 0.0)
    z[7, 9, n - 1] = complex(float("inf"), float("nan"))
    _case(f"ks from complex stack n={n}")
    amp = mr.rslc2amp(z)
    d1, p1 = mr.ks_test(z, 5, 7, return_dist=True)
    d2, p2 = mr.ks_test(amp, 5, 7, return_dist=True)
    _bits_equal(p1.cpu().numpy(), p2.cpu().numpy())
    _bits_equal(d1.cpu().numpy(), d2.cpu().numpy())
    q1, m1, c1 = mr.ks_test_select_shp(z, 5, 7, 0.05)
    q2, m2, c2 = mr.ks_test_select_shp(amp, 5, 7, 0.05)
    _bits_equal(q1.cpu().numpy(), q2.cpu().numpy())
    assert torch.equal(m1, m2) and torch.equal(c1, c2)
    assert torch.isnan(p1[3, 5]).all()


def test_amp_disp_gpu():
    """Amplitude dispersion vs the oracle on the DS/PS model with NaN samples
    (1e-6 relative: the reference sums in FP32, both sides here in FP64)."""
    mr = _mr()
    z = synth.ds_stack(33, 47, 60, seed=4, ps_frac=0.1)
    z[2, 3, 7] = np.nan
    _case("amp_disp 33x47x60")
    d = mr.amp_disp(z)
    want = R.amp_disp(z)
    assert np.isnan(d[2, 3]) and np.isnan(want[2, 3])
    np.testing.assert_allclose(d, want, rtol=1e-6, atol=1e-7)


def test_ad_intf_pc_gpu():
    """Adaptive multilooked interferogram vs the oracle: random SHP masks, points
    at the raster edges, one point without SHPs (NaN)."""
    mr = _mr()
    z = synth.ds_stack(30, 40, 12, seed=8, ps_frac=0.1)
    rng = np.random.default_rng(8)
    idx = np.concatenate([np.stack([rng.integers(0, 30, 200), rng.integers(0, 40, 200)], -1),
                          [[0, 0], [29, 39], [0, 39], [15, 20]]]).astype(np.int32)
    shp = rng.random((idx.shape[0], 15, 21)) < 0.5
    shp[-1] = False
    _case("ad_intf_pc 204 points 15x21")
    got = mr.ad_intf_pc(z[..., 2], z[..., 7], idx, shp)
    want = R.ad_intf_pc(z[..., 2], z[..., 7], idx, shp)
    assert np.isnan(got[-1].real) and np.isnan(want[-1].real)
    # FP64 sums on both sides, complex64 out
    np.testing.assert_allclose(got[:-1], want[:-1], rtol=0, atol=1e-6)


def test_regularize_spectral_gpu():
    """Scalar and per-matrix beta vs the oracle (FP64 arithmetic, complex64 out:
    one float32 rounding)."""
    mr = _mr()
    rng = np.random.default_rng(6)
    c = (rng.standard_normal((7, 3, 60, 60)) + 1j * rng.standard_normal((7, 3, 60, 60))).astype(np.complex64)
    b = rng.random((7, 3))
    _case("regularize_spectral 21 x 60 x 60")
    for beta in (0.25, b):
        got = mr.regularize_spectral(c, beta)
        want = R.regularize_spectral(c, beta).astype(np.complex64)
        assert got.dtype == np.complex64 and got.shape == c.shape
        np.testing.assert_array_equal(got, want)


def test_neighbours_empty_inputs_gpu():
    """Empty stacks, point lists and matrix stacks give empty outputs of the right shape and dtype."""
    mr = _mr()
    _case("neighbour functions on empty inputs")
    a = mr.rslc2amp(np.zeros((0, 4, 6), np.complex64))
    assert a.shape == (0, 4, 6) and a.dtype == np.float32
    d = mr.amp_disp(np.zeros((0, 5, 6), np.complex64))
    assert d.shape == (0, 5) and d.dtype == np.float32
    z = synth.ds_stack(6, 7, 5, seed=1)
    i = mr.ad_intf_pc(z[..., 0], z[..., 1], np.zeros((0, 2), np.int32), np.zeros((0, 3, 3), bool))
    assert i.shape == (0,) and i.dtype == np.complex64
    r = mr.regularize_spectral(np.zeros((0, 4, 4), np.complex64), 0.5)
    assert r.shape == (0, 4, 4)


def test_uncompress_gpu():
    mr = _mr()
    c = (np.random.default_rng(1).random((3, 4, 45)) + 1j).astype(np.complex64)
    _case("uncompress_coh")
    np.testing.assert_array_equal(mr.uncompress_coh(c), R.uncompress_coh(c))
    pairs = R.image_pairs_from_bandwidth(5, 2)
    c2 = np.random.default_rng(2).random((6, 7)).astype(np.complex64)
    np.testing.assert_array_equal(mr.uncompress_coh(c2, pairs), R.uncompress_coh(c2, pairs))


@pytest.mark.parametrize("n", [2, 31, 60, 128, 129, 200, 230])
def test_uncompress_sizes_gpu(n):
    # one row band of the LDS code table up to n = 128, several above
    mr = _mr()
    P = n * (n - 1) // 2
    rng = np.random.default_rng(n)
    c = (rng.standard_normal((3, P)) + 1j * rng.standard_normal((3, P))).astype(np.complex64)
    _case(f"uncompress_coh n={n}")
    np.testing.assert_array_equal(mr.uncompress_coh(c), R.uncompress_coh(c))


@pytest.mark.parametrize("n", [5, 210])
def test_uncompress_colliding_pairs_gpu(n):
    # reversed, repeated, diagonal and negative-index pairs: later writes win,
    # conjugates after pairs, the unit diagonal last (moraine/co.py:296-298)
    mr = _mr()
    rng = np.random.default_rng(7)
    base = np.array([[0, 1], [1, 0], [0, 0], [0, 1], [2, -1], [n - 1, 2], [3, 3], [1, n - 1]], np.int32)
    extra = rng.integers(0, n, (40, 2)).astype(np.int32)
    pairs = np.concatenate([base, extra, [[0, n - 1]]]).astype(np.int32)
    c = (rng.standard_normal((4, pairs.shape[0])) + 1j * rng.standard_normal((4, pairs.shape[0]))).astype(np.complex64)
    _case(f"uncompress_coh colliding pairs n={n}")
    want = R.uncompress_coh(c, np.where(pairs < 0, pairs + n, pairs))
    np.testing.assert_array_equal(mr.uncompress_coh(c, pairs), want)
    with pytest.raises(IndexError):
        mr.uncompress_coh(c, np.concatenate([pairs[:-1], [[n, 0]], pairs[-1:]]))


# ---------------------------------------------------------------------------- EMI
@pytest.mark.parametrize("n", [2, 3, 17, 30, 32, 33, 59, 60, 61, 64, 65, 90, 96, 97, 120, 128, 129, 150, 160, 161, 200,
                               201])
def test_emi_closure_free(n):
    mr = _mr()
    coh, exp = synth.closure_free_coh(24, n, seed=n)
    _case(f"emi closure-free n={n}")
    ph, q = mr.emi(coh)
    err = np.abs(np.angle(ph * np.conj(exp)))
    # closed form: ph = theta conj(theta_ref)/|.|, quality 1 (coherence stored in c64)
    assert err.max() < 2e-4, err.max()
    np.testing.assert_allclose(q, 1.0, atol=2e-3)
    np.testing.assert_allclose(np.abs(ph), 1.0, atol=1e-5)


@pytest.mark.parametrize("n,win", [(17, (5, 7)), (30, (5, 7)), (57, (5, 7)), (60, (5, 7)), (61, (5, 7)),
                                   (64, (5, 7)), (90, (5, 7)), (90, (11, 15)),
                                   (120, (11, 15)), (150, (11, 15)), (160, (11, 15)),
                                   (200, (11, 15))])
def test_emi_ds_vs_fp64(n, win):
    mr = _mr()
    h, w = (6, 8) if win == (5, 7) else (12, 16)
    z = synth.ds_stack(h, w, n, seed=200 + n, ps_frac=0.1)
    rng = np.random.default_rng(n)
    # mixed SHP counts: rank-deficient (indefinite |Gamma|) and full-rank windows
    shp = rng.random((48, *win)) < rng.uniform(0.1, 1.0, (48, 1, 1))
    shp[:, win[0] // 2, win[1] // 2] = True
    ii, jj = np.nonzero(np.ones((h, w), bool))
    sel = np.linspace(0, ii.size - 1, 48).astype(int)
    idx = np.stack([ii[sel], jj[sel]], -1).astype(np.int32)
    c = R.emperical_co_pc(z, idx, shp)
    _case(f"emi DS n={n} window={win}")
    ph, q = mr.emi(c)
    ph64, q64 = R.emi64(c)
    e = np.angle(ph * np.conj(ph64))
    rms = np.sqrt(np.mean(e ** 2, axis=1))
    # linked-phase RMS error vs FP64 <= 1e-3 rad (median far below)
    assert np.nanmax(rms) < 1e-3, np.nanmax(rms)
    np.testing.assert_allclose(q, q64, rtol=1e-3, atol=1e-3 * np.abs(q64).max())
    # the reference's own arithmetic (float32 inv of |Gamma|, complex64 eigh) on the same points:
    # the GPU's per-point tail is no worse (floor 2e-6 rad: both below it is a tie)
    ph32, _ = R.emi_ref32(c)
    rms32 = np.sqrt(np.nanmean(np.angle(ph32 * np.conj(ph64)) ** 2, axis=1))
    g99, r99 = np.nanquantile(rms, 0.99), np.nanquantile(rms32, 0.99)
    print(f"  per-point phase RMS vs FP64: GPU p99 {g99:.2e} max {np.nanmax(rms):.2e}; "
          f"reference FP32 LAPACK p99 {r99:.2e} max {np.nanmax(rms32):.2e}")
    assert g99 <= max(r99, 2e-6) and np.nanmax(rms) <= max(np.nanmax(rms32), 2e-6)


def test_emi_golden_and_ref():
    mr = _mr()
    g = np.load(os.path.join(GOLD, "emi_small.npz"))
    for n in (17, 30, 60):
        _case(f"emi golden n={n}")
        ph, q = mr.emi(g[f"ds_coh_{n}"], ref=0)
        assert np.abs(np.angle(ph * np.conj(g[f"ds_ph64_{n}"]))).max() < 1e-3
        ph5, _ = mr.emi(g[f"ds_coh_{n}"], ref=5)
        np.testing.assert_allclose(ph5[:, 5], 1.0, atol=1e-6)
        # changing ref only rotates every phase by conj(ph[ref])
        rot = ph * np.conj(ph[:, 5:6])
        assert np.abs(np.angle(rot * np.conj(ph5))).max() < 1e-4


# ---------------------------------------------------------------------------- temporal coherence
def test_tcoh_vs_oracle_and_no_mutation():
    mr = _mr()
    g = np.load(os.path.join(GOLD, "tcoh_small.npz"))
    coh, ph = g["coh"].copy(), g["ph"].copy()
    _case("ds_temp_coh / temp_coh")
    t = mr.ds_temp_coh(coh, ph)
    np.testing.assert_allclose(t, g["t_ds"], rtol=1e-5, atol=1e-6)
    t2 = mr.temp_coh(coh, ph)
    np.testing.assert_allclose(t2, g["t_pq"], rtol=1e-5, atol=1e-6)
    np.testing.assert_array_equal(coh, g["coh"])
    np.testing.assert_array_equal(ph, g["ph"])
    ras = mr.temp_coh(coh.reshape(6, 6, -1), ph.reshape(6, 6, -1))
    assert ras.shape == (6, 6)
    np.testing.assert_allclose(ras.ravel(), t2, rtol=0, atol=0)


@pytest.mark.parametrize("n", [17, 45, 60, 64, 90])
def test_fused_batch_driver_equals_composition(n):
    """nbs/API/pl.ipynb:704-714: the batch driver equals the separate calls exactly
    (for 32 < N <= 64 EMI and the temporal coherence run as one kernel)."""
    mr = _mr()
    z = synth.ds_stack(12, 16, n, seed=9, ps_frac=0.1)
    rmli = (np.abs(z) ** 2).astype(np.float32)
    _case(f"fused batch driver, N={n}")
    _, is_shp, shp_num = mr.ks_test_select_shp(rmli, 2, 3, 0.05, return_p=False)
    cand = shp_num >= 3
    idx = np.stack(np.nonzero(cand), -1).astype(np.int32)
    pc_shp = is_shp[cand]
    ph, q, t = mr.emperical_co_emi_temp_coh_pc(z, idx, pc_shp, batch_size=50)
    c = mr.emperical_co_pc(z, idx, pc_shp)
    ph2, q2 = mr.emi(c)
    t2 = mr.ds_temp_coh(c, ph2)
    np.testing.assert_array_equal(ph, ph2)
    np.testing.assert_array_equal(q, q2)
    np.testing.assert_array_equal(t, t2)
    from moraine_amd.pl import emi_ds_temp_coh

    ph3, q3, t3 = emi_ds_temp_coh(c)
    np.testing.assert_array_equal(ph3, ph2)
    np.testing.assert_array_equal(q3, q2)
    np.testing.assert_array_equal(t3, t2)


def test_emi_on_ks_selected_shps():
    """Realistic SHP masks from the KS test (few SHPs -> near-singular |Gamma| at some
    points).  EMI is compared with FP64 on the same coherence input."""
    mr = _mr()
    z = synth.ds_stack(24, 32, 17, seed=4, ps_frac=0.1)
    rmli = (np.abs(z) ** 2).astype(np.float32)
    _case("emi on KS-selected SHPs, N=17")
    _, is_shp, shp_num = mr.ks_test_select_shp(rmli, 3, 4, 0.05, return_p=False)
    cand = shp_num >= 5
    idx = np.stack(np.nonzero(cand), -1).astype(np.int32)
    coh = mr.emperical_co_pc(z, idx, is_shp[cand])
    ph, q = mr.emi(coh)
    ph64, q64 = R.emi64(coh)
    rms = np.sqrt(np.nanmean(np.angle(ph * np.conj(ph64)) ** 2, axis=1))
    assert np.nanmax(rms) < 1e-3, np.nanmax(rms)


def test_precision_form_n120_large_window_mixed_coherence():
    """Precision-form conditions at test size: N=120, 21x41 windows, gamma0 ~ U[0.2, 0.9]
    plus near-singular pixels.  Coherence within 1e-5 relative of FP64 on the same SHP
    masks; EMI within 1e-3 rad RMS of FP64 EMI on the same coherence (per-pixel errors
    printed, near-singular pixels listed separately)."""
    mr = _mr()
    z, ns = synth.ds_stack_mixed(14, 20, 120, seed=31, near_singular_frac=0.1)
    rmli = (np.abs(z) ** 2).astype(np.float32)
    _case("precision form N=120 14x20 window 21x41")
    _, is_shp, shp_num = mr.ks_test_select_shp(rmli, 10, 20, 0.05, return_p=False)
    idx = np.stack(np.nonzero(np.ones((14, 20), bool)), -1).astype(np.int32)
    pc = is_shp.reshape(-1, 21, 41)
    c = mr.emperical_co_pc(z, idx, pc)
    c64 = R.emperical_co_pc64(z, idx, pc)
    scale = np.abs(c64).max(axis=1, keepdims=True)
    assert np.nanmax(np.abs(c - c64) / scale) < 1e-5
    ph, q = mr.emi(c)
    ph64, q64 = R.emi64(c.astype(np.complex128))
    rms = np.sqrt(np.nanmean(np.angle(ph * np.conj(ph64)) ** 2, axis=1))
    nsf = ns.ravel()
    print(f"phase RMS vs FP64: regular max {np.nanmax(rms[~nsf]):.2e} median {np.nanmedian(rms[~nsf]):.2e}; "
          f"near-singular ({nsf.sum()}) max {np.nanmax(rms[nsf]) if nsf.any() else 0:.2e}; "
          f"SHP count {shp_num.min()}..{shp_num.max()}")
    assert np.nanmax(rms) < 1e-3
