"""CPU-side checks of the product: the C ABI library loads and exports every
declared symbol, host-side tables and pair logic match the oracle/reference.
No compute call touches a GPU here."""
import ctypes
import json
import os
import re

import numpy as np
import pytest

import moraine_amd as mr
from moraine_amd import _lib
from oracle import moraine_ref as R

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PRINTED = json.load(open(os.path.join(ROOT, "tests", "golden", "printed.json")))


def _declared_symbols():
    src = open(os.path.join(ROOT, "include", "moraine_hip.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(mr_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    syms = _declared_symbols()
    assert len(syms) >= 14
    for s in syms:
        assert hasattr(lib, s), s
        assert s in _lib.SIGNATURES, f"{s} has no ctypes signature"
    assert lib.mr_abi_version() == 1


def test_ks_tables_bit_exact_with_reference_series():
    lib = _lib.load()
    for n in list(range(1, 65)) + [90, 120, 150, 200, 256]:
        p = np.empty(n + 1, np.float32)
        d = np.empty(n + 1, np.float32)
        s = np.empty(n + 1, np.uint8)
        rc = lib.mr_ks_tables_f32(n, 0.05, p.ctypes.data_as(ctypes.c_void_p), d.ctypes.data_as(ctypes.c_void_p),
                                  s.ctypes.data_as(ctypes.c_void_p))
        assert rc == 0
        pt, dt = R.ks_tables(n)
        np.testing.assert_array_equal(p.view(np.uint32), pt.view(np.uint32))
        np.testing.assert_array_equal(d.view(np.uint32), dt.view(np.uint32))
        np.testing.assert_array_equal(s.astype(bool), pt.astype(np.float64) < 0.05)


def test_ks_tables_reject_bad_n():
    lib = _lib.load()
    assert lib.mr_ks_tables_f32(0, 0.05, None, None, None) == 1
    assert lib.mr_ks_tables_f32(257, 0.05, None, None, None) == 1
    assert b"outside" in lib.mr_last_error_string()


def test_argument_errors_raise_without_gpu():
    # shape errors are detected on the host before any device work
    with pytest.raises(ValueError):
        _lib.call("mr_emi", None, 10, 1, 0, None, None, None, 0, None)  # n < 2
    with pytest.raises(ValueError):
        _lib.call("mr_emi", None, 10, 17, 17, None, None, None, 0, None)  # ref out of range
    with pytest.raises(NotImplementedError):
        _lib.call("mr_ks_test_f32", ctypes.c_void_p(1), 4, 4, 300, 1, 1, 0.05, None, None, None, None,
                  ctypes.c_void_p(1), 1 << 30, None)


def test_tempnet_matches_printed_and_oracle():
    np.testing.assert_array_equal(mr.TempNet.from_bandwidth(5, 2).image_pairs, PRINTED["tnet_from_bandwidth_5_2"])
    t = mr.TempNet(np.array(PRINTED["tnet_custom"]))
    assert t.image_pairs_idx(ref=1, sec=2) == PRINTED["tnet_custom_idx_1_2"]
    np.testing.assert_array_equal(t.image_pairs_idx(ref=[1, 2], sec=[2, 3]), PRINTED["tnet_custom_idx_12_23"])
    assert mr.nimage_from_npair(6) == PRINTED["nimage_from_npair_6"]
    for n in (2, 17, 60, 200):
        np.testing.assert_array_equal(mr.TempNet.from_bandwidth(n).image_pairs, R.image_pairs_from_bandwidth(n))
        assert mr.nimage_from_npair(n * (n - 1) // 2) == n
    with pytest.raises(ValueError):
        mr.nimage_from_npair(7)
    # every pair count up to n = 2000: triangular numbers map to their n, the rest raise
    tri = {n * (n - 1) // 2: n for n in range(2, 2001)}
    for npair in range(1, 1999 * 2000 // 2 + 1, 97):
        if npair in tri:
            assert mr.nimage_from_npair(npair) == tri[npair]
        else:
            with pytest.raises(ValueError):
                mr.nimage_from_npair(npair)
    assert all(mr.nimage_from_npair(p) == n for p, n in tri.items())
    assert mr.nimage_from_npair(0) == 0  # the reference's guess loop accepts 0 images first
    with pytest.raises(ValueError):
        mr.TempNet(np.array([[0, 1], [0, 1], [2, 1]]))
    with pytest.raises(ValueError):
        mr.TempNet(np.array([[0, 2], [1, 2]]))


def test_api_signatures_match_reference_names():
    import inspect

    sig = {
        "ks_test": ["rmli", "az_half_win", "r_half_win", "block_size", "return_dist"],
        "select_shp": ["p", "p_max"],
        "emperical_co": ["rslc", "is_shp", "block_size"],
        "emperical_co_pc": ["rslc", "idx", "pc_is_shp", "block_size", "image_pairs"],
        "emi": ["coh", "ref"],
        "ds_temp_coh": ["coh", "ph", "image_pairs", "block_size"],
        "temp_coh": ["intf", "rslc", "image_pairs", "block_size"],
        "emperical_co_emi_temp_coh_pc": ["rslc", "idx", "pc_is_shp", "batch_size"],
        "uncompress_coh": ["coh", "image_pairs"],
        "rslc2amp": ["rslc"],
        "amp_disp": ["rslc"],
        "ad_intf_pc": ["ref_rslc", "sec_rslc", "idx", "pc_is_shp", "block_size"],
        "regularize_spectral": ["coh", "beta"],
    }
    for name, params in sig.items():
        ps = inspect.signature(getattr(mr, name)).parameters
        # the reference's parameters in order; extras only keyword-only with a default
        assert [k for k, v in ps.items() if v.kind != v.KEYWORD_ONLY] == params, name
        for k, v in ps.items():
            if v.kind == v.KEYWORD_ONLY:
                assert v.default is not inspect.Parameter.empty, (name, k)


def test_refine_argument_rules():
    # the precision form's argument checks run on the host before any GPU call
    from moraine_amd.pl import _check_refine

    assert _check_refine(None, 60) is None
    assert _check_refine(1e-3, 60) == 1e-3
    assert _check_refine(2e-4, 33) == 2e-4 and _check_refine(2e-4, 64) == 2e-4
    for bad in (0.0, -1e-3, float("nan")):
        with pytest.raises(ValueError):
            _check_refine(bad, 60)
    assert _check_refine(1e-3, 2) == 1e-3 and _check_refine(1e-3, 200) == 1e-3
    assert _check_refine(1e-3, 201) is None  # FP64 throughout there


def test_noalloc_ks_rejects_oversized_stack_on_host():
    """mr_ks_test_f32_noalloc checks the stack against its static scratch before any device work."""
    lib = _lib.load()
    rc = lib.mr_ks_test_f32_noalloc(None, 2048, 2048, 60, 7, 10, 0.05, None, None, None, None, None)
    assert rc == 3
    assert b"static scratch" in lib.mr_last_error_string()


def test_pair_indices_follow_numpy_rules():
    from moraine_amd.co import _pairs

    got = _pairs(np.array([[0, 1], [-3, -1], [2, 3]]), 4)
    np.testing.assert_array_equal(got, [[0, 1], [1, 3], [2, 3]])
    assert got.dtype == np.int32
    with pytest.raises(IndexError):
        _pairs(np.array([[0, 4]]), 4)
    with pytest.raises(IndexError):
        _pairs(np.array([[-5, 1]]), 4)
    with pytest.raises(ValueError):
        _pairs(np.array([0, 1, 2]), 4)


def test_batch_points_respects_hbm_budget():
    """The batch driver's batch: at least 65536 points, never more coherence
    bytes than a quarter of the free HBM, at least one point."""
    from moraine_amd.pl import BATCH_HBM_SHARE, MIN_BATCH_POINTS, batch_points

    gib = 1 << 30
    for n in (17, 60, 200, 1000):
        for free in (1 * gib, 16 * gib, 250 * gib):
            for bsz in (1, 1000, 1 << 20):
                bs = batch_points(bsz, n, 10 ** 7, free)
                per = 16 * (n * (n - 1) // 2) + 16 * n
                assert bs >= 1
                assert bs == 1 or bs * per <= BATCH_HBM_SHARE * free
                assert bs >= min(max(bsz, MIN_BATCH_POINTS), int(BATCH_HBM_SHARE * free) // per, 10 ** 7) or bs == 1
    assert batch_points(1000, 60, 5000, 250 * gib) == 5000  # never more than the point count
    assert batch_points(1000, 1000, 10 ** 6, 16 * gib) == 1 * 16 * gib // 4 // (16 * 499500 + 16000)


def test_fused_driver_step_argument_checks():
    """mr_co_emi_ds_temp_coh_pc validates its arguments before touching the GPU."""
    call = lambda *a: _lib.call("mr_co_emi_ds_temp_coh_pc", *a)  # noqa: E731
    # (rslc, nl, w, n, idx, mask, n_pc, aw, rw, ref, ph, q, t, ws, ws_bytes, stream)
    with pytest.raises(ValueError, match="ref"):
        call(None, 4, 4, 60, None, None, 10, 3, 3, 60, None, None, None, None, 0, None)
    with pytest.raises(ValueError, match="window"):
        call(None, 4, 4, 60, None, None, 10, 65, 65, 0, None, None, None, None, 0, None)
    with pytest.raises(ValueError, match="empty raster"):
        call(None, 0, 4, 60, None, None, 10, 3, 3, 0, None, None, None, None, 0, None)
    with pytest.raises(ValueError, match="workspace"):
        x = ctypes.c_void_p(16)  # never dereferenced: the workspace check fails first
        call(x, 4, 4, 60, x, x, 10, 3, 3, 0, x, x, x, None, 0, None)
    assert _lib.call("mr_co_emi_ds_temp_coh_pc", None, 4, 4, 60, None, None, 0, 3, 3, 0, None, None, None, None, 0,
                     None) == 0  # no points: nothing to do
    P = 60 * 59 // 2
    need = _lib.query("mr_co_emi_ds_temp_coh_pc_workspace_bytes", 1000, 60)
    assert need >= 1000 * P * 8 + P * 8 + 1000 * 8
