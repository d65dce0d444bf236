"""GPU parity of the coherence kernels against the FP64 oracle, per entry.

* raster emperical_co (cov and coh, N x N per pixel) on every kernel path:
  single-tile MFMA (N <= 32), two-tile MFMA with mirrored off-diagonal blocks
  (33 <= N <= 64) and the VALU raster kernel (N > 64), with all-true and random
  SHP masks;
* the row range (row0, nrows) of mr_emperical_co through the C ABI: row bands
  computed separately equal the whole raster bit for bit;
* emperical_co_pc with all-SHP windows at the largest SHP counts of the
  15x21 (K = 315) and 21x41 (K = 861) windows.

Tolerances (written at each assertion): coherence entries satisfy |coh| <= 1,
so the per-entry bar is absolute, plus a relative bar for entries >= 0.1.
"""
import sys

import numpy as np
import pytest
import torch

from moraine_amd import _dev, _lib, synth
from oracle import moraine_ref as R

pytestmark = pytest.mark.gpu


def _case(msg):
    print(f"[gpu case] {msg}", flush=True)
    sys.stdout.flush()


def _mr():
    import moraine_amd as mr

    return mr


def _coh_errors(c, c64):
    c64 = np.asarray(c64)
    ok = np.isfinite(c64)
    err = np.abs(np.asarray(c, np.complex128) - c64)[ok]
    mag = np.abs(c64)[ok]
    big = mag >= 0.1
    rel = err[big] / mag[big]
    return err.max(), (rel.max() if rel.size else 0.0)


@pytest.mark.parametrize("n", [17, 33, 60, 64, 65, 120])
@pytest.mark.parametrize("mask", ["all", "random"])
def test_co_raster_paths_vs_fp64(n, mask):
    mr = _mr()
    nl, w, az, r = 9, 11, 3, 5
    z = synth.ds_stack(nl, w, n, seed=300 + n, ps_frac=0.1)
    rng = np.random.default_rng(n)
    if mask == "all":
        shp = np.ones((nl, w, 2 * az + 1, 2 * r + 1), bool)
    else:
        shp = rng.random((nl, w, 2 * az + 1, 2 * r + 1)) < 0.5
    _case(f"emperical_co raster n={n} {nl}x{w} window {2*az+1}x{2*r+1} mask={mask}")
    cov, coh = mr.emperical_co(z, shp)
    cov64, coh64 = R.emperical_co64(z, shp)
    ea, er = _coh_errors(coh, coh64)
    print(f"  coh per-entry: max abs {ea:.2e}, max rel (|coh| >= 0.1) {er:.2e}")
    # per entry: 2e-6 absolute (|coh| <= 1), 1e-5 relative where |coh| >= 0.1
    assert ea < 2e-6 and er < 1e-5
    # the diagonal of coh is 1 to FP32 rounding (its imaginary part is the FP16 hi+lo
    # split's residue, not exactly 0); off the diagonal both outputs are exactly
    # Hermitian (the mirrored blocks are written as conjugates)
    d = np.diagonal(coh, axis1=2, axis2=3)
    assert np.abs(d - 1).max() < 1e-6
    off = ~np.eye(n, dtype=bool)
    np.testing.assert_array_equal(coh[..., off], np.conj(np.swapaxes(coh, 2, 3))[..., off])
    np.testing.assert_array_equal(cov[..., off], np.conj(np.swapaxes(cov, 2, 3))[..., off])
    # cov = S / K: per entry relative to sqrt(cov_mm cov_jj) (the Cauchy-Schwarz bound of |cov_mj|)
    dg = np.real(np.diagonal(cov64, axis1=2, axis2=3))
    scale = np.sqrt(dg[..., :, None] * dg[..., None, :])
    ecov = np.abs(cov - cov64) / scale
    print(f"  cov per-entry error / sqrt(cov_mm cov_jj): max {np.nanmax(ecov):.2e}")
    assert np.nanmax(ecov) < 2e-6


@pytest.mark.parametrize("n,az,r", [(60, 7, 10), (120, 3, 4)])
def test_co_raster_row_range_equals_whole(n, az, r):
    """mr_emperical_co on row bands [row0, row0 + nrows) equals the whole raster."""
    nl, w = 13, 10
    z = torch.from_numpy(synth.ds_stack(nl, w, n, seed=7 + n, ps_frac=0.1)).cuda()
    shp = torch.from_numpy(np.random.default_rng(n).random((nl, w, 2 * az + 1, 2 * r + 1)) < 0.7).cuda()
    aw, rw = 2 * az + 1, 2 * r + 1
    _case(f"emperical_co row ranges n={n} {nl}x{w}")
    cov = torch.empty((nl, w, n, n), dtype=torch.complex64, device="cuda")
    coh = torch.empty_like(cov)
    _lib.call("mr_emperical_co", _dev.ptr(z), _dev.ptr(shp), nl, w, n, aw, rw, 0, nl, _dev.ptr(cov), _dev.ptr(coh),
              _dev.stream())
    for row0, nrows in [(0, 4), (4, 5), (9, 4), (12, 1)]:
        cv = torch.full((nrows, w, n, n), float("nan"), dtype=torch.complex64, device="cuda")
        ch = torch.full_like(cv, float("nan"))
        _lib.call("mr_emperical_co", _dev.ptr(z), _dev.ptr(shp), nl, w, n, aw, rw, row0, nrows, _dev.ptr(cv),
                  _dev.ptr(ch), _dev.stream())
        torch.cuda.synchronize()
        np.testing.assert_array_equal(ch.cpu().numpy().view(np.uint64),
                                      coh[row0:row0 + nrows].cpu().numpy().view(np.uint64))
        np.testing.assert_array_equal(cv.cpu().numpy().view(np.uint64),
                                      cov[row0:row0 + nrows].cpu().numpy().view(np.uint64))
    # cov may be NULL: coh alone is unchanged
    ch = torch.empty((3, w, n, n), dtype=torch.complex64, device="cuda")
    _lib.call("mr_emperical_co", _dev.ptr(z), _dev.ptr(shp), nl, w, n, aw, rw, 5, 3, None, _dev.ptr(ch),
              _dev.stream())
    torch.cuda.synchronize()
    np.testing.assert_array_equal(ch.cpu().numpy().view(np.uint64), coh[5:8].cpu().numpy().view(np.uint64))


def test_co_raster_row_range_rejects_bad_range():
    with pytest.raises(ValueError):
        _lib.call("mr_emperical_co", None, None, 10, 10, 17, 3, 5, 8, 3, None, None, None)


@pytest.mark.parametrize("n,az,r,k", [(60, 7, 10, 315), (120, 10, 20, 861)])
def test_co_pc_all_shp_largest_k_per_entry(n, az, r, k):
    """All-SHP windows in the raster interior: K = W at 15x21 (315) and 21x41 (861)."""
    mr = _mr()
    nl, w = 2 * az + 6, 2 * r + 8
    z = synth.ds_stack(nl, w, n, seed=k, ps_frac=0.1)
    idx = np.array([[a, c] for a in range(az, nl - az) for c in range(r, w - r)], np.int32)
    shp = np.ones((idx.shape[0], 2 * az + 1, 2 * r + 1), bool)
    assert shp[0].sum() == k
    _case(f"emperical_co_pc all-SHP K={k} n={n} points={idx.shape[0]}")
    c = mr.emperical_co_pc(z, idx, shp)
    c64 = R.emperical_co_pc64(z, idx, shp)
    ea, er = _coh_errors(c, c64)
    print(f"  K={k}: per-entry max abs {ea:.2e}, max rel (|coh| >= 0.1) {er:.2e}")
    # per entry: 4e-6 absolute, 2e-5 relative where |coh| >= 0.1 (FP16 hi+lo products summed in FP32 over 2K slots)
    assert ea < 4e-6 and er < 2e-5


@pytest.mark.parametrize("n,nl,w,first,last", [(60, 48, 37, 23, 5), (17, 40, 64, 0, 0), (90, 30, 41, 7, 11),
                                               (33, 9, 300, 5, 0)])
def test_co_pc_point_order_does_not_change_results(n, nl, w, first, last):
    """The point-cloud kernels visit a run of consecutive raster pixels in
    strips (a run of at least two raster rows per XCD range) and any other
    point list in list order.  Each point's coherence is computed alone, so a
    dense run starting and ending mid-row and the same points shuffled agree bit
    for bit (the last case is shorter than two rows per range: list order)."""
    mr = _mr()
    z = synth.ds_stack(nl, w, n, seed=n + w, ps_frac=0.1)
    rng = np.random.default_rng(w)
    px = np.arange(first, nl * w - last)
    idx = np.stack([px // w, px % w], -1).astype(np.int32)
    shp = rng.random((px.size, 7, 9)) < 0.5
    shp[:, 3, 4] = True
    perm = rng.permutation(px.size)
    _case(f"emperical_co_pc point order n={n} raster {nl}x{w} points={px.size}")
    dense = np.asarray(mr.emperical_co_pc(z, idx, shp))
    shuffled = np.asarray(mr.emperical_co_pc(z, idx[perm], shp[perm]))
    np.testing.assert_array_equal(dense[perm].view(np.uint32), shuffled.view(np.uint32))
    sub = rng.choice(px.size, 64, replace=False)
    c64 = R.emperical_co_pc64(z, idx[sub], shp[sub])
    ea, er = _coh_errors(dense[sub], c64)
    # per entry: 4e-6 absolute, 2e-5 relative where |coh| >= 0.1
    assert ea < 4e-6 and er < 2e-5
