"""GPU: one batch of emperical_co_emi_temp_coh_pc in one launch
(mr_co_emi_ds_temp_coh_pc) against the separate calls it replaces.

moraine's own test of the batch driver is exact equality with the composition
emperical_co_pc -> emi -> ds_temp_coh (nbs/API/pl.ipynb:704-714); the fused
launch computes the same coherence values into on-chip memory, so the bar here
is bit equality too.  The composition itself is checked against the FP64 oracle
in test_gpu_parity.py.
"""
import sys

import numpy as np
import pytest

from moraine_amd import synth

pytestmark = pytest.mark.gpu


def _case(msg):
    print(f"[gpu case] {msg}", flush=True)
    sys.stdout.flush()


def _composition(mr, z, idx, m):
    c = mr.emperical_co_pc(z, idx, m)
    ph, q = mr.emi(c)
    t = mr.ds_temp_coh(c, ph)
    return ph, q, t


def _assert_equal(got, want):
    for a, b in zip(got, want):
        a = np.asarray(a)
        b = np.asarray(b)
        assert a.shape == b.shape and a.dtype == b.dtype
        # NaN == NaN bit for bit: compare the raw words
        np.testing.assert_array_equal(a.view(np.uint32), b.view(np.uint32))


@pytest.mark.parametrize("n", [33, 45, 60, 61, 64])
def test_fused_step_all_pixels_in_raster_order(n):
    """Every pixel of a 48 x 64 raster in raster order (the XCD ranges then take
    the strip order), KS-selected SHPs of 7 x 11 windows: some points have K < N
    and go through the Householder pass from the rows the fused launch wrote."""
    import moraine_amd as mr

    _case(f"fused co+emi+tcoh, all pixels, N={n}")
    z = synth.ds_stack(48, 64, n, seed=31 + n, ps_frac=0.1)
    rmli = (np.abs(z) ** 2).astype(np.float32)
    _, is_shp, shp_num = mr.ks_test_select_shp(rmli, 3, 5, 0.05, return_p=False)
    ii, jj = np.meshgrid(np.arange(48), np.arange(64), indexing="ij")
    idx = np.stack([ii.ravel(), jj.ravel()], -1).astype(np.int32)
    m = is_shp.reshape(-1, 7, 11)
    got = mr.emperical_co_emi_temp_coh_pc(z, idx, m)
    hh = int(mr.pl.last_householder.item())
    _assert_equal(got, _composition(mr, z, idx, m))
    assert hh > 0, "no point took the Householder pass: the decline path is untested"


@pytest.mark.parametrize("n", [17, 40, 60, 90])
def test_fused_step_c_abi_random_points(n):
    """Points in random order (identity point order), random masks with empty and
    full windows, windows hanging over every raster edge; N = 17 and 90 run the
    staged calls inside the same entry point."""
    import torch

    import moraine_amd as mr
    from moraine_amd.pl import co_emi_tcoh_dev

    _case(f"fused co+emi+tcoh via the C ABI, random points, N={n}")
    rng = np.random.default_rng(100 + n)
    nl, w = 20, 30
    z = synth.ds_stack(nl, w, n, seed=n, ps_frac=0.1)
    npt = 700
    idx = np.stack([rng.integers(0, nl, npt), rng.integers(0, w, npt)], -1).astype(np.int32)
    m = rng.random((npt, 9, 13)) < rng.uniform(0.05, 1.0, (npt, 1, 1))
    m[:5] = False  # no SHP: NaN rows, as the composition gives
    m[5:10] = True
    zt = torch.from_numpy(z).cuda()
    it = torch.from_numpy(idx).cuda()
    mt = torch.from_numpy(m).cuda()
    ph = torch.empty((npt, n), dtype=torch.complex64, device="cuda")
    q = torch.empty(npt, dtype=torch.float32, device="cuda")
    t = torch.empty(npt, dtype=torch.float32, device="cuda")
    co_emi_tcoh_dev(zt, it, mt, 3, ph, q, t)
    c = mr.emperical_co_pc(zt, it, mt)
    ph2, q2 = mr.emi(c, ref=3)
    t2 = mr.ds_temp_coh(c, ph2)
    _assert_equal((ph.cpu().numpy(), q.cpu().numpy(), t.cpu().numpy()),
                  (ph2.cpu().numpy(), q2.cpu().numpy(), t2.cpu().numpy()))


def test_fused_step_rejects_bad_arguments():
    import moraine_amd as mr
    from moraine_amd import _lib

    _case("fused co+emi+tcoh argument checks")
    with pytest.raises(ValueError):
        _lib.call("mr_co_emi_ds_temp_coh_pc", None, 4, 4, 60, None, None, 10, 3, 3, 60, None, None, None, None, 0,
                  None)
    with pytest.raises(ValueError):
        _lib.call("mr_co_emi_ds_temp_coh_pc", None, 4, 4, 60, None, None, 10, 65, 65, 0, None, None, None, None, 0,
                  None)
    assert mr is not None


def test_fused_batches_on_two_streams_equal_one_call():
    """Batches alternating over the two side streams (each batch's tail overlapping
    the next batch's launch) give the outputs of one call over every point."""
    import torch

    import moraine_amd as mr
    from moraine_amd.pl import co_emi_tcoh_batches, co_emi_tcoh_dev

    _case("fused co+emi+tcoh, batches on two streams, N=60")
    n = 60
    z = synth.ds_stack(40, 50, n, seed=77, ps_frac=0.1)
    rmli = (np.abs(z) ** 2).astype(np.float32)
    _, is_shp, _ = mr.ks_test_select_shp(rmli, 3, 5, 0.05, return_p=False)
    zt = torch.from_numpy(z).cuda()
    ii, jj = torch.meshgrid(torch.arange(40, device="cuda", dtype=torch.int32),
                            torch.arange(50, device="cuda", dtype=torch.int32), indexing="ij")
    it = torch.stack([ii.reshape(-1), jj.reshape(-1)], -1).contiguous()
    mt = torch.from_numpy(is_shp.reshape(-1, 7, 11)).cuda()
    npt = it.shape[0]
    outs = []
    for bs in (npt, 333):
        ph = torch.full((npt, n), float("nan"), dtype=torch.complex64, device="cuda")
        q = torch.full((npt,), float("nan"), device="cuda")
        t = torch.full((npt,), float("nan"), device="cuda")
        if bs == npt:
            co_emi_tcoh_dev(zt, it, mt, 0, ph, q, t)
        else:
            co_emi_tcoh_batches(zt, it, mt, 0, ph, q, t, bs)
        outs.append((ph.cpu().numpy(), q.cpu().numpy(), t.cpu().numpy()))
    _assert_equal(outs[1], outs[0])
