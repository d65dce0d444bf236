"""GPU tests of the batched DS form (csrc/batch.hip through the C ABI).

The batched path shares every arithmetic step with the drop-in kernels, so a
sample matrix made of a pixel's SHP stacks in window order must give exactly
the drop-in composition emperical_co_pc -> emi -> ds_temp_coh (bitwise,
NaN == NaN).  Synthetic ragged batches are checked against FP64.
"""
import sys

import numpy as np
import pytest

from moraine_amd import synth
from oracle import moraine_ref as R

pytestmark = pytest.mark.gpu


def _case(msg):
    print(f"[gpu case] {msg}", flush=True)
    sys.stdout.flush()


def _bits(a):
    a = np.ascontiguousarray(a)
    return a.view(np.uint32) if a.dtype in (np.float32, np.complex64) else a


def _plain_case(n, nl, w, az, r, seed, min_shp):
    import moraine_amd as mr
    from moraine_amd.batched import raster_sample_matrices

    z = synth.ds_stack(nl, w, n, seed=seed, ps_frac=0.1)
    rmli = (np.abs(z) ** 2).astype(np.float32)
    _, is_shp, shp_num = mr.ks_test_select_shp(rmli, az, r, 0.05, return_p=False)
    cand = shp_num >= min_shp
    idx = np.stack(np.nonzero(cand), -1).astype(np.int32)
    pc = is_shp[cand]
    c = mr.emperical_co_pc(z, idx, pc)
    ph, q = mr.emi(c)
    t = mr.ds_temp_coh(c, ph)
    return raster_sample_matrices(z, idx, pc), ph, q, t


def test_batched_equals_plain_composition_mixed_n():
    from moraine_amd.batched import ds_batch

    cases = [(17, 20, 24, 3, 4, 11, 3), (60, 16, 20, 3, 5, 12, 4), (90, 12, 16, 4, 5, 13, 4),
             (120, 10, 14, 3, 4, 14, 4), (150, 8, 10, 3, 4, 15, 4)]
    mats, exp = [], []
    for n, nl, w, az, r, seed, mins in cases:
        _case(f"plain composition N={n} {nl}x{w} window {2*az+1}x{2*r+1}")
        m, ph, q, t = _plain_case(n, nl, w, az, r, seed, mins)
        mats += m
        exp += [(ph[i], q[i], t[i]) for i in range(len(m))]
    order = np.random.default_rng(0).permutation(len(mats))
    _case(f"ds_batch mixed N, {len(mats)} pixels")
    ph, q, t = ds_batch([mats[i] for i in order])
    for j, i in enumerate(order):
        np.testing.assert_array_equal(_bits(ph[j]), _bits(exp[i][0]))
        np.testing.assert_array_equal(_bits(q[j]), _bits(exp[i][1]))
        np.testing.assert_array_equal(_bits(t[j]), _bits(exp[i][2]))


def test_batched_graph_replay_and_in_place_update():
    """All group launches in one HIP graph: every replay equals the eager run,
    with unrelated eager work between replays; new data written into the bound
    buffers is picked up by the next replay."""
    import torch

    from moraine_amd.batched import DSBatch, DSBatchRunner

    s, ro, ni, ns = synth.ds_samples_torch(600, n_choices=(17, 30, 60, 90, 150), k_range=(20, 200), seed=3,
                                           device="cuda")
    b = DSBatch.from_tensors(s, ro, ni, ns)
    _case("ds_batch eager reference")
    eager = DSBatchRunner(b)
    ph0, _, q0, t0 = [x.clone() for x in eager.run()]
    _case("ds_batch graph capture + 3 replays")
    r = DSBatchRunner(b)
    r.capture()
    for k in range(3):
        x = torch.randn(512, 512, device="cuda")
        (x @ x).sum().item()
        ph, _, q, t = r.replay()
        torch.cuda.synchronize()
        assert torch.equal(ph.view(torch.int64), ph0.view(torch.int64)), k
        assert torch.equal(q.view(torch.int32), q0.view(torch.int32)), k
        assert torch.equal(t.view(torch.int32), t0.view(torch.int32)), k
    _case("ds_batch replay after in-place data update")
    b.samples.mul_(torch.exp(1j * torch.tensor(0.3, device="cuda")).to(torch.complex64))
    ph1, _, q1, t1 = [x.clone() for x in eager.run()]
    ph, _, q, t = r.replay()
    torch.cuda.synchronize()
    assert torch.equal(ph.view(torch.int64), ph1.view(torch.int64))
    assert torch.equal(t.view(torch.int32), t1.view(torch.int32))


def test_batched_launch_shapes_do_not_change_results():
    import torch

    from moraine_amd.batched import AutoTuner, DSBatch, DSBatchRunner

    s, ro, ni, ns = synth.ds_samples_torch(3000, n_choices=(30, 45, 90), k_range=(20, 120), seed=4, device="cuda")
    b = DSBatch.from_tensors(s, ro, ni, ns)
    _case("ds_batch default shapes")
    ref = [x.clone() for x in DSBatchRunner(b).run()]
    _case("ds_batch autotuned shapes")
    tuner = AutoTuner()
    r = DSBatchRunner(b, tuner=tuner)
    out = r.run()
    torch.cuda.synchronize()
    for g in r.groups:
        # every group (N = 30, 45, 90) was timed fused and staged
        k = AutoTuner.key("fused_or_staged", g.n, g.count)
        assert any(v < 0 for v in tuner.trials[k]) and any(v >= 0 for v in tuner.trials[k])
        if not g.staged:
            assert g.blocks in AutoTuner.fused_blocks(g.n)
        else:
            assert g.chunk in AutoTuner.CHUNKS or g.chunk == g.count
    # a staged N = 90 group also picks its coherence kernel's tile layout
    g90 = r.groups[-1]
    assert (AutoTuner.key("coh_waves", 90, g90.count) in tuner.picks) == g90.staged
    assert len(tuner.picks) == 3 + int(g90.staged)
    assert torch.equal(out[0].view(torch.int64), ref[0].view(torch.int64))
    assert torch.equal(out[2].view(torch.int32), ref[2].view(torch.int32))
    assert torch.equal(out[3].view(torch.int32), ref[3].view(torch.int32))


def test_fused_equals_per_stage():
    """The fused one-wave-per-pixel launch and the per-stage sequence
    (coherence to HBM, EMI, temporal coherence) give identical outputs."""
    import torch

    from moraine_amd.batched import DSBatch, DSBatchRunner

    s, ro, ni, ns = synth.ds_samples_torch(4000, n_choices=(17, 30, 45, 60, 90), k_range=(20, 315), seed=5,
                                           device="cuda")
    b = DSBatch.from_tensors(s, ro, ni, ns)
    _case("ds_batch fused vs per-stage, 4000 pixels")
    a = [x.clone() for x in DSBatchRunner(b).run()]
    c = DSBatchRunner(b, per_stage=True).run()
    torch.cuda.synchronize()
    assert torch.equal(a[0].view(torch.int64), c[0].view(torch.int64))
    assert torch.equal(a[2].view(torch.int32), c[2].view(torch.int32))
    assert torch.equal(a[3].view(torch.int32), c[3].view(torch.int32))


def test_fused_small_and_wide_equal_per_stage():
    """The fused launches for N <= 32 (one wave per pixel, coherence in an LDS
    tile) and 64 < N <= 200 (one workgroup per pixel, coherence row in the
    workgroup's slot) equal the per-stage sequence bit for bit, at every tile
    grid and EMI register layout (N at each range's ends) and at grids small
    enough that each workgroup takes many pixels (its slot re-used)."""
    import torch

    from moraine_amd.batched import DSBatch, DSBatchRunner

    ns_ = (2, 17, 32, 65, 96, 97, 128, 129, 160, 161, 192, 193, 200)
    s, ro, ni, ns = synth.ds_samples_torch(1300, n_choices=ns_, k_range=(20, 300), seed=21, device="cuda")
    b = DSBatch.from_tensors(s, ro, ni, ns)
    _case("ds_batch per-stage reference, N in " + ",".join(map(str, ns_)))
    c = [x.clone() for x in DSBatchRunner(b, per_stage=True).run()]
    r = DSBatchRunner(b)
    r.tune()
    assert not any(g.staged for g in r.groups)
    for blocks in (0, 3):
        _case(f"ds_batch fused N<=32 and 64<N<=200, blocks={blocks}")
        r.ph.zero_()
        r.quality.zero_()
        r.t_coh.zero_()
        for g in r.groups:
            g.blocks = blocks
        a = r.run()
        torch.cuda.synchronize()
        assert torch.equal(a[0].view(torch.int64), c[0].view(torch.int64)), blocks
        assert torch.equal(a[2].view(torch.int32), c[2].view(torch.int32)), blocks
        assert torch.equal(a[3].view(torch.int32), c[3].view(torch.int32)), blocks


def _coh64(z):
    z = z.astype(np.complex128)
    s = z.T @ z.conj()
    d = np.sqrt(np.real(np.diag(s)))
    iu = np.triu_indices(z.shape[1], 1)
    return (s / np.outer(d, d))[iu]


def test_batched_vs_fp64():
    """Synthetic ragged pixels (N in {30, 60, 90, 150}, K in [20, 315]) against
    FP64 coherence + FP64 EMI + FP64 temporal coherence.  Linked-phase RMS
    <= 1e-3 rad where K >= N (well-posed coherence); rank-deficient pixels are
    reported, not asserted (their FP64 answer moves by more than that under
    FP32 rounding of the coherence itself)."""
    from moraine_amd.batched import ds_batch

    mats = synth.ds_samples(48, seed=7)
    _case("ds_batch synthetic vs FP64, 48 pixels")
    ph, q, t = ds_batch(mats)
    rms_ok, rms_all = [], []
    for m, p, qq, tt in zip(mats, ph, q, t):
        c = _coh64(m)[None]
        p64, q64 = R.emi64(c)
        e = np.angle(p * np.conj(p64[0]))
        rms = float(np.sqrt(np.mean(e ** 2)))
        rms_all.append(rms)
        if m.shape[0] >= m.shape[1]:
            rms_ok.append(rms)
            assert abs(qq - q64[0]) <= 1e-3 * max(1.0, abs(q64[0]))
            t64 = R.ds_temp_coh64(c, p64)[0]
            assert abs(tt - t64) <= 1e-4
    print(f"phase RMS: K>=N max {max(rms_ok):.2e} (n={len(rms_ok)}), all median {np.median(rms_all):.2e}")
    assert len(rms_ok) >= 10 and max(rms_ok) <= 1e-3


def test_batched_edges():
    import torch

    from moraine_amd.batched import DSBatch, DSBatchRunner, ds_batch

    _case("ds_batch empty and single pixel")
    ph, q, t = ds_batch([])
    assert len(ph) == 0 and q.shape == (0,)
    one = synth.ds_samples(1, n_choices=(60,), seed=9)
    ph1, q1, t1 = ds_batch(one)
    assert ph1[0].shape == (60,) and np.allclose(np.abs(ph1[0]), 1, atol=1e-5)
    with pytest.raises(ValueError):
        DSBatchRunner(DSBatch.from_matrices([np.zeros((5, 1), np.complex64)]))


def test_fused_resolve_equals_per_stage():
    """N = 60 pixels with fewer SHPs than images (K in [20, 59]: |Gamma| near
    singular): the fused launch re-solves the pixels its Householder pass flags
    from tiles rebuilt on chip, the per-stage path from the stored coherence
    rows; the outputs are identical, and some pixels take that path."""
    import torch

    from moraine_amd.batched import DSBatch, DSBatchRunner

    s, ro, ni, ns = synth.ds_samples_torch(6000, n_choices=(60,), k_range=(20, 59), seed=11, device="cuda")
    b = DSBatch.from_tensors(s, ro, ni, ns)
    _case("ds_batch fused re-solve vs per-stage, N=60, 6000 pixels, K in [20, 59]")
    r = DSBatchRunner(b)
    a = [x.clone() for x in r.run()]
    n_hh, n_res = (int(v) for v in r.ws[:8].view(torch.int32).cpu())
    print(f"  Householder pass {n_hh}, FP64 re-solve {n_res} of 6000")
    assert 0 < n_res <= n_hh
    c = DSBatchRunner(b, per_stage=True).run()
    torch.cuda.synchronize()
    assert torch.equal(a[0].view(torch.int64), c[0].view(torch.int64))
    assert torch.equal(a[2].view(torch.int32), c[2].view(torch.int32))
    assert torch.equal(a[3].view(torch.int32), c[3].view(torch.int32))


def test_batched_coherence_tile_layouts_equal():
    """Every tile layout of the per-stage coherence kernel (waves per workgroup
    over its 32 x 32 tile grid) gives the fused launch's outputs bit for bit;
    a layout the grid does not have is rejected."""
    import torch

    from moraine_amd.batched import AutoTuner, DSBatch, DSBatchRunner

    s, ro, ni, ns = synth.ds_samples_torch(900, n_choices=(70, 150, 200), k_range=(20, 240), seed=8, device="cuda")
    b = DSBatch.from_tensors(s, ro, ni, ns)
    ref = [x.clone() for x in DSBatchRunner(b).run()]  # the fused launches
    r = DSBatchRunner(b, per_stage=True)
    r.tune()
    for g in r.groups:
        for v in AutoTuner.coh_waves(g.n):
            _case(f"ds_batch coherence layout N={g.n} waves={v}")
            r.ph.zero_()
            r.quality.zero_()
            r.t_coh.zero_()
            r._call(g, -1, g.chunk, v)
            torch.cuda.synchronize()
            sel = g.pix.long()
            assert torch.equal(r.quality[sel].view(torch.int32), ref[2][sel].view(torch.int32))
            assert torch.equal(r.t_coh[sel].view(torch.int32), ref[3][sel].view(torch.int32))
            offs = ref[1][sel]
            for o in offs[:64].tolist():
                assert torch.equal(r.ph[o:o + g.n].view(torch.int64), ref[0][o:o + g.n].view(torch.int64))
    assert r.groups[1].n == 150
    with pytest.raises(NotImplementedError):
        r._call(r.groups[1], -1, r.groups[1].chunk, 2)  # no two-wave layout for a 5 x 5 tile grid
