"""CPU checks of the batched form's host logic: ragged layout, the sample
matrices built from a raster window (row order pinned against the oracle's
coherence), the autotuner and the synthetic generators.  No GPU."""
import json

import numpy as np

from moraine_amd import synth
from moraine_amd.batched import AutoTuner, DSBatch, raster_sample_matrices
from oracle import moraine_ref as R


def test_layout_offsets():
    row_off, ph_off = DSBatch.layout([30, 60, 30], [20, 5, 7])
    np.testing.assert_array_equal(row_off, [0, 600, 900])
    np.testing.assert_array_equal(ph_off, [0, 30, 90])
    r0, p0 = DSBatch.layout([], [])
    assert r0.size == 0 and p0.size == 0


def _coh64_from_samples(z):
    """FP64 coherence of the triu pairs from a (K, N) sample matrix."""
    z = z.astype(np.complex128)
    s = z.T @ z.conj()  # S_ab = sum_k z_ka conj(z_kb)
    d = np.sqrt(np.real(np.diag(s)))
    iu = np.triu_indices(z.shape[1], 1)
    return (s / np.outer(d, d))[iu]


def test_raster_samples_reproduce_point_cloud_coherence():
    """The window-order SHP rows of a pixel give the oracle's emperical_co_pc
    (moraine/co.py:209-226) coherence, including windows clipped at the border."""
    z = synth.ds_stack(9, 11, 7, seed=5)
    rng = np.random.default_rng(0)
    idx = np.array([[0, 0], [4, 5], [8, 10], [2, 9]], np.int32)
    shp = rng.random((4, 5, 7)) < 0.6
    shp[:, 2, 3] = True
    mats = raster_sample_matrices(z, idx, shp)
    ref = R.emperical_co_pc(z, idx, shp)
    for m, c in zip(mats, ref):
        np.testing.assert_allclose(_coh64_from_samples(m), c, rtol=2e-6, atol=2e-7)
    # clipped window at the corner holds fewer rows than the mask count
    assert mats[0].shape[0] < shp[0].sum()


def test_autotuner_picks_fastest_and_persists(tmp_path):
    cost = {0: 3.0, 4096: 1.0, 16384: 2.0}
    calls = []

    def run(c):
        calls.append(c)

    t = AutoTuner(path=str(tmp_path / "tune.json"), reps=2,
                  timer=lambda fn: (fn(), cost[calls[-1]])[1])
    key = AutoTuner.key("fused", 60, 1000)
    assert key == "fused:n=60:px<=1024"
    assert t.choose(key, (0, 4096, 16384), run) == 4096
    assert t.trials[key] == cost
    assert json.load(open(tmp_path / "tune.json")) == {key: 4096}
    # a new tuner starts from the persisted pick without timing anything
    n_calls = len(calls)
    t2 = AutoTuner(path=str(tmp_path / "tune.json"), timer=lambda fn: 1 / 0)
    assert t2.choose(key, (0, 4096, 16384), run) == 4096
    assert len(calls) == n_calls


def test_sample_generators_shapes():
    mats = synth.ds_samples(20, seed=2)
    for m in mats:
        assert m.shape[1] in (30, 60, 90, 150) and 20 <= m.shape[0] <= 315 and m.dtype == np.complex64
    s, ro, ni, ns = synth.ds_samples_torch(50, seed=1, device="cpu", chunk=16)
    ns = ns.numpy()
    assert s.numel() == int((ni.astype(np.int64) * ns).sum())
    np.testing.assert_array_equal(ro.numpy(), DSBatch.layout(ni, ns)[0])
    assert np.isfinite(s.numpy()).all()


def test_fused_workspace_sizes():
    """Workspace of the fused launches (a host-side query, no GPU needed): the
    pixel lists for 32 < N <= 64, one coherence row per workgroup for
    64 < N <= 200 (2048 workgroups by default, never more than the pixels),
    nothing for N <= 32 and no fused launch above N = 200."""
    from moraine_amd import _lib

    def q(n_pix, n, blocks):
        return _lib.query("mr_ds_batch_fused_workspace_bytes", n_pix, n, blocks)

    assert q(1000, 17, 0) == 0
    assert q(1000, 60, 0) == 256 + 2 * ((1000 * 4 + 255) // 256 * 256)
    assert q(5000, 150, 0) == 2048 * (150 * 149 // 2) * 8
    assert q(100, 150, 0) == 100 * (150 * 149 // 2) * 8
    assert q(5000, 200, 7) == 7 * (200 * 199 // 2) * 8
    assert q(5000, 201, 0) == 0
    assert q(0, 150, 0) == 0
    assert AutoTuner.fused_blocks(150) == AutoTuner.WIDE_BLOCKS
    assert AutoTuner.fused_blocks(30) == AutoTuner.FUSED_BLOCKS
