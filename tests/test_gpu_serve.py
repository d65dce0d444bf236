"""GPU tests of the serving form (moraine_amd/serve.py).

A request's answer must not depend on the other requests sharing its ticks:
every request served in a mix equals the same request served alone, bit for
bit, and both equal the drop-in composition ks_test_select_shp ->
emperical_co_pc -> emi -> ds_temp_coh.
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _case(msg):
    print(f"[gpu case] {msg}", flush=True)
    sys.stdout.flush()


def _specs():
    from moraine_amd.serve import PatchSpec

    # (n, h, w, az, r): differing N and windows; tiles small enough for the test
    shapes = [(17, 40, 48, 3, 4), (30, 32, 32, 5, 7), (60, 48, 40, 2, 3), (90, 24, 32, 3, 5), (17, 16, 16, 5, 7),
              (120, 20, 24, 2, 4)]
    return [PatchSpec(i, n, h, w, az, r, seed=50 + i) for i, (n, h, w, az, r) in enumerate(shapes)]


def _same(a, b):
    import torch

    for x, y in ((a.ph, b.ph), (a.quality, b.quality), (a.t_coh, b.t_coh), (a.shp_num, b.shp_num)):
        x, y = x.contiguous(), y.contiguous()
        if x.is_complex():
            x, y = torch.view_as_real(x), torch.view_as_real(y)
        assert torch.equal(x.view(torch.int32) if x.dtype == torch.float32 else x,
                           y.view(torch.int32) if y.dtype == torch.float32 else y)


def test_alone_equals_mix_and_drop_in():
    import torch

    import moraine_amd as mr
    from moraine_amd import synth
    from moraine_amd.serve import DSServer

    _case("serve: mix of 6 requests, 700-pixel ticks")
    srv = DSServer(max_tick_px=700, coh_budget=64 << 20, max_inflight=2)
    for s in _specs():
        srv.submit(s)
    srv.drain()
    torch.cuda.synchronize()
    mix = {r.rid: r for r in srv.results}
    assert sorted(mix) == list(range(6))
    assert srv.max_concurrent >= 2  # requests really shared the device
    for s in _specs():
        _case(f"serve: request {s.rid} alone (N={s.n}, {s.h}x{s.w}, window {2*s.az+1}x{2*s.r+1})")
        one = DSServer(coh_budget=64 << 20)
        one.submit(s)
        one.drain()
        torch.cuda.synchronize()
        _same(mix[s.rid], one.results[0])
    # the drop-in composition on the same stack
    s = _specs()[2]
    _case("serve vs drop-in composition, request 2")
    z = synth.ds_stack_torch(s.h, s.w, s.n, seed=s.seed, ps_frac=0.1, device="cuda")
    x = torch.view_as_real(z)
    rmli = (x[..., 0] * x[..., 0] + x[..., 1] * x[..., 1]).contiguous()
    _, m, cnt = mr.ks_test_select_shp(rmli, s.az, s.r, 0.05, return_p=False)
    ii, jj = np.meshgrid(np.arange(s.h), np.arange(s.w), indexing="ij")
    idx = torch.from_numpy(np.stack([ii.ravel(), jj.ravel()], -1).astype(np.int32)).cuda()
    c = mr.emperical_co_pc(z, idx, m.view(s.npx, 2 * s.az + 1, 2 * s.r + 1))
    ph, q = mr.emi(c)
    t = mr.ds_temp_coh(c, ph)
    r = mix[2]
    assert torch.equal(torch.view_as_real(r.ph.reshape(-1, s.n)).view(torch.int32),
                       torch.view_as_real(ph).view(torch.int32))
    assert torch.equal(r.quality.reshape(-1).view(torch.int32), q.view(torch.int32))
    assert torch.equal(r.t_coh.reshape(-1).view(torch.int32), t.view(torch.int32))
    assert torch.equal(r.shp_num, cnt)


def test_admission_is_bounded_by_hbm_budget():
    import torch

    from moraine_amd.serve import DSServer, footprint_bytes

    specs = _specs()
    budget = max(footprint_bytes(s) for s in specs) + 1
    _case(f"serve: HBM budget {budget} bytes (one request at a time)")
    srv = DSServer(hbm_budget=budget, coh_budget=64 << 20, max_tick_px=4096)
    for s in specs:
        srv.submit(s)
    srv.drain()
    torch.cuda.synchronize()
    assert len(srv.results) == len(specs)
    assert srv.peak_used <= budget and srv.used == 0
    with pytest.raises(ValueError):
        srv.submit(type(specs[0])(99, 60, 512, 512, 7, 10))  # larger than the budget


def test_poisson_stream_and_serial_baseline():
    import torch

    from moraine_amd.serve import DSServer, run_serial, run_stream, stream_specs, stream_stats

    specs = stream_specs(8, rate=50.0, seed=3, n_choices=(17, 30, 60), edge_range=(32, 96), az_range=(2, 4),
                         r_range=(3, 6))
    _case("serve: 8-request Poisson stream")
    res = run_stream(specs, DSServer(coh_budget=256 << 20, keep_outputs=True))
    st = stream_stats(res, specs)
    assert st["requests"] == 8 and st["px_per_s"] > 0 and st["p99_latency_s"] >= st["p50_latency_s"] > 0
    _case("serve: serial plain-form baseline of the same stream")
    ser = run_serial(specs, keep_outputs=True)
    for a, b in zip(res, ser):
        assert torch.equal(torch.view_as_real(a.ph).view(torch.int32), torch.view_as_real(b.ph).view(torch.int32))
    print(stream_stats(ser, specs))


def test_config5_shaped_mix_against_oracle(port):
    """A served mix at the serving workload's shapes: an N=200 request, 21x41 and
    11x15 windows, tiles of 128^2 and larger.  SHP counts equal the oracle's KS
    (C restatement of the reference's numba path) bit for bit; sampled points'
    phases are within 1e-3 rad RMS of FP64 EMI on the same coherence rows."""
    import torch

    import moraine_amd as mr
    from moraine_amd import synth
    from moraine_amd.serve import DSServer, PatchSpec
    from oracle import moraine_ref as R

    shapes = [(200, 128, 128, 10, 20), (30, 160, 128, 5, 7), (60, 128, 144, 7, 10), (120, 136, 128, 10, 20)]
    specs, stacks = [], {}
    for i, (n, h, w, az, r) in enumerate(shapes):
        z = synth.ds_stack(h, w, n, seed=900 + i, ps_frac=0.1)
        stacks[i] = z
        specs.append(PatchSpec(i, n, h, w, az, r, seed=900 + i, rslc=z))
    _case("serve: config-5-shaped mix (N=200 128x128 21x41, N=30 11x15, N=60 15x21, N=120 21x41)")
    srv = DSServer(coh_budget=1 << 30, max_tick_px=1 << 15)
    for s in specs:
        srv.submit(s)
    srv.drain()
    torch.cuda.synchronize()
    res = {x.rid: x for x in srv.results}
    assert sorted(res) == list(range(len(specs)))
    rng = np.random.default_rng(5)
    for s in specs:
        z = stacks[s.rid]
        rmli = z.real.astype(np.float32) ** 2 + z.imag.astype(np.float32) ** 2
        _case(f"serve: request {s.rid} N={s.n} {s.h}x{s.w} window {2*s.az+1}x{2*s.r+1} against the oracle")
        p_ref = port.ks_test(rmli, s.az, s.r)
        m_ref, n_ref = R.select_shp(p_ref, 0.05)
        np.testing.assert_array_equal(res[s.rid].shp_num.cpu().numpy().reshape(s.h, s.w), n_ref)
        pts = rng.choice(s.npx, 48, replace=False)
        idx = np.stack([pts // s.w, pts % s.w], -1).astype(np.int32)
        c = mr.emperical_co_pc(z, idx, m_ref[idx[:, 0], idx[:, 1]])
        ph64, _ = R.emi64(c)
        ph = res[s.rid].ph.reshape(s.npx, s.n).cpu().numpy()[pts]
        rms = np.sqrt(np.nanmean(np.angle(ph * np.conj(ph64)) ** 2, axis=1))
        print(f"  phase RMS vs FP64 on 48 points: max {np.nanmax(rms):.2e}")
        assert np.nanmax(rms) < 1e-3
