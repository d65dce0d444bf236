#!/usr/bin/env python
"""DS hot-path throughput on one MI355X: KS + SHP selection + coherence + EMI + temporal coherence.

Workload (default): N=60 complex64 SLCs over a 2500x2500 tile, DS model with 10%
PS-like pixels, 15x21 search window (az_half=7, r_half=10), p_max = 0.05, every
pixel taken through the whole path.  One step = one pass over the tile:

    rmli = |rslc|^2
    p, is_shp, shp_num = ks_test + select_shp              (p materialised, as the drop-in returns it)
    per batch of pixels: (ph, quality, t) = emperical_co_pc -> emi -> ds_temp_coh, the batch
                         driver's step (for 32 < N <= 64 one launch computes each point's
                         coherence into LDS and runs EMI and the temporal coherence from there;
                         batches alternate over two streams)

Inputs are generated on the device before timing (synthetic, seeded).  Prints
one JSON line.  Algorithmic units per pixel: KS 4N + 5W + 4 bytes, coherence 4K N(N+1)
flops, EMI N^3 (FP64) + (4/3) N^3 (FP32) + 8 N^2 x 2 iterations flops (the formulation-
independent count: Cholesky and inverse of |Gamma|, a complex LDL^H factorisation and two
inverse-iteration sweeps), temporal coherence 8P + 8N + 4 bytes (W window entries, K SHPs,
P = N(N-1)/2 pairs).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6  # MI355X FP64 (vector and matrix); also FP32 vector without packed math


def emi_algorithmic_flops(n: int) -> float:
    """Algorithmic EMI flops per point: N^3 (FP64 Cholesky + SPD inverse of |Gamma|)
    + (4/3) N^3 (FP32 complex LDL^H) + 8 N^2 per inverse-iteration sweep, two sweeps."""
    return n ** 3 + (4.0 / 3.0) * n ** 3 + 2 * 8.0 * n ** 2




def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--size", type=int, default=2500, help="tile is size x size pixels")
    ap.add_argument("--n", type=int, default=60)
    ap.add_argument("--az", type=int, default=7)
    ap.add_argument("--r", type=int, default=10)
    ap.add_argument("--ps", type=float, default=0.10)
    ap.add_argument("--batch", type=int, default=1 << 21, help="pixels per coherence/EMI batch")
    ap.add_argument("--batch-streams", type=int, default=2, choices=(1, 2),
                    help="streams the plain pass's batches alternate over (1: one after the other, for "
                         "per-kernel profiles without overlap)")
    ap.add_argument("--cpu-sample", type=int, default=256, help="CPU baseline: interior sub-tile edge")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--err-sample", type=int, default=2048)
    ap.add_argument("--forms", default="plain,batched,serving,rows,streamed,precision",
                    help="comma list: plain (drop-in pipeline, the headline), batched, serving, rows (each "
                         "drop-in function timed alone), streamed (the same tile from host memory in bands), "
                         "precision (the N=120, 21x41, 1000x4000 mixed-coherence workload)")
    ap.add_argument("--batched-px", type=int, default=200000, help="pixels of the ragged batch")
    ap.add_argument("--serve-requests", type=int, default=16)
    ap.add_argument("--serve-edge", type=int, default=1024, help="largest tile edge of the request stream")
    ap.add_argument("--c3-lines", type=int, default=1000, help="precision workload: tile lines (x 4000 columns)")
    return ap.parse_args()


def progress(msg):
    """One line per bench stage on stderr (the JSON result is the only stdout line)."""
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def measured_traffic(kernels, args):
    """HBM bytes per launch of the given kernels (summed) from the committed FETCH_SIZE/WRITE_SIZE
    counter summary (copy-kernel calibrated), valid only for the workload it was measured on."""
    if (args.size, args.n, args.az, args.r, args.batch) != (2500, 60, 7, 10, 1 << 21):
        return None
    try:
        with open(os.path.join(ROOT, "profiles", "traffic.json")) as f:
            per = json.load(f)["per_dispatch"]
        return sum(per[k]["read_bytes"] + per[k]["write_bytes"] for k in kernels)
    except (OSError, KeyError, ValueError):
        return None


def ensure_built():
    so = os.path.join(ROOT, "moraine_amd", "libmoraine_hip.so")
    if not os.path.exists(so):
        from moraine_amd import _build

        _build.build()


# ---------------------------------------------------------------------------- CPU baseline
def _cpu_pipeline(z, az, r, rows, cols, cores):
    """The reference CPU pipeline on the pixels rows x cols of stack z: KS with the p
    series per pixel pair and select_shp (C/OpenMP port), emperical_co_pc over the
    full pair network (port), EMI as moraine/pl.py:45-65 computes it (float32 inv,
    complex64 eigh, LAPACK, one thread per chunk), ds_temp_coh (port).  Returns the
    seconds taken and the SHP masks of those pixels."""
    import concurrent.futures as cf

    import numpy as np

    from oracle import moraine_ref as R
    from oracle import port

    n = z.shape[-1]
    rmli = (np.abs(z) ** 2).astype(np.float32)
    t0 = time.perf_counter()
    p = port.ks_test(rmli, az, r, nthreads=cores, per_pair=True)
    is_shp, _ = R.select_shp(p, 0.05)
    ii, jj = np.meshgrid(rows, cols, indexing="ij")
    idx = np.stack([ii.ravel(), jj.ravel()], -1).astype(np.int32)
    pc_shp = is_shp[ii.ravel(), jj.ravel()]
    pairs = R.image_pairs_from_bandwidth(n)
    coh = port.emperical_co_pc(z, idx, pc_shp, pairs, nthreads=cores)
    chunks = np.array_split(np.arange(idx.shape[0]), cores * 4)
    with cf.ThreadPoolExecutor(cores) as ex:
        res = list(ex.map(lambda c: R.emi_ref32(coh[c]), chunks))
    ph = np.concatenate([a for a, _ in res])
    port.ds_temp_coh(coh, ph, pairs, nthreads=cores)
    return time.perf_counter() - t0, pc_shp


def _gpu_pipeline_config1(z, az, r, torch):
    """The drop-in pipeline on the GPU for a whole small stack (KS + select_shp on the
    CPU path's amplitudes |z|^2, emperical_co_pc, EMI + ds_temp_coh): best of five
    timed runs (amplitudes uploaded once, outside them) and the SHP masks."""
    import numpy as np

    import moraine_amd as mr

    zt = torch.from_numpy(z).cuda()
    rmli = torch.from_numpy((np.abs(z) ** 2).astype(np.float32)).cuda()
    nl, w, n = z.shape
    ii, jj = torch.meshgrid(torch.arange(nl, device="cuda", dtype=torch.int32),
                            torch.arange(w, device="cuda", dtype=torch.int32), indexing="ij")
    idx = torch.stack([ii.reshape(-1), jj.reshape(-1)], -1).contiguous()

    def run():
        _, shp, _ = mr.ks_test_select_shp(rmli, az, r, 0.05, return_p=False)
        coh = mr.emperical_co_pc(zt, idx, shp.view(-1, 2 * az + 1, 2 * r + 1))
        mr.pl.emi_ds_temp_coh(coh)
        return shp

    run()
    torch.cuda.synchronize()
    best = float("inf")
    for _ in range(5):
        t0 = time.perf_counter()
        shp = run()
        torch.cuda.synchronize()
        best = min(best, time.perf_counter() - t0)
    return best, shp.view(nl * w, 2 * az + 1, 2 * r + 1).cpu().numpy()


def cpu_baseline(args, torch):
    """The reference CPU path restated (oracle/port.c, C/OpenMP, the p series per
    pixel pair as the reference evaluates it; LAPACK EMI as moraine/pl.py:45-65) on
    all host cores: on a bounded interior sub-tile of the headline workload (the
    value), and on the whole small CPU workload (200x200, 17 images, 11x15 windows)
    with the GPU drop-in timed on the same stack beside it."""
    import numpy as np

    from moraine_amd import synth
    from oracle import port

    port.build()
    cores = len(os.sched_getaffinity(0))
    s, az, r, n = args.cpu_sample, args.az, args.r, args.n
    z = synth.ds_stack(s + 2 * az, s + 2 * r, n, seed=123, ps_frac=args.ps)
    dt, _ = _cpu_pipeline(z, az, r, np.arange(az, az + s), np.arange(r, r + s), cores)
    npx = s * s
    # the whole small workload: 200 x 200 pixels, 17 images, 11 x 15 windows, no PS pixels
    z1 = synth.ds_stack(200, 200, 17, seed=1)
    dt1, shp_cpu = _cpu_pipeline(z1, 5, 7, np.arange(200), np.arange(200), cores)
    gt1, shp_gpu = _gpu_pipeline_config1(z1, 5, 7, torch)
    return {
        "value": npx / dt, "unit": "DS pixels/s", "cores": cores, "kind": "port",
        "sample": f"{s}x{s} interior pixels (+{az}/{r} halo) of the N={n}, {2*az+1}x{2*r+1}, {int(args.ps*100)}% PS "
                  f"workload: C/OpenMP KS (p series per pixel pair) + select + coherence + tcoh, numpy LAPACK "
                  f"EMI (f32 inv, c64 eigh); {dt:.1f} s",
        "small_workload": {
            "workload": "200x200 pixels, 17 images, 11x15 windows, p_max=0.05, no PS pixels, every pixel",
            "cpu_s": dt1, "cpu_px_per_s": 40000 / dt1, "gpu_s": gt1, "gpu_px_per_s": 40000 / gt1,
            "gpu_over_cpu": dt1 / gt1, "shp_masks_equal": bool(np.array_equal(shp_cpu, shp_gpu)),
            "note": "same pipeline and stack; GPU: best of 5 after a warm-up, launch-bound at this size",
        },
    }


# ---------------------------------------------------------------------------- other forms
def batched_leg(args, torch):
    """Ragged batch: N in {30, 60, 90, 150}, K in [20, 315], one sample matrix per
    pixel.  Fused + autotuned + HIP-graph replay against the per-stage launches
    (coherence through HBM, EMI, temporal coherence per image-count group)."""
    from moraine_amd import synth
    from moraine_amd.batched import AutoTuner, DSBatch, DSBatchRunner

    s, ro, ni, ns = synth.ds_samples_torch(args.batched_px, seed=0, device="cuda")
    b = DSBatch.from_tensors(s, ro, ni, ns)
    torch.cuda.synchronize()

    def best(fn, reps=3):
        fn()
        torch.cuda.synchronize()
        ts = []
        for _ in range(reps):
            t0 = time.perf_counter()
            fn()
            torch.cuda.synchronize()
            ts.append(time.perf_counter() - t0)
        return min(ts)

    tuner = AutoTuner()
    fused = DSBatchRunner(b, tuner=tuner)
    fused.capture()
    t_graph = best(fused.replay)
    per_stage = DSBatchRunner(b, per_stage=True)
    t_stage = best(per_stage.run)
    groups = {}
    for g in fused.groups:
        groups[str(g.n)] = {"pixels": g.count, "staged": g.staged,
                            "s": best(lambda g=g: fused._call(g, -1 if g.staged else g.blocks, g.chunk), 2)}
    # phase error vs FP64 (complex128 coherence, inverse and eigh) on a sample per group
    ph, ph_off, q, _ = fused.replay()
    torch.cuda.synchronize()
    off = ph_off.cpu().numpy()
    rms_all = []
    import numpy as np

    for g in fused.groups:
        ids = g.pix[:64].cpu().numpy()
        for i in ids:
            k, n = int(ns[i]), int(ni[i])
            Z = s[int(ro[i]):int(ro[i]) + k * n].view(k, n).to(torch.complex128)
            S = Z.T @ Z.conj()
            d = torch.sqrt(torch.diagonal(S).real)
            G = S / torch.outer(d, d)
            M = torch.linalg.inv(G.abs()).to(torch.complex128) * G
            _, V = torch.linalg.eigh(M)
            v = V[:, 0] * V[0, 0].conj()
            e = torch.angle(ph[off[i]:off[i] + n].to(torch.complex128) * (v / v.abs()).conj())
            rms_all.append(float(torch.sqrt(torch.mean(e ** 2))))
    rms_all = np.array(rms_all)
    progress("batched form: N=60 window batch")
    n60 = _n60_window_batch(args, torch)
    return {
        "n60_15x21_window_batch": n60,
        "workload": f"{args.batched_px} pixels, N in {{30,60,90,150}}, K in [20,315], gamma0~U[0.3,0.9], "
                    f"tau~U[12,96] d",
        "value": args.batched_px / t_graph, "unit": "DS pixels/s", "s_per_batch": t_graph,
        "per_stage_value": args.batched_px / t_stage, "fused_vs_per_stage": t_stage / t_graph,
        "tuner_picks": tuner.picks,
        "tuner_trials_ms": {k: {str(c): round(v * 1e3, 3) for c, v in t.items()} for k, t in tuner.trials.items()},
        "groups": groups,
        "phase_rms_vs_fp64": {"median": float(np.median(rms_all)), "p99": float(np.quantile(rms_all, 0.99)),
                              "max": float(rms_all.max()), "pixels": int(rms_all.size)},
    }


def _n60_window_batch(args, torch):
    """The batched form on the plain workload's own pixels: N = 60, SHP sets from the KS test over
    15x21 windows of a 384x384 DS stack; each pixel's sample matrix is its in-bounds SHP stacks in
    window order, gathered on the device.  Fused kernel (autotuned, HIP-graph replay) against the
    plain drop-in on the same pixels (emperical_co_emi_temp_coh_pc on the raster: coherence, EMI and
    DS temporal coherence): both compute the same outputs (bit for bit, tests/test_gpu_batched.py)."""
    import numpy as np

    import moraine_amd as mr
    from moraine_amd import synth
    from moraine_amd.batched import AutoTuner, DSBatch, DSBatchRunner

    S, N, az, r = 384, 60, 7, 10
    aw, rw = 2 * az + 1, 2 * r + 1
    z = synth.ds_stack_torch(S, S, N, seed=7, ps_frac=args.ps, device="cuda")
    x = torch.view_as_real(z)
    rmli = (x[..., 0] * x[..., 0] + x[..., 1] * x[..., 1]).contiguous()
    _, shp, cnt = mr.ks_test_select_shp(rmli, az, r, 0.05, return_p=False)
    del rmli, x
    npx = S * S
    ii, jj = torch.meshgrid(torch.arange(S, device="cuda"), torch.arange(S, device="cuda"), indexing="ij")
    ka, kr = torch.meshgrid(torch.arange(aw, device="cuda") - az, torch.arange(rw, device="cuda") - r, indexing="ij")
    sa = ii.reshape(-1, 1) + ka.reshape(1, -1)
    sr = jj.reshape(-1, 1) + kr.reshape(1, -1)
    valid = shp.reshape(npx, -1) & (sa >= 0) & (sa < S) & (sr >= 0) & (sr < S)
    src = (sa * S + sr)[valid]  # window order within each pixel, pixels in index order
    k = valid.sum(1).to(torch.int64)
    samples = z.reshape(-1, N)[src].reshape(-1).contiguous()
    row_off = torch.zeros(npx, dtype=torch.int64, device="cuda")
    row_off[1:] = torch.cumsum(k * N, 0)[:-1]
    b = DSBatch.from_tensors(samples, row_off, np.full(npx, N, np.int32), k.to(torch.int32))
    del sa, sr, valid, src
    idx = torch.stack([ii.reshape(-1), jj.reshape(-1)], -1).to(torch.int32).contiguous()
    pcs = shp.reshape(npx, aw, rw)

    def best(fn, reps=3):
        fn()
        torch.cuda.synchronize()
        ts = []
        for _ in range(reps):
            t0 = time.perf_counter()
            fn()
            torch.cuda.synchronize()
            ts.append(time.perf_counter() - t0)
        return min(ts)

    def plain():
        mr.emperical_co_emi_temp_coh_pc(z, idx, pcs)

    runner = DSBatchRunner(b, tuner=AutoTuner())
    runner.capture()
    t_b = best(runner.replay)
    t_p = best(plain)
    return {"workload": f"{S}x{S} pixels, N={N}, 15x21 windows, SHP sets from the KS test (mean K "
                        f"{float(k.float().mean()):.1f})", "pixels": npx,
            "batched_px_per_s": npx / t_b, "plain_px_per_s": npx / t_p, "batched_vs_plain": t_p / t_b,
            "note": "coherence + EMI + temporal coherence only (KS excluded from both)"}


def streamed_leg(args, torch):
    """The plain workload with the stack in host memory: ds_pipeline_streamed
    copies row bands (with halo) through pinned buffers on a copy stream while
    the previous band computes, and returns host results.  PCIe- and
    host-copy-inclusive; reported beside the headline, never as it."""
    from moraine_amd import synth
    from moraine_amd.stream import ds_pipeline_streamed

    S, N = args.size, args.n
    z = synth.ds_stack_torch(S, S, N, seed=0, ps_frac=args.ps, device="cuda").cpu().numpy()
    ds_pipeline_streamed(z[: min(S, 64)], args.az, args.r, 0.05, band_rows=32)  # warm-up
    t0 = time.perf_counter()
    ph, q, t, cnt = ds_pipeline_streamed(z, args.az, args.r, 0.05, band_rows=256)
    dt = time.perf_counter() - t0
    del ph, q, t, cnt, z
    return {"workload": f"{S}x{S}x{N} complex64 stack in host memory, 15x21 windows, 256-row bands + halo, "
                        "results to host (ph, quality, t_coh, shp_num)",
            "value": S * S / dt, "unit": "DS pixels/s", "s": dt,
            "note": "host -> pinned -> HBM copies on a copy stream overlap the previous band's compute; "
                    "includes host copies and PCIe both ways"}


def rows_leg(args, torch):
    """Device time of each drop-in function on its own (synthetic N=60 inputs, 15x21 windows), with its
    algorithmic bytes or flops per call and the fraction of the MI355X peak that bounds it (HBM 8 TB/s;
    FP64 / non-packed FP32 vector 78.6 TFLOP/s for EMI; FP16 dense MFMA 2.5 PFLOP/s for coherence)."""
    import moraine_amd as mr
    from moraine_amd import synth
    from moraine_amd.pl import emi_ds_temp_coh

    def dtime(fn, reps=3):
        fn()
        torch.cuda.synchronize()
        best = float("inf")
        for _ in range(reps):
            a = torch.cuda.Event(enable_timing=True)
            b = torch.cuda.Event(enable_timing=True)
            a.record()
            fn()
            b.record()
            b.synchronize()
            best = min(best, a.elapsed_time(b) / 1e3)
        return best

    def hbm(name, t, nbytes, what):
        rows[name] = {"s": t, "GB/s": nbytes / t / 1e9, "frac": nbytes / t / (HBM_PEAK_GBS * 1e9), "bound": "hbm",
                      "per_call": what}

    def flops(name, t, nflop, peak, what):
        rows[name] = {"s": t, "TFLOP/s": nflop / t / 1e12, "frac": nflop / t / (peak * 1e12),
                      "bound": "mfma" if peak > 1000 else "valu", "per_call": what}

    rows = {}
    N, az, r = 60, 7, 10
    aw, rw = 2 * az + 1, 2 * r + 1
    W, P = aw * rw, N * (N - 1) // 2
    S = 1024
    z = synth.ds_stack_torch(S, S, N, seed=5, ps_frac=0.1, device="cuda")
    x = torch.view_as_real(z)
    rmli = (x[..., 0] * x[..., 0] + x[..., 1] * x[..., 1]).contiguous()
    del x
    npx = S * S
    hbm("ks_test", dtime(lambda: mr.ks_test(rmli, az, r)), npx * (4 * N + 4 * W),
        f"{S}x{S}x{N}, rmli in, p out")
    p = mr.ks_test(rmli, az, r)
    hbm("select_shp", dtime(lambda: mr.select_shp(p, 0.05)), npx * (4 * W + W + 4), "p in, mask + count out")
    del p
    hbm("ks_test_select_shp", dtime(lambda: mr.ks_test_select_shp(rmli, az, r, 0.05, return_p=False)),
        npx * (4 * N + W + 4), "rmli in, mask + count out")
    hbm("rslc2amp", dtime(lambda: mr.rslc2amp(z)), npx * N * (8 + 4), f"{S}x{S}x{N} complex64 in, float32 out")
    hbm("amp_disp", dtime(lambda: mr.amp_disp(z)), npx * (8 * N + 4), f"{S}x{S}x{N} complex64 in, float32 out")
    hbm("ks_test_select_shp_rslc", dtime(lambda: mr.ks_test_select_shp(z, az, r, 0.05, return_p=False)),
        npx * (8 * N + W + 4), "complex64 stack in (amplitude while staging), mask + count out")
    _, shp, cnt = mr.ks_test_select_shp(rmli, az, r, 0.05, return_p=False)
    kmean = float(cnt.float().mean())
    co_fl = 4.0 * kmean * N * (N + 1)
    # raster coherence on a 256x256 corner (cov and coh: 2 N^2 complex outputs per pixel)
    zs, ss = z[:256, :256].contiguous(), shp[:256, :256].contiguous()
    ks = float(ss.reshape(256 * 256, -1).sum(1).float().mean())  # SHP flags per pixel (the kernel skips those outside the corner)
    flops("emperical_co", dtime(lambda: mr.emperical_co(zs, ss)), 256 * 256 * 4.0 * ks * N * (N + 1), 2500.0,
          "256x256 raster, cov + coh (N x N each)")
    ii, jj = torch.meshgrid(torch.arange(512, device="cuda", dtype=torch.int32),
                            torch.arange(512, device="cuda", dtype=torch.int32), indexing="ij")
    idx = torch.stack([ii.reshape(-1), jj.reshape(-1)], -1).contiguous()
    pcs = shp[:512, :512].reshape(-1, aw, rw).contiguous()
    npt = idx.shape[0]
    flops("emperical_co_pc", dtime(lambda: mr.emperical_co_pc(z, idx, pcs)), npt * co_fl, 2500.0,
          f"{npt} points (512x512), full pair network")
    z0, z1 = z[..., 0].contiguous(), z[..., 1].contiguous()
    hbm("ad_intf_pc", dtime(lambda: mr.ad_intf_pc(z0, z1, idx, pcs)), npt * (W + 8 + 8 + 16),
        f"{npt} points, one pair (SHP mask + index in, intf out, each pixel's two samples once)")
    del z0, z1
    coh = mr.emperical_co_pc(z, idx, pcs)
    ch = coh[:65536].contiguous()
    hbm("uncompress_coh", dtime(lambda: mr.uncompress_coh(ch)), 65536 * (8 * P + 8 * N * N),
        "65536 points, (P) -> (N, N)")
    emi_fl = emi_algorithmic_flops(N)
    flops("emi", dtime(lambda: mr.emi(coh)), npt * emi_fl, FP64_PEAK_TFLOPS, f"{npt} points")
    ph, q = mr.emi(coh)
    hbm("ds_temp_coh", dtime(lambda: mr.ds_temp_coh(coh, ph)), npt * (8 * P + 8 * N + 4), f"{npt} points")
    rs = z[:512, :512].reshape(-1, N).contiguous()
    hbm("temp_coh", dtime(lambda: mr.temp_coh(coh, rs)), npt * (8 * P + 8 * N + 4), f"{npt} points")
    flops("emi_ds_temp_coh", dtime(lambda: emi_ds_temp_coh(coh)), npt * emi_fl, FP64_PEAK_TFLOPS,
          f"{npt} points, EMI flops only")
    t = dtime(lambda: mr.emperical_co_emi_temp_coh_pc(z, idx, pcs, batch_size=npt), reps=2)
    rows["emperical_co_emi_temp_coh_pc"] = {"s": t, "points/s": npt / t, "per_call": f"{npt} points, one batch"}
    # the reference's own GPU EMI algorithm (moraine/pl.py:18-42: float32 inv of |Gamma|, complex64
    # eigh of inv o Gamma) run through the vendor library (rocSOLVER via torch) on the same rows:
    # what a library-call port of moraine's CuPy path gives on this GPU
    nref = 8192  # rocSOLVER's batched eigh takes about a millisecond per matrix at N = 60
    cr = coh[:nref].contiguous()

    def emi_reference_algorithm():
        G = mr.uncompress_coh(cr)
        M = torch.linalg.inv(G.abs()).to(torch.complex64) * G
        _, V = torch.linalg.eigh(M)
        v = V[..., :, 0] * V[..., 0:1, 0].conj()
        return v / v.abs()

    progress("rows: reference EMI algorithm through rocSOLVER")
    emi_reference_algorithm()  # library warm-up
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    emi_reference_algorithm()
    e1.record()
    e1.synchronize()
    t_ref = e0.elapsed_time(e1) / 1e3
    t_ours = dtime(lambda: mr.emi(cr))
    rows["emi_reference_algorithm_rocsolver"] = {
        "s": t_ref, "points/s": nref / t_ref, "vs_emi_kernel": t_ref / t_ours,
        "per_call": f"{nref} points, N=60: uncompress, float32 inv(|Gamma|), complex64 eigh (rocSOLVER, batched)"}
    for n2, npt2 in ((120, 32768), (200, 16384)):
        c2 = torch.randn(npt2, n2 * (n2 - 1) // 2, dtype=torch.complex64, device="cuda") * 0.1
        flops(f"emi_n{n2}", dtime(lambda: mr.emi(c2), reps=2), npt2 * emi_algorithmic_flops(n2),
              FP64_PEAK_TFLOPS, f"{npt2} points, N={n2} (random coherence rows)")
        del c2
    return {"workload": f"N={N}, 15x21 windows, synthetic DS stack; each function timed alone (HIP events)",
            "rows": rows}


def _phase_rms(ph, ph64, torch):
    err = torch.angle(ph.to(torch.complex128) * ph64.conj())
    return torch.sqrt(torch.nanmean(err ** 2, dim=1))


def _rms_stats(rms, torch):
    ok = rms[~torch.isnan(rms)]
    return {"median": float(ok.median()), "p99": float(torch.quantile(ok, 0.99)), "max": float(ok.max()),
            "all": float(torch.sqrt(torch.mean(ok ** 2))), "points": int(ok.numel())}


def config3_leg(args, torch):
    """Precision workload: 1000x4000 tile, N=120, 21x41 windows, gamma0 ~ U[0.2, 0.9] per pixel and
    2% near-singular pixels (gamma0 in [0.98, 0.999] at every lag).  The plain drop-in pass (KS +
    selection with p materialised, coherence, EMI + temporal coherence in 65536-point batches) and
    the same pass with emi(refine=1e-3); per-point phase error against an FP64 solve (complex128
    inverse and eigh) on a sample of the last batch, beside the reference's own arithmetic (float32
    inv, complex64 eigh on the CPU) on the same points."""
    import moraine_amd as mr
    from moraine_amd import synth
    from moraine_amd.pl import _emi_tcoh_dev as emi_tcoh

    nl, w, n, az, r = args.c3_lines, 4000, 120, 10, 20
    aw, rw = 2 * az + 1, 2 * r + 1
    z, ns = synth.ds_stack_mixed_torch(nl, w, n, seed=3, device="cuda")
    npx = nl * w
    ii, jj = torch.meshgrid(torch.arange(nl, device="cuda", dtype=torch.int32),
                            torch.arange(w, device="cuda", dtype=torch.int32), indexing="ij")
    idx = torch.stack([ii.reshape(-1), jj.reshape(-1)], -1).contiguous()
    del ii, jj
    ph = torch.empty((npx, n), dtype=torch.complex64, device="cuda")
    q = torch.empty(npx, dtype=torch.float32, device="cuda")
    tc = torch.empty(npx, dtype=torch.float32, device="cuda")
    bs = 1 << 16
    st = {}

    def run(tol):
        x = torch.view_as_real(z)
        rmli = (x[..., 0] * x[..., 0] + x[..., 1] * x[..., 1]).contiguous()
        p, m, cnt = mr.ks_test_select_shp(rmli, az, r, 0.05, return_p=True)
        del rmli, p
        pcs = m.view(npx, aw, rw)
        for s0 in range(0, npx, bs):
            e0 = min(npx, s0 + bs)
            c = mr.emperical_co_pc(z, idx[s0:e0], pcs[s0:e0])
            emi_tcoh(c, n, 0, ph[s0:e0], q[s0:e0], tc[s0:e0], tol)
        st["coh"], st["s0"], st["cnt"] = c, s0, cnt

    # warm-up on a few lines (kernel loading), then one timed pass per form
    out = {"workload": f"{nl}x{w} tile, N={n}, {aw}x{rw} windows, gamma0~U[0.2,0.9], 2% near-singular, "
                       "p_max=0.05, all pixels", "pixels": npx}
    for name, tol in (("plain", None), ("refine_1e-3", 1e-3)):
        progress(f"precision workload: {name} pass")
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        run(tol)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        c, s0 = st["coh"], st["s0"]
        m = min(512, c.shape[0])
        if "fp64" not in st:
            G = mr.uncompress_coh(c[:m]).to(torch.complex128)
            M = torch.linalg.inv(G.abs()).to(torch.complex128) * G
            _, V = torch.linalg.eigh(M)
            v = V[..., :, 0] * V[..., 0:1, 0].conj()
            st["fp64"] = v / v.abs()
            G32 = G.to(torch.complex64).cpu()
            _, V32 = torch.linalg.eigh(torch.linalg.inv(G32.abs()).to(torch.complex64) * G32)
            v32 = (V32[..., :, 0] * V32[..., 0:1, 0].conj()).to(torch.complex128)
            out["phase_rms_vs_fp64_reference_fp32_lapack"] = _rms_stats(
                _phase_rms(v32 / v32.abs(), st["fp64"].cpu(), torch), torch)
        rms = _phase_rms(ph[s0:s0 + m], st["fp64"], torch)
        nsm = ns.reshape(-1)[s0:s0 + m]
        out[name] = {"s": dt, "value": npx / dt, "unit": "DS pixels/s", "phase_rms_vs_fp64": _rms_stats(rms, torch),
                     "near_singular_points_max_rms": float(rms[nsm].max()) if bool(nsm.any()) else None}
        if tol is not None:
            out[name]["re_solved_last_batch"] = int(mr.pl.last_refined)
    out["mean_shp_count"] = float(st["cnt"].float().mean())
    out["max_shp_count"] = int(st["cnt"].max())
    return out


def serving_leg(args, torch):
    """Poisson stream (50 req/s) of patch requests, N in {30, 60, 120, 200}, square
    tiles 128..serve_edge, windows 11x15..21x41, through the continuous-batching
    server and through the plain functions one request at a time."""
    from moraine_amd.serve import DSServer, run_serial, run_stream, stream_specs, stream_stats

    warm = stream_specs(4, seed=99, edge_range=(64, 64))
    run_stream(warm, DSServer(keep_outputs=False))
    run_serial(warm)
    specs = stream_specs(args.serve_requests, rate=50.0, seed=11, edge_range=(128, args.serve_edge))
    srv = run_stream([type(x)(**x.__dict__) for x in specs], DSServer(keep_outputs=False))
    ser = run_serial(specs)
    a, b = stream_stats(srv, specs), stream_stats(ser, specs)
    return {
        "workload": f"{args.serve_requests} requests at 50 req/s, N in {{30,60,120,200}}, tiles 128^2.."
                    f"{args.serve_edge}^2, windows 11x15..21x41, seed per request",
        "served": a, "serial_plain": b, "throughput_vs_serial": a["px_per_s"] / b["px_per_s"],
    }


# ---------------------------------------------------------------------------- GPU
def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "single-GPU build: only --gpus 1 is supported (tile sharding "
                                   "across GPUs is not built)"}))
        sys.exit(2)
    ensure_built()
    import numpy as np
    import torch

    import moraine_amd as mr
    from moraine_amd import synth
    from moraine_amd.pl import co_emi_tcoh_batches

    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    S, N, az, r = args.size, args.n, args.az, args.r
    aw, rw = 2 * az + 1, 2 * r + 1
    npx = S * S

    rslc = synth.ds_stack_torch(S, S, N, seed=0, ps_frac=args.ps, device=dev)
    ii, jj = torch.meshgrid(torch.arange(S, device=dev, dtype=torch.int32),
                            torch.arange(S, device=dev, dtype=torch.int32), indexing="ij")
    idx = torch.stack([ii.reshape(-1), jj.reshape(-1)], -1).contiguous()
    del ii, jj
    ph = torch.empty((npx, N), dtype=torch.complex64, device=dev)
    q = torch.empty(npx, dtype=torch.float32, device=dev)
    tc = torch.empty(npx, dtype=torch.float32, device=dev)
    torch.cuda.synchronize()

    batches = [(s, min(npx, s + args.batch)) for s in range(0, npx, args.batch)]
    ev = {}

    def mark(name):
        if ev is None:
            return
        e = torch.cuda.Event(enable_timing=True)
        e.record()
        ev.setdefault(name, []).append(e)

    state = {}

    def step(instrument=False):
        nonlocal ev
        if not instrument:
            ev = None
        else:
            ev = {}
        mark("start")
        x = torch.view_as_real(rslc)
        rmli = (x[..., 0] * x[..., 0] + x[..., 1] * x[..., 1]).contiguous()
        mark("rmli")
        p, is_shp, shp_num = mr.ks_test_select_shp(rmli, az, r, 0.05, return_p=True)
        mark("ks")
        pcs = is_shp.view(npx, aw, rw)
        mark("co0")
        # emperical_co_emi_temp_coh_pc's batches (coherence, EMI, temporal coherence; one launch
        # per batch for 32 < N <= 64, batches alternating over two streams so one batch's
        # Householder pass and re-solve overlap the next launch), writing straight into ph, q, tc
        co_emi_tcoh_batches(rslc, idx, pcs, 0, ph, q, tc, args.batch, args.batch_streams)
        mark("emi1")
        state["p"], state["shp_num"], state["pcs"] = p, shp_num, pcs
        if mr.pl.last_householder is not None:
            state["householder_last"] = mr.pl.last_householder
            state["resolved_last"] = mr.pl.last_resolved

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    times = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ts = time.perf_counter()
        step()
        torch.cuda.synchronize()
        times.append(time.perf_counter() - ts)
    t1 = time.perf_counter()
    dt = (t1 - t0) / args.steps

    # instrumented step: per-stage device time from HIP events on the launch stream
    step(instrument=True)
    torch.cuda.synchronize()
    e = ev

    def span(a, b):
        return sum(x.elapsed_time(y) for x, y in zip(e[a], e[b])) / 1e3

    t_rmli = span("start", "rmli")
    t_ks = span("rmli", "ks")
    t_cet = span("co0", "emi1")
    nb = len(batches)

    shp_num = state["shp_num"].float()
    kmean = float(shp_num.mean())
    # algorithmic units per pixel: bytes for ks/tcoh, flops for coherence/EMI
    ks_bytes = 4 * N + 4 * aw * rw + aw * rw + 4  # rmli in, p out, mask out, count out
    co_flops = 4.0 * kmean * N * (N + 1)  # Hermitian rank-K update (unique entries)
    emi_flops = emi_algorithmic_flops(N)
    stages = {
        "rmli": {"s": t_rmli},
        "ks_test+select_shp": {"s": t_ks, "GB/s": ks_bytes * npx / t_ks / 1e9,
                               "hbm_frac": ks_bytes * npx / t_ks / 1e9 / HBM_PEAK_GBS},
        "emperical_co_pc+emi+ds_temp_coh": {
            "s": t_cet, "emi_TFLOP/s": emi_flops * npx / t_cet / 1e12, "coh_TFLOP/s": co_flops * npx / t_cet / 1e12,
            "note": "one launch per batch computes each point's coherence into LDS, the shift-invert EMI "
                    "pass and the temporal coherence; the Householder pass and FP64 re-solve follow"},
    }
    dom = max(["ks_test+select_shp", "emperical_co_pc+emi+ds_temp_coh"], key=lambda k: stages[k]["s"])
    if dom == "emperical_co_pc+emi+ds_temp_coh":
        ach = emi_flops * (npx / nb) / (t_cet / nb) / 1e12
        # the fused launch keeps 60 EMI columns per lane up to N = 60, else 64
        ek = [f"co_emi_pc_kernel<{60 if N <= 60 else 64}>", "emi_fast_kernel<64, true, true, 2>",
              "emi_general_kernel<true, float2 >"]
        # the same launches also do the coherence's 4KN(N+1) flops (MFMA work): the share of the
        # stage time an ideal machine would need for both at their peaks
        ideal = (emi_flops / (FP64_PEAK_TFLOPS * 1e12) + co_flops / 2.5e15) * npx
        roof = {"kernel": " + ".join(ek) if 32 < N <= 64 else "co + emi kernels", "bound": "valu", "achieved": ach,
                "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": ach / FP64_PEAK_TFLOPS,
                "frac_with_coherence": ideal / t_cet,
                "traffic": measured_traffic(ek, args),
                "note": "coherence, EMI and the DS temporal coherence: the fused launch (coherence on MFMA into LDS, "
                        "the shift-invert EMI pass, temporal coherence) over every point, the Householder pass over "
                        "the points it leaves and the FP64 re-solve of those the Householder pass flags, timed "
                        "together with HIP events around the batch loop (batches alternate over two streams, "
                        "so a batch's tail overlaps the next launch; per batch = stage time / batches); "
                        "VALU-bound (EMI); achieved = EMI algorithmic flops "
                        "N^3 + (4/3)N^3 + 16N^2 per point only (the coherence's MFMA flops are in "
                        "frac_with_coherence); the passes execute more (an FP64 sweep, 2N^3, and a complex FP32 "
                        "sweep, 8N^3, in place of Cholesky factors); peak = MI355X FP64 = FP32 non-packed vector rate"}
    elif dom == "ks_test+select_shp":
        ach = ks_bytes * npx / t_ks / 1e9
        ks = ["ks_sort_kernel<float, 1, false>", "ks_pairs_kernel<float, 4, unsigned char>",
              "ks_emit_kernel<float, unsigned char>"]
        roof = {"kernel": " + ".join(ks), "bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": ach / HBM_PEAK_GBS, "traffic": measured_traffic(ks, args)}

    # phase error of the last batch vs an FP64 solve (torch/rocSOLVER, complex128) on a sample;
    # the last batch's coherence rows (the fused launch keeps them on chip) computed again here,
    # outside the timing
    s0, s1 = batches[-1]
    coh = mr.emperical_co_pc(rslc, idx[s0:s1], state["pcs"][s0:s1])
    m = min(args.err_sample, coh.shape[0])
    G = mr.uncompress_coh(coh[:m]).to(torch.complex128)
    A = G.abs()
    M = torch.linalg.inv(A).to(torch.complex128) * G
    lam, V = torch.linalg.eigh(M)
    v = V[..., :, 0] * V[..., 0:1, 0].conj()
    ph64 = v / v.abs()
    err = torch.angle(ph[s0:s0 + m].to(torch.complex128) * ph64.conj())
    rms = torch.sqrt(torch.nanmean(err ** 2, dim=1))
    phase = {"rms_median_rad": float(torch.nanmedian(rms)), "rms_p99_rad": float(torch.nanquantile(rms, 0.99)),
             "rms_max_rad": float(rms[~torch.isnan(rms)].max()),
             "rms_all_rad": float(torch.sqrt(torch.nanmean(rms ** 2))), "points": m}
    # the reference's own arithmetic on the same points (moraine/pl.py:57-59: float32 inv of |Gamma|,
    # complex64 eigh; CPU LAPACK through torch) against the same FP64 solve
    G32 = G.to(torch.complex64).cpu()
    M32 = torch.linalg.inv(G32.abs()).to(torch.complex64) * G32
    _, V32 = torch.linalg.eigh(M32)
    v32 = (V32[..., :, 0] * V32[..., 0:1, 0].conj()).to(torch.complex128)
    err32 = torch.angle((v32 / v32.abs()) * ph64.conj().cpu())
    rms32 = torch.sqrt(torch.nanmean(err32 ** 2, dim=1))
    phase["reference_fp32_lapack"] = {
        "rms_median_rad": float(torch.nanmedian(rms32)), "rms_p99_rad": float(torch.nanquantile(rms32, 0.99)),
        "rms_max_rad": float(rms32[~torch.isnan(rms32)].max()),
        "rms_all_rad": float(torch.sqrt(torch.nanmean(rms32 ** 2))),
        "note": "float32 inv(|Gamma|) and complex64 eigh (LAPACK, CPU) on the same coherence rows"}
    # how often the matrices EMI factors are positive definite on this workload: a Cholesky-based
    # formulation (FP64 Cholesky of |Gamma|, shift-invert iteration on M) needs both
    Gs = mr.uncompress_coh(coh[:m])
    phase["abs_gamma_not_pd_frac"] = float((~mr.isPD(Gs.abs().to(torch.float64))).float().mean())
    phase["lambda_min_negative_frac"] = float((q[s0:s0 + m] < 0).float().mean())
    phase["k_below_n_frac"] = float((state["shp_num"].reshape(-1)[s0:s0 + m] < N).float().mean())
    if "householder_last" in state:  # points of the last batch the shift-invert EMI left to Householder
        phase["householder_pass_frac"] = int(state["householder_last"]) / coh.shape[0]
        phase["fp64_resolve_frac"] = int(state["resolved_last"]) / coh.shape[0]
    del Gs
    phase["gpu_over_reference"] = {
        "p99": phase["rms_p99_rad"] / phase["reference_fp32_lapack"]["rms_p99_rad"],
        "max": phase["rms_max_rad"] / phase["reference_fp32_lapack"]["rms_max_rad"]}
    # precision form on the same batch: FP64 re-solve of points whose estimated
    # phase error exceeds tol (emi(..., refine=tol)); error on the same sample
    if 32 < args.n <= 64:
        from moraine_amd.pl import emi_ds_temp_coh
        tol = 1e-3
        emi_ds_temp_coh(coh, refine=tol)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        ph_r, _, _ = emi_ds_temp_coh(coh, refine=tol)
        e1.record()
        torch.cuda.synchronize()
        k = int(mr.pl.last_refined)
        err_r = torch.angle(ph_r[:m].to(torch.complex128) * ph64.conj())
        rms_r = torch.sqrt(torch.nanmean(err_r ** 2, dim=1))
        phase["refined"] = {"tol_rad": tol, "ms_per_batch": e0.elapsed_time(e1), "points_per_batch": coh.shape[0],
                            "re_solved": k, "rms_max_rad": float(rms_r[~torch.isnan(rms_r)].max()),
                            "rms_all_rad": float(torch.sqrt(torch.nanmean(rms_r ** 2))),
                            "note": "EMI + temporal coherence of the last batch with refine=tol, HIP events"}

    progress("plain pass timed; CPU baseline")
    cpu = None if args.no_cpu_baseline else cpu_baseline(args, torch)
    forms = {}
    wanted = {f.strip() for f in args.forms.split(",") if f.strip()}
    del state, rslc, idx, ph, q, tc, coh
    torch.cuda.empty_cache()
    if "batched" in wanted:
        progress("batched form")
        forms["batched"] = batched_leg(args, torch)
        torch.cuda.empty_cache()
    if "serving" in wanted:
        progress("serving form")
        forms["serving"] = serving_leg(args, torch)
        torch.cuda.empty_cache()
    if "rows" in wanted:
        progress("per-function rows")
        forms["rows"] = rows_leg(args, torch)
        torch.cuda.empty_cache()
    if "streamed" in wanted:
        progress("streamed form")
        forms["streamed"] = streamed_leg(args, torch)
        torch.cuda.empty_cache()
    if "precision" in wanted:
        progress("precision workload")
        forms["precision"] = config3_leg(args, torch)
        torch.cuda.empty_cache()
    value = npx / dt
    out = {
        "metric": "DS pixels/s (KS+coh+EMI+tcoh) at N=60; p50/p99 latency; phase err vs FP64",
        "value": value, "unit": "DS pixels/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": dt * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 (KS: sorted on int keys, integer step counts; coherence: fp16 hi+lo MFMA, f32 sums; EMI: f64 inverse + f32 eigen)", "data": "synthetic (seeded DS/PS model)",
        "config": {"workload": f"plain drop-in pipeline, N={N}, {S}x{S} tile, {aw}x{rw} window, "
                               f"{int(args.ps*100)}% PS, p_max=0.05, all pixels", "pixels": npx,
                   "n_images": N, "window": [aw, rw], "batch": args.batch, "batch_streams": args.batch_streams},
        "latency_ms": {"p50": float(np.percentile(times, 50) * 1e3), "p99": float(np.percentile(times, 99) * 1e3),
                       "of": "plain-form tile steps; request latency of the serving form is under forms.serving"},
        "phase_err_vs_fp64": phase,
        "mean_shp_count": kmean,
        "stages": stages,
        "roofline": roof,
        "cpu_baseline": cpu,
        "forms": forms,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
