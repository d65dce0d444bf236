"""Serving form: a continuous-batching request loop for patch solves.

A request is one patch: an SLC stack (h, w, N) complex64 and a search window
(2*az+1, 2*r+1).  Its answer is what the drop-in pipeline gives for every pixel
of the patch: KS + SHP selection (moraine/shp.py:269-323), coherence
(moraine/co.py:230-266), EMI (moraine/pl.py:68-77) and DS temporal coherence
(moraine/pl.py:126-145).

The loop (:class:`DSServer`):
  * admits queued requests in arrival order while their device footprint fits
    the HBM budget (free HBM at start-up, less the shared coherence buffer);
    admission uploads (or generates) the stack and runs the KS test on its own
    stream, so it overlaps the DS work of requests already on the device;
  * keeps every admitted request's stack, SHP mask and outputs resident until
    it completes;
  * runs "ticks": each tick takes pixels from the admitted requests (oldest
    first) up to a pixel budget and a coherence-buffer budget, computes their
    coherence into one shared buffer (one launch per request chunk), then one
    EMI and one temporal-coherence launch per image count over the whole tick,
    and scatters the results into the requests' outputs; for 32 < N <= 64 each
    request chunk instead runs coherence, EMI and temporal coherence in one
    launch (mr_co_emi_ds_temp_coh_pc) straight into the request's outputs,
    chunks alternating over two side streams;
  * keeps up to ``max_inflight`` ticks queued on the device and frees a request
    (and its HBM) as soon as the tick holding its last pixels has finished.
Every kernel is per-pixel and deterministic, so a request's result does not
depend on what else shares its ticks (tested bit for bit).
"""
from __future__ import annotations

import math
import time
from collections import deque
from dataclasses import dataclass, field

import numpy as np
import torch

from . import _dev, _lib
from .co import _pairs
from .pl import _side_streams


@dataclass
class PatchSpec:
    """A patch-solve request.  ``rslc`` (h, w, n) complex64 numpy/torch, or None
    for a seeded synthetic stack (DS model with 10% PS-like pixels) generated
    on the device at admission."""
    rid: int
    n: int
    h: int
    w: int
    az: int
    r: int
    seed: int = 0
    arrival: float = 0.0
    rslc: object = None
    p_max: float = 0.05

    @property
    def npx(self) -> int:
        return self.h * self.w

    @property
    def win(self) -> int:
        return (2 * self.az + 1) * (2 * self.r + 1)


def footprint_bytes(s: PatchSpec) -> int:
    """Peak device bytes a request holds while admitted: stack, |z|^2 and KS
    workspace during admission; SHP mask, counts, pixel indices and outputs
    until it completes."""
    px = s.npx
    stack = px * s.n * 8
    ks = px * s.n * 4 + px * s.n * 8 + px
    keep = px * s.win + px * 4 + px * 8 + px * s.n * 8 + 2 * px * 4
    return stack + ks + keep


@dataclass
class PatchResult:
    rid: int
    ph: torch.Tensor | None
    quality: torch.Tensor | None
    t_coh: torch.Tensor | None
    shp_num: torch.Tensor | None
    arrival: float
    admitted: float
    done: float

    @property
    def latency(self) -> float:
        return self.done - self.arrival


# ---------------------------------------------------------------------------- policy (pure, CPU-testable)
def admit_order(queue, used: int, budget: int, max_active: int, n_active: int):
    """FIFO admission: take queued (rid, bytes) in order while they fit; stop at
    the first that does not (a large request is not overtaken forever)."""
    out = []
    for rid, nb in queue:
        if n_active + len(out) >= max_active or used + nb > budget:
            break
        out.append(rid)
        used += nb
    return out


def plan_tick(active, px_budget: int, coh_budget: int):
    """Split a tick among admitted requests, oldest first.

    ``active``: (rid, n, remaining_px) in age order.  Returns (rid, take) with
    sum(take) <= px_budget and sum(take * 8 n(n-1)/2) <= coh_budget.
    """
    plan = []
    px_left, coh_left = px_budget, coh_budget
    for rid, n, rem in active:
        if rem <= 0:
            continue
        per = 8 * (n * (n - 1) // 2)
        take = min(rem, px_left, coh_left // per)
        if take <= 0:
            break
        plan.append((rid, take))
        px_left -= take
        coh_left -= take * per
        if px_left == 0:
            break
    return plan


# ---------------------------------------------------------------------------- server
@dataclass
class _Active:
    spec: PatchSpec
    rslc: torch.Tensor
    mask: torch.Tensor  # (npx, aw, rw) bool
    shp_num: torch.Tensor
    idx: torch.Tensor  # (npx, 2) int32
    ph: torch.Tensor  # (npx, n) complex64
    q: torch.Tensor
    t: torch.Tensor
    ready: torch.cuda.Event
    admitted: float
    nbytes: int
    cursor: int = 0
    last_tick: int = -1


class DSServer:
    def __init__(self, device=None, hbm_budget: int | None = None, coh_budget: int = 8 << 30,
                 max_tick_px: int = 1 << 20, max_inflight: int = 2, max_active: int = 64, keep_outputs: bool = True,
                 ps_frac: float = 0.1):
        self.dev = torch.device(device) if device is not None else _dev.device()
        self.coh_budget = int(coh_budget)
        self.coh_buf = torch.empty(self.coh_budget // 8, dtype=torch.complex64, device=self.dev)
        if hbm_budget is None:
            free, _ = torch.cuda.mem_get_info(self.dev)
            hbm_budget = int(free * 0.8)
        self.hbm_budget = int(hbm_budget)
        self.max_tick_px = int(max_tick_px)
        self.max_inflight = int(max_inflight)
        self.max_active = int(max_active)
        self.keep_outputs = keep_outputs
        self.ps_frac = ps_frac
        self.queue: deque[PatchSpec] = deque()
        self.active: dict[int, _Active] = {}
        self.used = 0
        self.inflight: deque = deque()  # (event, tick id)
        self.tick_id = 0
        self.done_ticks = -1
        self.results: list[PatchResult] = []
        self.compute = torch.cuda.current_stream(self.dev)
        self.admit_stream = torch.cuda.Stream(self.dev)  # default priority
        self._pairs: dict[int, torch.Tensor] = {}
        self._emi_ws = None
        self._fused_ws = [None, None]
        self.t0 = time.perf_counter()
        self.max_concurrent = 0
        self.peak_used = 0

    # ------------------------------------------------------------ requests
    def submit(self, spec: PatchSpec, now: float | None = None):
        if spec.n < 2:
            raise ValueError("request needs at least 2 images")
        if 8 * (spec.n * (spec.n - 1) // 2) > self.coh_budget:
            raise ValueError(f"request {spec.rid}: one pixel's coherence exceeds the tick buffer")
        if footprint_bytes(spec) > self.hbm_budget:
            raise ValueError(f"request {spec.rid}: {footprint_bytes(spec)} bytes exceed the HBM budget")
        spec.arrival = (time.perf_counter() - self.t0) if now is None else now
        self.queue.append(spec)

    def _pairs_for(self, n: int) -> torch.Tensor:
        if n not in self._pairs:
            self._pairs[n] = _dev.to_dev(_pairs(None, n)).to(self.dev)
        return self._pairs[n]

    def _admit(self, s: PatchSpec, now: float):
        from . import shp, synth

        with torch.cuda.stream(self.admit_stream):
            if s.rslc is None:
                z = synth.ds_stack_torch(s.h, s.w, s.n, seed=s.seed, ps_frac=self.ps_frac, device=self.dev)
            else:
                z = _dev.to_dev(s.rslc, torch.complex64).to(self.dev)
                if tuple(z.shape) != (s.h, s.w, s.n):
                    raise ValueError(f"request {s.rid}: stack shape {tuple(z.shape)} != ({s.h}, {s.w}, {s.n})")
            x = torch.view_as_real(z)
            rmli = (x[..., 0] * x[..., 0] + x[..., 1] * x[..., 1]).contiguous()
            _, mask, cnt = shp.ks_test_select_shp(rmli, s.az, s.r, s.p_max, return_p=False)
            del rmli
            ii, jj = torch.meshgrid(torch.arange(s.h, device=self.dev, dtype=torch.int32),
                                    torch.arange(s.w, device=self.dev, dtype=torch.int32), indexing="ij")
            idx = torch.stack([ii.reshape(-1), jj.reshape(-1)], -1).contiguous()
            ph = torch.empty((s.npx, s.n), dtype=torch.complex64, device=self.dev)
            q = torch.empty(s.npx, dtype=torch.float32, device=self.dev)
            t = torch.empty(s.npx, dtype=torch.float32, device=self.dev)
            ev = torch.cuda.Event()
            ev.record(self.admit_stream)
        # tensors made on the admission stream are used on the compute stream
        for tt in (z, mask, cnt, idx, ph, q, t):
            tt.record_stream(self.compute)
        aw, rw = 2 * s.az + 1, 2 * s.r + 1
        a = _Active(s, z, mask.view(s.npx, aw, rw), cnt, idx, ph, q, t, ev, now, footprint_bytes(s))
        self.active[s.rid] = a
        self.used += a.nbytes
        self.peak_used = max(self.peak_used, self.used)
        self.max_concurrent = max(self.max_concurrent, len(self.active))

    # ------------------------------------------------------------ ticks
    def _emi_workspace(self, rows: int, n: int):
        need = _lib.query("mr_emi_workspace_bytes", rows, n)
        if need == 0:
            return None
        if self._emi_ws is None or self._emi_ws.numel() < need:
            self._emi_ws = _dev.workspace(need, self.dev)
        return self._emi_ws

    def _fused_workspace(self, rows: int, n: int, slot: int):
        """Workspace of side stream `slot` (allocated on that stream, which is current)."""
        need = _lib.query("mr_co_emi_ds_temp_coh_pc_workspace_bytes", rows, n)
        if self._fused_ws[slot] is None or self._fused_ws[slot].numel() < need:
            self._fused_ws[slot] = _dev.workspace(need, self.dev)
        return self._fused_ws[slot]

    def _run_tick(self, plan):
        with torch.cuda.stream(self.compute):
            self._run_tick_on_stream(plan)

    def _run_tick_on_stream(self, plan):
        by_n: dict[int, list] = {}
        for rid, take in plan:
            a = self.active[rid]
            by_n.setdefault(a.spec.n, []).append((a, take))
        st = _dev.stream()
        for a, _ in (x for v in by_n.values() for x in v):
            self.compute.wait_event(a.ready)
        off_bytes = 0
        for n, chunks in by_n.items():
            if 32 < n <= 64:
                # coherence, EMI and the DS temporal coherence of each request chunk in one
                # launch (mr_co_emi_ds_temp_coh_pc), straight into the request's outputs; the
                # chunks alternate over the two side streams (forked from and joined back to
                # the compute stream), so one chunk's Householder pass and re-solve overlap the
                # next chunk's launch, as in the drop-in driver
                sides = _side_streams(self.dev)
                cur = torch.cuda.current_stream(self.dev)
                for sd in sides:
                    sd.wait_stream(cur)
                for i, (a, take) in enumerate(chunks):
                    s = a.spec
                    p0 = a.cursor
                    with torch.cuda.stream(sides[i % 2]):
                        ws = self._fused_workspace(take, n, i % 2)
                        _lib.call("mr_co_emi_ds_temp_coh_pc", _dev.ptr(a.rslc), s.h, s.w, n,
                                  _dev.ptr(a.idx[p0:p0 + take]), _dev.ptr(a.mask[p0:p0 + take]), take,
                                  2 * s.az + 1, 2 * s.r + 1, 0, _dev.ptr(a.ph[p0:p0 + take]),
                                  _dev.ptr(a.q[p0:p0 + take]), _dev.ptr(a.t[p0:p0 + take]), _dev.ptr(ws),
                                  ws.numel(), _dev.stream())
                    a.cursor += take
                    a.last_tick = self.tick_id
                for sd in sides:
                    cur.wait_stream(sd)
                continue
            P = n * (n - 1) // 2
            rows = sum(t for _, t in chunks)
            base = off_bytes // 8
            coh = self.coh_buf[base:base + rows * P].view(rows, P)
            off_bytes += rows * P * 8
            pairs = self._pairs_for(n)
            r0 = 0
            for a, take in chunks:
                s = a.spec
                p0 = a.cursor
                _lib.call("mr_emperical_co_pc", _dev.ptr(a.rslc), s.h, s.w, n, _dev.ptr(a.idx[p0:p0 + take]),
                          _dev.ptr(a.mask[p0:p0 + take]), take, 2 * s.az + 1, 2 * s.r + 1, _dev.ptr(pairs), P,
                          _dev.ptr(coh[r0:r0 + take]), st)
                r0 += take
            ph = torch.empty((rows, n), dtype=torch.complex64, device=self.dev)
            q = torch.empty(rows, dtype=torch.float32, device=self.dev)
            t = torch.empty(rows, dtype=torch.float32, device=self.dev)
            ws = self._emi_workspace(rows, n)
            # EMI then the DS temporal coherence (one kernel for 32 < n <= 64)
            _lib.call("mr_emi_ds_temp_coh", _dev.ptr(coh), rows, n, 0, _dev.ptr(pairs), P, _dev.ptr(ph), _dev.ptr(q),
                      _dev.ptr(t), _dev.ptr(ws), 0 if ws is None else ws.numel(), st)
            r0 = 0
            for a, take in chunks:
                p0 = a.cursor
                a.ph[p0:p0 + take].copy_(ph[r0:r0 + take])
                a.q[p0:p0 + take].copy_(q[r0:r0 + take])
                a.t[p0:p0 + take].copy_(t[r0:r0 + take])
                a.cursor += take
                a.last_tick = self.tick_id
                r0 += take
        ev = torch.cuda.Event()
        ev.record(self.compute)
        self.inflight.append((ev, self.tick_id))
        self.tick_id += 1

    def _retire(self, now: float):
        while self.inflight and self.inflight[0][0].query():
            _, tid = self.inflight.popleft()
            self.done_ticks = tid
        self._finish(now)

    def _finish(self, now: float):
        for rid in [r for r, a in self.active.items() if a.cursor >= a.spec.npx and a.last_tick <= self.done_ticks]:
            a = self.active.pop(rid)
            s = a.spec
            keep = self.keep_outputs
            self.results.append(PatchResult(
                rid, a.ph.view(s.h, s.w, s.n) if keep else None, a.q.view(s.h, s.w) if keep else None,
                a.t.view(s.h, s.w) if keep else None, a.shp_num if keep else None, s.arrival, a.admitted, now))
            self.used -= a.nbytes

    def step(self) -> bool:
        """One scheduling round; returns True if it enqueued or retired work."""
        now = time.perf_counter() - self.t0
        did = False
        n0 = len(self.results)
        self._retire(now)
        did |= len(self.results) > n0
        # admission by free HBM, oldest first
        q = [(s.rid, footprint_bytes(s)) for s in self.queue]
        for rid in admit_order(q, self.used, self.hbm_budget, self.max_active, len(self.active)):
            s = self.queue.popleft()
            assert s.rid == rid
            self._admit(s, now)
            did = True
        # ticks: keep up to max_inflight queued on the device
        while len(self.inflight) < self.max_inflight:
            order = sorted(self.active.values(), key=lambda a: (a.spec.arrival, a.spec.rid))
            plan = plan_tick([(a.spec.rid, a.spec.n, a.spec.npx - a.cursor) for a in order], self.max_tick_px,
                             self.coh_budget)
            if not plan:
                break
            self._run_tick(plan)
            did = True
        return did

    def idle(self) -> bool:
        return not self.queue and not self.active and not self.inflight

    def drain(self):
        while not self.idle():
            if not self.step() and self.inflight:
                self.inflight[0][0].synchronize()


# ---------------------------------------------------------------------------- stream drivers
def stream_specs(n_req: int, rate: float = 50.0, seed: int = 0, n_choices=(30, 60, 120, 200),
                 edge_range=(128, 1024), az_range=(5, 10), r_range=(7, 20)):
    """Poisson stream of synthetic patch requests: N from ``n_choices``, square
    tiles with edge uniform in ``edge_range`` (multiples of 32), windows from
    (2*az_range[0]+1) x (2*r_range[0]+1) to (2*az_range[1]+1) x (2*r_range[1]+1),
    one seed per request."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, n_req))
    out = []
    for i in range(n_req):
        e = int(rng.integers(edge_range[0] // 32, edge_range[1] // 32 + 1)) * 32
        out.append(PatchSpec(i, int(rng.choice(n_choices)), e, e, int(rng.integers(az_range[0], az_range[1] + 1)),
                             int(rng.integers(r_range[0], r_range[1] + 1)), seed=1000 + i, arrival=float(t[i])))
    return out


def run_stream(specs, server: DSServer):
    """Replay ``specs`` (arrival offsets in seconds) against a server in real
    time; returns the server's results sorted by request id."""
    pending = deque(sorted(specs, key=lambda s: s.arrival))
    server.t0 = time.perf_counter()
    while pending or not server.idle():
        now = time.perf_counter() - server.t0
        while pending and pending[0].arrival <= now:
            s = pending.popleft()
            server.submit(s, now=s.arrival)
        did = server.step()
        if not did:
            if server.inflight:
                server.inflight[0][0].synchronize()
            elif pending:
                time.sleep(max(0.0, min(pending[0].arrival - (time.perf_counter() - server.t0), 0.01)))
    torch.cuda.synchronize()
    return sorted(server.results, key=lambda r: r.rid)


def run_serial(specs, ps_frac: float = 0.1, keep_outputs: bool = False):
    """The same stream through the plain drop-in functions (ks_test_select_shp,
    then emperical_co_emi_temp_coh_pc), one request at a time in arrival order
    (the request waits while the previous one runs)."""
    from . import pl, shp, synth

    dev = _dev.device()
    out = []
    t0 = time.perf_counter()
    for s in sorted(specs, key=lambda s: s.arrival):
        now = time.perf_counter() - t0
        if now < s.arrival:
            time.sleep(s.arrival - now)
        start = time.perf_counter() - t0
        z = (synth.ds_stack_torch(s.h, s.w, s.n, seed=s.seed, ps_frac=ps_frac, device=dev) if s.rslc is None
             else _dev.to_dev(s.rslc, torch.complex64))
        x = torch.view_as_real(z)
        rmli = (x[..., 0] * x[..., 0] + x[..., 1] * x[..., 1]).contiguous()
        _, mask, cnt = shp.ks_test_select_shp(rmli, s.az, s.r, s.p_max, return_p=False)
        del rmli
        ii, jj = torch.meshgrid(torch.arange(s.h, device=dev, dtype=torch.int32),
                                torch.arange(s.w, device=dev, dtype=torch.int32), indexing="ij")
        idx = torch.stack([ii.reshape(-1), jj.reshape(-1)], -1).contiguous()
        m = mask.view(s.npx, 2 * s.az + 1, 2 * s.r + 1)
        ph, q, t = pl.emperical_co_emi_temp_coh_pc(z, idx, m)
        torch.cuda.synchronize()
        done = time.perf_counter() - t0
        out.append(PatchResult(s.rid, ph.view(s.h, s.w, s.n) if keep_outputs else None,
                               q.view(s.h, s.w) if keep_outputs else None, t.view(s.h, s.w) if keep_outputs else None,
                               cnt if keep_outputs else None, s.arrival, start, done))
    return sorted(out, key=lambda r: r.rid)


def stream_stats(results, specs):
    """Throughput (DS pixels/s over the makespan) and latency percentiles."""
    px = sum(s.npx for s in specs)
    lat = np.array([r.latency for r in results])
    span = max(r.done for r in results) - min(s.arrival for s in specs)
    return {"requests": len(results), "pixels": int(px), "makespan_s": float(span),
            "px_per_s": float(px / span), "p50_latency_s": float(np.percentile(lat, 50)),
            "p99_latency_s": float(np.percentile(lat, 99)), "mean_latency_s": float(lat.mean())}
