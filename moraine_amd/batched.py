"""Batched DS form: many independent DS pixels with ragged inputs.

Each pixel carries its own complex64 sample matrix (K_i SHP stacks of N_i
images, row-major K_i x N_i) and goes through coherence -> EMI -> temporal
coherence, as moraine's batch driver does for a point cloud
(moraine/pl.py:148-168).  Pixels are grouped by image count; each group is one
call of ``mr_ds_batch`` (csrc/batch.hip): fused launches for N <= 200 (one wave
per pixel with the coherence in LDS for N <= 64, one workgroup per pixel with
the coherence row in a per-workgroup slot above), chunked coherence / EMI /
temporal-coherence launches for N > 200 or in the per-stage mode.  Launch
shapes (fused or staged, grid size of the fused kernel, chunk size, tile layout
of the staged coherence kernel) come from :class:`AutoTuner`, and a
:class:`DSBatchRunner` can capture all group launches in one HIP graph and
replay it on new data written into the same buffers.

Layout of a batch (all device tensors except ``n_images``):
  samples    complex64, flat; pixel i's rows start at complex element row_off[i]
  row_off    int64 (n_px,)
  n_samples  int32 (n_px,)   K_i
  n_images   int32 numpy (n_px,) N_i  (host copy: it decides the grouping)
Outputs: ph complex64 flat with pixel i's N_i phases at ph_off[i], quality and
t_coh float32 (n_px,).
"""
from __future__ import annotations

import json
import os
import time
from dataclasses import dataclass, field

import numpy as np
import torch

from . import _dev, _lib
from .co import _pairs


@dataclass
class DSBatch:
    samples: torch.Tensor
    row_off: torch.Tensor
    n_samples: torch.Tensor
    n_images: np.ndarray

    @property
    def n_px(self) -> int:
        return int(self.n_images.shape[0])

    @staticmethod
    def layout(n_images, n_samples):
        """(row_off, ph_off) int64 numpy offsets for pixels stored in index order."""
        n_images = np.asarray(n_images, np.int64)
        n_samples = np.asarray(n_samples, np.int64)
        row_off = np.zeros(n_images.size, np.int64)
        ph_off = np.zeros(n_images.size, np.int64)
        if n_images.size:
            row_off[1:] = np.cumsum(n_images * n_samples)[:-1]
            ph_off[1:] = np.cumsum(n_images)[:-1]
        return row_off, ph_off

    @classmethod
    def from_matrices(cls, mats, device=None) -> "DSBatch":
        """Batch from a list of (K_i, N_i) complex arrays (host or device)."""
        dev = device or _dev.device()
        n_img = np.array([m.shape[1] for m in mats], np.int32)
        n_smp = np.array([m.shape[0] for m in mats], np.int32)
        row_off, _ = cls.layout(n_img, n_smp)
        flat = [(m if _dev.is_torch(m) else torch.from_numpy(np.ascontiguousarray(m))).reshape(-1) for m in mats]
        samples = (torch.cat([f.to(dev, torch.complex64) for f in flat]) if flat
                   else torch.zeros(0, dtype=torch.complex64, device=dev))
        return cls(samples.contiguous(), torch.from_numpy(row_off).to(dev), torch.from_numpy(n_smp).to(dev), n_img)

    @classmethod
    def from_tensors(cls, samples, row_off, n_images, n_samples) -> "DSBatch":
        dev = samples.device
        return cls(samples.contiguous(), _dev.to_dev(row_off, torch.int64).to(dev),
                   _dev.to_dev(n_samples, torch.int32).to(dev), np.asarray(n_images, np.int32))


def raster_sample_matrices(rslc: np.ndarray, idx: np.ndarray, pc_is_shp: np.ndarray):
    """Sample matrices of point-cloud pixels: the in-bounds SHP stacks of each
    pixel's window in window order (az offset, then range offset), as
    emperical_co_pc consumes them (moraine/co.py:209-226)."""
    nl, w, n = rslc.shape
    aw, rw = pc_is_shp.shape[-2:]
    az, r = aw // 2, rw // 2
    out = []
    for (i, j), m in zip(np.asarray(idx), np.asarray(pc_is_shp)):
        rows = []
        for k in range(aw):
            for l in range(rw):
                a, b = i - az + k, j - r + l
                if m[k, l] and 0 <= a < nl and 0 <= b < w:
                    rows.append(rslc[a, b])
        out.append(np.asarray(rows, np.complex64).reshape(-1, n))
    return out


class AutoTuner:
    """Picks launch shapes by timing them on the workload itself.

    For each key (kernel family, N, pixel-count class) every candidate runs once
    untimed and then ``reps`` times under a timer; the fastest is kept.  Picks
    persist in memory and, when ``path`` is given (or MORAINE_AMD_TUNE_CACHE is
    set), in a JSON file, so a later process starts from them.
    """

    FUSED_BLOCKS = (0, 4096, 16384, 65536)  # one-wave fused launches, N <= 64 (0: one per pixel, capped)
    WIDE_BLOCKS = (0, 1024, 4096)  # workgroup-per-pixel fused launch, 64 < N <= 200 (0: 2048)
    CHUNKS = (4096, 8192, 16384, 32768, 65536)

    @staticmethod
    def fused_blocks(n: int) -> tuple:
        """Grid sizes tried for the fused launch of an N-image group."""
        return AutoTuner.WIDE_BLOCKS if n > 64 else AutoTuner.FUSED_BLOCKS

    @staticmethod
    def coh_waves(n: int) -> tuple:
        """Tile layouts of the per-stage coherence kernel for n images: waves per
        workgroup over its grid of 32 x 32 tiles (() where it has one layout)."""
        if not 64 < n <= 224:
            return ()
        return (2, 3, 4, 8) if n <= 96 else (4, 8)

    def __init__(self, path: str | None = None, reps: int = 2, timer=None):
        self.path = path or os.environ.get("MORAINE_AMD_TUNE_CACHE")
        self.reps = reps
        self.timer = timer or _device_time
        self.picks: dict[str, int] = {}
        self.trials: dict[str, dict[int, float]] = {}
        if self.path and os.path.exists(self.path):
            with open(self.path) as f:
                self.picks.update({k: int(v) for k, v in json.load(f).items()})

    @staticmethod
    def key(kind: str, n: int, count: int) -> str:
        return f"{kind}:n={n}:px<={1 << max(int(count) - 1, 0).bit_length()}"

    def choose(self, key: str, candidates, run) -> int:
        if key in self.picks:
            return self.picks[key]
        times = {}
        for c in candidates:
            run(c)
            times[int(c)] = min(self.timer(lambda: run(c)) for _ in range(self.reps))
        best = min(times, key=times.get)
        self.picks[key] = best
        self.trials[key] = times
        if self.path:
            os.makedirs(os.path.dirname(os.path.abspath(self.path)), exist_ok=True)
            with open(self.path, "w") as f:
                json.dump(self.picks, f, indent=1, sort_keys=True)
        return best


def _device_time(fn) -> float:
    """Seconds of device time for fn() on the current stream (HIP events)."""
    a = torch.cuda.Event(enable_timing=True)
    b = torch.cuda.Event(enable_timing=True)
    a.record()
    fn()
    b.record()
    b.synchronize()
    return a.elapsed_time(b) / 1e3


@dataclass
class _Group:
    n: int
    pix: torch.Tensor
    count: int
    pairs: torch.Tensor
    blocks: int = 0
    chunk: int = 0
    staged: bool = True
    coh_waves: int = 0  # 0: the drop-in's tile layout


class DSBatchRunner:
    """Runs one DSBatch through the batched path; buffers are bound at
    construction, so the launches can be captured in a HIP graph and replayed
    after new data has been written into ``batch.samples`` in place."""

    def __init__(self, batch: DSBatch, ref: int = 0, tuner: AutoTuner | None = None,
                 max_workspace_bytes: int = 8 << 30, per_stage: bool = False):
        self.batch = batch
        self.ref = int(ref)
        # per_stage: coherence, EMI and temporal coherence as separate launches
        # through HBM for every group (the drop-in sequence), instead of the
        # fused launches for N <= 200
        self.per_stage = bool(per_stage)
        self.tuner = tuner
        dev = batch.samples.device
        _, ph_off = DSBatch.layout(batch.n_images, np.zeros(batch.n_px))
        self.ph_off = torch.from_numpy(ph_off).to(dev)
        total = int(np.asarray(batch.n_images, np.int64).sum())
        self.ph = torch.empty(total, dtype=torch.complex64, device=dev)
        self.quality = torch.empty(batch.n_px, dtype=torch.float32, device=dev)
        self.t_coh = torch.empty(batch.n_px, dtype=torch.float32, device=dev)
        self.groups: list[_Group] = []
        for n in sorted(set(np.asarray(batch.n_images).tolist())):
            if n < 2 or n > 1024:
                raise ValueError(f"ds_batch: N={n} outside [2, 1024]")
            ids = np.nonzero(batch.n_images == n)[0].astype(np.int32)
            pairs = _dev.to_dev(_pairs(None, n))
            self.groups.append(_Group(n, torch.from_numpy(ids).to(dev), ids.size, pairs))
        self.max_ws = int(max_workspace_bytes)
        self.ws = None
        self.graph = None
        self._tuned = False

    # ---------------------------------------------------------------- launches
    def _call(self, g: _Group, blocks: int, chunk: int, coh_waves: int | None = None):
        ws = self.ws
        _lib.call("mr_ds_batch", _dev.ptr(self.batch.samples), _dev.ptr(self.batch.row_off),
                  _dev.ptr(self.batch.n_samples), _dev.ptr(self.ph_off), _dev.ptr(g.pix), g.count, g.n, self.ref,
                  _dev.ptr(g.pairs), g.pairs.shape[0], _dev.ptr(self.ph), _dev.ptr(self.quality),
                  _dev.ptr(self.t_coh), int(blocks), int(chunk), int(g.coh_waves if coh_waves is None else coh_waves),
                  _dev.ptr(ws), 0 if ws is None else ws.numel(), _dev.stream())

    FUSED_MAX_N = 200  # largest N with a fused launch

    def _staged(self, g: _Group) -> bool:
        return self.per_stage or g.n > self.FUSED_MAX_N

    def _chunk_fits(self, n: int, chunk: int) -> bool:
        return _lib.query("mr_ds_batch_workspace_bytes", chunk, n) <= self.max_ws

    def tune(self):
        """Choose each group's launch shape (autotuner if given, else defaults)
        and allocate the shared workspace.  For N <= 200 the tuner also
        chooses between the fused launch (one of fused_blocks(N) grid sizes)
        and the staged one (coherence chunk through HBM, then EMI and temporal
        coherence; one of CHUNKS), encoded as -chunk: both give the same
        outputs bit for bit, and which is faster depends on N and the SHP
        counts.  For staged groups with 64 < N <= 224 it then chooses the tile
        layout of the coherence kernel (waves per workgroup over its 32 x 32
        tile grid, coh_waves), which does not change the outputs either."""
        dev = self.batch.samples.device
        need = 0
        for g in self.groups:
            g.staged = self._staged(g)
            if not g.staged:  # the fused launch's lists or slots, for every grid the tuner may try
                blocks = AutoTuner.fused_blocks(g.n) if self.tuner is not None else (0,)
                need = max([need] + [_lib.query("mr_ds_batch_fused_workspace_bytes", g.count, g.n, b)
                                     for b in blocks])
            if g.staged or self.tuner is not None:
                cands = [c for c in AutoTuner.CHUNKS if self._chunk_fits(g.n, c)] or [1024]
                g.chunk = min(cands[-1], max(g.count, 1))
                need = max(need, _lib.query("mr_ds_batch_workspace_bytes", g.chunk, g.n))
        self.ws = _dev.workspace(need, dev) if need else None
        if self.tuner is not None:
            for g in self.groups:
                if g.count == 0:
                    continue
                chunks = [c for c in AutoTuner.CHUNKS if self._chunk_fits(g.n, c) and c <= g.chunk] or [g.chunk]
                if not g.staged:
                    cands = list(AutoTuner.fused_blocks(g.n)) + [-c for c in chunks]
                    pick = self.tuner.choose(
                        AutoTuner.key("fused_or_staged", g.n, g.count), cands,
                        lambda v, g=g: self._call(g, v, 0) if v >= 0 else self._call(g, -1, -v))
                    if pick >= 0:
                        g.blocks = pick
                    else:
                        g.staged, g.chunk = True, -pick
                else:
                    g.chunk = self.tuner.choose(AutoTuner.key("chunk", g.n, g.count), chunks,
                                                lambda c, g=g: self._call(g, -1, c))
                waves = AutoTuner.coh_waves(g.n) if g.staged else ()
                if waves:  # the coherence kernel's tile layout, at the chosen chunk
                    g.coh_waves = self.tuner.choose(AutoTuner.key("coh_waves", g.n, g.count), waves,
                                                    lambda v, g=g: self._call(g, -1, g.chunk, v))
        self._tuned = True

    def launch(self):
        """Enqueue every group on the current stream."""
        if not self._tuned:
            self.tune()
        for g in self.groups:
            if g.count:
                self._call(g, -1 if g.staged else g.blocks, g.chunk)

    def run(self):
        """Eager run; returns (ph, ph_off, quality, t_coh)."""
        self.launch()
        return self.ph, self.ph_off, self.quality, self.t_coh

    def capture(self):
        """Record all group launches in one HIP graph (tuning and workspace
        allocation happen before capture, so the graph allocates nothing)."""
        if not self._tuned:
            self.tune()
        s = torch.cuda.Stream()
        s.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(s):
            self.launch()  # warm-up outside the graph
        torch.cuda.current_stream().wait_stream(s)
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            self.launch()
        self.graph = g
        return g

    def replay(self):
        if self.graph is None:
            self.capture()
        self.graph.replay()
        return self.ph, self.ph_off, self.quality, self.t_coh


def ds_batch(mats_or_batch, ref: int = 0, tuner: AutoTuner | None = None):
    """Batched coherence -> EMI -> temporal coherence.

    ``mats_or_batch``: a DSBatch or a list of (K_i, N_i) complex sample matrices.
    Returns (ph list of (N_i,) complex64, quality (n_px,), t_coh (n_px,)) in the
    input's kind (numpy for a list of numpy arrays, torch otherwise).
    """
    if isinstance(mats_or_batch, DSBatch):
        batch, as_torch = mats_or_batch, True
    else:
        mats = list(mats_or_batch)
        as_torch = bool(mats) and _dev.is_torch(mats[0])
        batch = DSBatch.from_matrices(mats)
    r = DSBatchRunner(batch, ref=ref, tuner=tuner)
    ph, ph_off, q, t = r.run()
    offs = r.ph_off.cpu().numpy()
    nimg = np.asarray(batch.n_images, np.int64)
    phs = [ph[o:o + k] for o, k in zip(offs.tolist(), nimg.tolist())]
    if not as_torch:
        phn = ph.cpu().numpy()
        phs = [phn[o:o + k] for o, k in zip(offs.tolist(), nimg.tolist())]
        return phs, q.cpu().numpy(), t.cpu().numpy()
    return phs, q, t
