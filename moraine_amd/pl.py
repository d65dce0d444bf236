"""Phase linking on the GPU: EMI, DS temporal coherence and the fused batch driver.

Drop-in for moraine.emi (moraine/pl.py:68-77), moraine.ds_temp_coh
(moraine/pl.py:126-145) and moraine.emperical_co_emi_temp_coh_pc
(moraine/pl.py:148-168).  Kernels: csrc/emi.hip, csrc/tcoh.hip.
"""
from __future__ import annotations

import numpy as np
import torch

from . import _dev, _lib
from .co import _full_pairs_dev, _pairs, emperical_co_pc
from .tnet import nimage_from_npair


#: after a call with ``refine`` set: a 0-d int32 device tensor, the number of
#: points that were solved again in FP64 (reading it synchronises the stream)
last_refined = None
#: ... and an int32 device tensor whose first ``last_refined`` entries are those
#: points' indices (in no particular order)
last_refined_points = None
#: after a plain EMI call with 32 < N <= 64: a 0-d int32 device tensor, the
#: number of points the shift-invert pass left to the Householder pass
last_householder = None
#: ... and a 0-d int32 device tensor, the number of those the Householder
#: pass's conditioning test (phase tolerance 1e-3 rad) sent to the FP64 re-solve
last_resolved = None


def _check_refine(refine, n: int):
    if refine is None:
        return None
    tol = float(refine)
    if not tol > 0.0:
        raise ValueError(f"refine must be a positive phase tolerance in radians, got {refine}")
    if n > 200:
        return None  # EMI runs in FP64 throughout above 200 images
    return tol


def _emi_refined_dev(c: torch.Tensor, n: int, ref: int, tol: float, ph: torch.Tensor, q: torch.Tensor, t=None):
    """mr_emi_refined: the one-wave solve, then an FP64 re-solve of every point
    whose estimated phase error exceeds tol (with t: the fused temporal coherence)."""
    global last_refined
    npts = c.shape[0]
    pairs = _full_pairs_dev(n, c.device) if t is not None else None
    wsb = _lib.query("mr_emi_refined_workspace_bytes", npts, n)
    ws = _dev.workspace(wsb, c.device)
    _lib.call("mr_emi_refined", _dev.ptr(c), npts, n, int(ref), tol, _dev.ptr(pairs) if pairs is not None else None,
              pairs.shape[0] if pairs is not None else 0, _dev.ptr(ph), _dev.ptr(q),
              _dev.ptr(t) if t is not None else None, _dev.ptr(ws), ws.numel(), _dev.stream())
    global last_refined_points
    last_refined = ws[:4].view(torch.int32)[0]
    last_refined_points = ws[256:256 + 4 * npts].view(torch.int32)


def _emi_dev(c: torch.Tensor, n: int, ref: int, tol=None):
    npts = c.shape[0]
    ph = torch.empty((npts, n), dtype=c.dtype, device=c.device)
    q = torch.empty(npts, dtype=torch.float32, device=c.device)
    if _dev.is_c128(c):  # FP64 throughout: refine has nothing to do
        ws = _dev.workspace(_lib.query("mr_emi_c128_workspace_bytes", npts, n), c.device)
        _lib.call("mr_emi_c128", _dev.ptr(c), npts, n, int(ref), _dev.ptr(ph), _dev.ptr(q), _dev.ptr(ws), ws.numel(),
                  _dev.stream())
        return ph, q
    if tol is not None:
        _emi_refined_dev(c, n, ref, tol, ph, q)
        return ph, q
    wsb = _lib.query("mr_emi_workspace_bytes", npts, n)
    ws = _dev.workspace(wsb, c.device)
    _lib.call("mr_emi", _dev.ptr(c), npts, n, int(ref), _dev.ptr(ph), _dev.ptr(q), _dev.ptr(ws), ws.numel(),
              _dev.stream())
    _note_householder(ws, n)
    return ph, q


def _note_householder(ws: torch.Tensor, n: int):
    global last_householder, last_resolved
    if 32 < n <= 64:
        last_householder = ws[:4].view(torch.int32)[0]
        last_resolved = ws[4:8].view(torch.int32)[0]


def emi(coh, ref: int = 0, *, refine=None):
    """EMI phase history (n_pts, N) with coh's dtype and quality (n_pts,) float32
    from compressed coherence (n_pts, N(N-1)/2) in triu order (moraine/pl.py:45-65).
    complex128 coherence runs every stage in FP64.

    refine (not in moraine): a phase tolerance in radians, e.g. 1e-3.  Points
    whose phase error from the FP32 stages may exceed it are solved again in
    FP64 throughout; the others keep the plain call's outputs.  The test is an
    FP64 eigen-residual bound for the shift-invert points (32 < N <= 64) and a
    conditioning estimate from the tridiagonal form for the Householder ones.
    (Above 200 images EMI is FP64 throughout and refine has nothing to do.)"""
    as_torch = _dev.is_torch(coh)
    c = _dev.to_dev_complex(coh)
    n = nimage_from_npair(c.shape[-1])
    tol = _check_refine(refine, n)
    lead = tuple(c.shape[:-1])
    ph, q = _emi_dev(c.reshape(-1, c.shape[-1]), n, ref, tol)
    return _dev.back(ph.reshape(lead + (n,)), as_torch), _dev.back(q.reshape(lead), as_torch)


def _tcoh(kind: str, a, b, image_pairs):
    as_torch = _dev.is_torch(a)
    x = _dev.to_dev_complex(a)
    y = _dev.to_dev_complex(b)
    if x.dtype != y.dtype:  # complex128 if either is: the FP64 kernel
        x, y = x.to(torch.complex128), y.to(torch.complex128)
    if _dev.is_c128(x):
        kind = kind + "_c128"
    n = y.shape[-1]
    lead = tuple(y.shape[:-1])
    x2 = x.reshape(-1, x.shape[-1])
    y2 = y.reshape(-1, n)
    pairs_np = _pairs(image_pairs, n)
    if x2.shape[-1] != pairs_np.shape[0]:
        raise ValueError(f"{kind}: last dim {x2.shape[-1]} != number of image pairs {pairs_np.shape[0]}")
    pairs = _dev.to_dev(pairs_np)
    out = torch.empty(y2.shape[0], dtype=torch.float32, device=y.device)
    _lib.call(kind, _dev.ptr(x2), _dev.ptr(y2), y2.shape[0], n, _dev.ptr(pairs), pairs_np.shape[0], _dev.ptr(out),
              _dev.stream())
    return _dev.back(out.reshape(lead), as_torch)


def ds_temp_coh(coh, ph, image_pairs=None, block_size: int = 128):
    """|sum_p coh_p/|coh_p| conj(ph_a) ph_b| / n_pairs per point (float32)."""
    return _tcoh("mr_ds_temp_coh", coh, ph, image_pairs)


def _emi_tcoh_dev(c: torch.Tensor, n: int, ref: int, ph: torch.Tensor, q: torch.Tensor, t: torch.Tensor, tol=None):
    """emi then ds_temp_coh on the same full-network rows, into ph, q, t (one
    kernel for 32 < n <= 64; results equal the two calls bit for bit)."""
    if _dev.is_c128(c):
        p2, q2 = _emi_dev(c, n, ref)
        ph.copy_(p2)
        q.copy_(q2)
        pairs = _full_pairs_dev(n, c.device)
        _lib.call("mr_ds_temp_coh_c128", _dev.ptr(c), _dev.ptr(ph), c.shape[0], n, _dev.ptr(pairs), pairs.shape[0],
                  _dev.ptr(t), _dev.stream())
        return
    if tol is not None:
        _emi_refined_dev(c, n, ref, tol, ph, q, t)
        return
    npts = c.shape[0]
    pairs = _full_pairs_dev(n, c.device)
    wsb = _lib.query("mr_emi_workspace_bytes", npts, n)
    ws = _dev.workspace(wsb, c.device)
    _lib.call("mr_emi_ds_temp_coh", _dev.ptr(c), npts, n, int(ref), _dev.ptr(pairs), pairs.shape[0], _dev.ptr(ph),
              _dev.ptr(q), _dev.ptr(t), _dev.ptr(ws), ws.numel(), _dev.stream())
    _note_householder(ws, n)


def emi_ds_temp_coh(coh, ref: int = 0, *, refine=None):
    """(ph, quality, t_coh) = emi(coh, ref, refine) followed by ds_temp_coh(coh, ph):
    the inner stages of moraine.emperical_co_emi_temp_coh_pc (moraine/pl.py:165-167)
    for full-network coherence rows (n_pts, N(N-1)/2) in triu order."""
    as_torch = _dev.is_torch(coh)
    c = _dev.to_dev_complex(coh).reshape(-1, coh.shape[-1])
    n = nimage_from_npair(c.shape[-1])
    tol = _check_refine(refine, n)
    ph = torch.empty((c.shape[0], n), dtype=c.dtype, device=c.device)
    q = torch.empty(c.shape[0], dtype=torch.float32, device=c.device)
    t = torch.empty(c.shape[0], dtype=torch.float32, device=c.device)
    _emi_tcoh_dev(c, n, ref, ph, q, t, tol)
    lead = tuple(coh.shape[:-1])
    return (_dev.back(ph.reshape(lead + (n,)), as_torch), _dev.back(q.reshape(lead), as_torch),
            _dev.back(t.reshape(lead), as_torch))


#: a batch of the batch driver holds at least this many points (fewer only
#: lengthen the run: every point is independent) ...
MIN_BATCH_POINTS = 1 << 21
#: ... and its coherence buffer takes at most this share of the free HBM
BATCH_HBM_SHARE = 0.25


def batch_points(batch_size: int, n: int, n_pc: int, free_bytes: int) -> int:
    """Points per batch of emperical_co_emi_temp_coh_pc: the caller's batch_size
    raised to MIN_BATCH_POINTS, then capped so that the batch's coherence rows
    (8 N(N-1)/2 bytes per point, 16 for complex128 adds at most 2x: the cap uses
    16) stay within BATCH_HBM_SHARE of free_bytes; at least 1 point."""
    per_point = 16 * (n * (n - 1) // 2) + 16 * n
    cap = max(1, int(BATCH_HBM_SHARE * free_bytes) // per_point)
    return max(1, min(max(int(batch_size), MIN_BATCH_POINTS), cap, max(n_pc, 1)))


def emperical_co_emi_temp_coh_pc(rslc, idx, pc_is_shp, batch_size: int = 1000, *, refine=None):
    """Coherence -> EMI -> temporal coherence for a point cloud, in batches.

    Every point is independent and every kernel deterministic, so the result
    does not depend on the batch size.  The batch is batch_points(batch_size,
    ...): at least 2^21 points (smaller batches run slower: at N = 60, 15x21
    windows, 65536-point batches take 1.2x as long as 2^18-point ones, 2^20-point
    ones 0.97x, 2^21-point ones 0.95x), at most what a quarter of the free HBM
    holds.
    refine: as for emi (not in moraine).  complex128 input runs in FP64 and
    returns complex128 phases.
    """
    as_torch = _dev.is_torch(rslc)
    z = _dev.to_dev_complex(rslc)
    ix = _dev.to_dev(idx, torch.int32)
    m = _dev.to_dev(pc_is_shp, torch.bool)
    n = z.shape[2]
    tol = _check_refine(refine, n)
    n_pc = ix.shape[0]
    ph = torch.empty((n_pc, n), dtype=z.dtype, device=z.device)
    q = torch.empty(n_pc, dtype=torch.float32, device=z.device)
    t = torch.empty(n_pc, dtype=torch.float32, device=z.device)
    bs = batch_points(batch_size, n, n_pc, torch.cuda.mem_get_info(z.device)[0])
    if tol is None and not _dev.is_c128(z):
        co_emi_tcoh_batches(z, ix, m, 0, ph, q, t, bs)
    else:
        for s in range(0, n_pc, bs):
            e = min(n_pc, s + bs)
            c = emperical_co_pc(z, ix[s:e], m[s:e])
            _emi_tcoh_dev(c, n, 0, ph[s:e], q[s:e], t[s:e], tol)
    return _dev.back(ph, as_torch), _dev.back(q, as_torch), _dev.back(t, as_torch)


_SIDE_STREAMS: dict = {}


def _side_streams(device, count: int = 2) -> tuple:
    """`count` streams of default priority per device, created once and reused
    (so their cached workspace blocks are too)."""
    key = (str(device), count)
    if key not in _SIDE_STREAMS:
        _SIDE_STREAMS[key] = tuple(torch.cuda.Stream(device=device) for _ in range(count))
    return _SIDE_STREAMS[key]


def co_emi_tcoh_batches(z: torch.Tensor, ix: torch.Tensor, m: torch.Tensor, ref: int, ph: torch.Tensor,
                        q: torch.Tensor, t: torch.Tensor, bs: int, streams: int = 2):
    """co_emi_tcoh_dev over batches of bs points.  Batches alternate over
    `streams` side streams (each waits for the caller's stream first; the
    caller's stream waits for all of them at the end), so one batch's
    Householder pass and FP64 re-solve -- a small tail of low occupancy --
    overlap the next batch's fused launch; streams=1 runs every batch on the
    caller's stream, one after the other.  Two is the default: three or four
    measured the same at N = 60 with 2^19- and 2^20-point batches.  Every
    point is independent: the outputs do not depend on bs or on the overlap."""
    n_pc = ix.shape[0]
    if streams == 1:
        for s in range(0, n_pc, bs):
            e = min(n_pc, s + bs)
            co_emi_tcoh_dev(z, ix[s:e], m[s:e], ref, ph[s:e], q[s:e], t[s:e])
        return
    if streams < 1:
        raise ValueError(f"streams must be at least 1, got {streams}")
    cur = torch.cuda.current_stream(z.device)
    sides = _side_streams(z.device, streams)
    for st in sides:
        st.wait_stream(cur)
    for b, s in enumerate(range(0, n_pc, bs)):
        e = min(n_pc, s + bs)
        with torch.cuda.stream(sides[b % streams]):
            co_emi_tcoh_dev(z, ix[s:e], m[s:e], ref, ph[s:e], q[s:e], t[s:e])
    for st in sides:
        cur.wait_stream(st)


def co_emi_tcoh_dev(z: torch.Tensor, ix: torch.Tensor, m: torch.Tensor, ref: int, ph: torch.Tensor, q: torch.Tensor,
                    t: torch.Tensor):
    """One batch of emperical_co_emi_temp_coh_pc for complex64 device inputs
    (z (nl, w, N) complex64, ix (n_pc, 2) int32, m (n_pc, aw, rw) bool, all
    contiguous) into ph, q, t: coherence, EMI and DS temporal coherence in one
    launch for 32 < N <= 64 (mr_co_emi_ds_temp_coh_pc); equal bit for bit to
    emperical_co_pc followed by emi and ds_temp_coh."""
    if z.dtype != torch.complex64 or ix.dtype != torch.int32 or m.dtype not in (torch.bool, torch.uint8):
        raise TypeError("co_emi_tcoh_dev: complex64 stack, int32 points and a bool/uint8 mask expected")
    if not (z.is_contiguous() and ix.is_contiguous() and m.is_contiguous()):
        raise ValueError("co_emi_tcoh_dev: inputs must be contiguous")
    nl, w, n = z.shape
    aw, rw = m.shape[-2:]
    n_pc = ix.shape[0]
    if (tuple(ph.shape) != (n_pc, n) or tuple(q.shape) != (n_pc,) or tuple(t.shape) != (n_pc,)
            or ph.dtype != torch.complex64 or q.dtype != torch.float32 or t.dtype != torch.float32
            or not (ph.is_contiguous() and q.is_contiguous() and t.is_contiguous())):
        raise ValueError("co_emi_tcoh_dev: outputs must be contiguous ph (n_pc, N) complex64, q and t (n_pc,) "
                         "float32")
    ws = _dev.workspace(_lib.query("mr_co_emi_ds_temp_coh_pc_workspace_bytes", n_pc, n), z.device)
    _lib.call("mr_co_emi_ds_temp_coh_pc", _dev.ptr(z), nl, w, n, _dev.ptr(ix), _dev.ptr(m), n_pc, aw, rw, int(ref),
              _dev.ptr(ph), _dev.ptr(q), _dev.ptr(t), _dev.ptr(ws), ws.numel(), _dev.stream())
    _note_householder(ws, n)
