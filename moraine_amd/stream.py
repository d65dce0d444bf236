"""Chunk streaming of a host-resident stack through one GPU.

Replaces moraine's Dask chunking of the DS pipeline (moraine/cli/shp.py,
moraine/cli/co.py, moraine/cli/pl.py: ``map_blocks(ks_test)`` with a window
halo, then ``emperical_co_emi_temp_coh_pc`` per chunk) for a stack that stays in
host memory: the raster is cut into row bands; each band is copied with its
KS half-window of halo rows above and below through pinned buffers on a copy
stream while the previous band computes on the compute stream, and its results
come back through pinned buffers, one band behind.  Per band the chain is the
drop-in one (ks_test_select_shp, then emperical_co_emi_temp_coh_pc's batches:
coherence -> emi -> ds_temp_coh), so
every pixel's outputs equal the whole-tile computation bit for bit: the halo
gives each interior pixel its full window, and the clipped bands at the raster's
first and last rows meet the raster's own edges.
"""
from __future__ import annotations

import numpy as np
import torch

from .pl import co_emi_tcoh_batches
from .shp import ks_test_select_shp

__all__ = ["ds_pipeline_streamed"]


def ds_pipeline_streamed(rslc, az_half_win: int, r_half_win: int, p_max: float = 0.05, band_rows: int = 256,
                         batch: int = 1 << 20, device="cuda"):
    """DS phase linking of every pixel of a host (nlines, width, nimages)
    complex64 stack (numpy, memmap welcome), streamed through the GPU in bands.

    Returns numpy arrays (ph (nl, w, n) complex64, quality (nl, w) float32,
    t_coh (nl, w) float32, shp_num (nl, w) int32), equal to running
    ks_test_select_shp on the whole raster and emperical_co_emi_temp_coh_pc on
    all its pixels.  band_rows: interior rows per band; batch: points per
    coherence/EMI launch.
    """
    z = np.asarray(rslc) if not isinstance(rslc, np.memmap) else rslc
    if z.dtype != np.complex64 or z.ndim != 3:
        raise TypeError("ds_pipeline_streamed: rslc must be a (nlines, width, nimages) complex64 host array")
    nl, w, n = z.shape
    az, r = int(az_half_win), int(r_half_win)
    aw, rw = 2 * az + 1, 2 * r + 1
    B = max(1, min(int(band_rows), nl))
    dev = torch.device(device)
    halo_rows = B + 2 * az
    comp = torch.cuda.current_stream(dev)
    copy = torch.cuda.Stream(dev)  # default priority
    # two pinned input slots and two pinned output slots (double buffering)
    pin_in = [torch.empty((halo_rows, w, n), dtype=torch.complex64, pin_memory=True) for _ in range(2)]
    dev_in = [torch.empty((halo_rows, w, n), dtype=torch.complex64, device=dev) for _ in range(2)]
    pin_ph = [torch.empty((B, w, n), dtype=torch.complex64, pin_memory=True) for _ in range(2)]
    pin_qtc = [torch.empty((3, B, w), dtype=torch.float32, pin_memory=True) for _ in range(2)]
    ph = np.empty((nl, w, n), np.complex64)
    q = np.empty((nl, w), np.float32)
    t = np.empty((nl, w), np.float32)
    cnt = np.empty((nl, w), np.int32)
    bands = [(b0, min(nl, b0 + B)) for b0 in range(0, nl, B)]

    def rows_of(k):
        b0, b1 = bands[k]
        return max(0, b0 - az), min(nl, b1 + az)

    copy_ev = [None, None]  # last host -> device copy out of each pinned input slot

    def stage(k):  # host -> pinned (CPU copy), then pinned -> device on the copy stream
        h0, h1 = rows_of(k)
        s = k & 1
        if copy_ev[s] is not None:
            copy_ev[s].synchronize()  # the slot's previous copy has left the pinned buffer
        pin_in[s][: h1 - h0].copy_(torch.from_numpy(z[h0:h1]))  # torch's host copy runs on several threads
        with torch.cuda.stream(copy):
            # the slot's previous band must be done with dev_in before it is overwritten
            copy.wait_event(done_in[s])
            dev_in[s][: h1 - h0].copy_(pin_in[s][: h1 - h0], non_blocking=True)
            ev = torch.cuda.Event()
            ev.record(copy)
        copy_ev[s] = ev
        return ev

    done_in = [torch.cuda.Event(), torch.cuda.Event()]
    for e in done_in:
        e.record(comp)
    pending = [None, None]  # (k, event) of results waiting in the pinned output slots

    def drain(s):
        if pending[s] is None:
            return
        k, ev = pending[s]
        ev.synchronize()
        b0, b1 = bands[k]
        m = b1 - b0
        torch.from_numpy(ph[b0:b1]).copy_(pin_ph[s][:m])
        qt = pin_qtc[s][:, :m].numpy()
        q[b0:b1], t[b0:b1] = qt[0], qt[1]
        cnt[b0:b1] = qt[2].view(np.int32)
        pending[s] = None

    ready = stage(0)
    for k in range(len(bands)):
        s = k & 1
        comp.wait_event(ready)
        h0, h1 = rows_of(k)
        b0, b1 = bands[k]
        zb = dev_in[s][: h1 - h0]
        _, shp, c = ks_test_select_shp(zb, az, r, p_max, return_p=False)
        i0, m = b0 - h0, b1 - b0
        ii, jj = torch.meshgrid(torch.arange(i0, i0 + m, device=dev, dtype=torch.int32),
                                torch.arange(w, device=dev, dtype=torch.int32), indexing="ij")
        idx = torch.stack([ii.reshape(-1), jj.reshape(-1)], -1).contiguous()
        pcs = shp[i0:i0 + m].reshape(-1, aw, rw)
        npt = idx.shape[0]
        ph_b = torch.empty((npt, n), dtype=torch.complex64, device=dev)
        qtc = torch.empty((3, npt), dtype=torch.float32, device=dev)
        # emperical_co_emi_temp_coh_pc's batches (joined back to this stream before it returns)
        co_emi_tcoh_batches(zb, idx, pcs, 0, ph_b, qtc[0], qtc[1], batch)
        qtc[2].copy_(c[i0:i0 + m].reshape(-1).view(torch.float32))
        done_in[s].record(comp)  # this slot's input may be overwritten from here on
        # the next band's host copy and upload, and the host copy of band k-2's results,
        # run while the GPU computes band k (all of band k is enqueued above)
        nxt = stage(k + 1) if k + 1 < len(bands) else None
        drain(s)  # the pinned output slot s held band k-2
        pin_ph[s][:m].view(-1, n).copy_(ph_b, non_blocking=True)
        pin_qtc[s][:, :m].view(3, -1).copy_(qtc, non_blocking=True)
        ev = torch.cuda.Event()
        ev.record(comp)
        pending[s] = (k, ev)
        if nxt is not None:
            ready = nxt
    drain(0)
    drain(1)
    return ph, q, t, cnt
