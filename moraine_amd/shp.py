"""SHP identification: KS test and SHP selection on the GPU.

Drop-in for moraine.ks_test (moraine/shp.py:269-305) and moraine.select_shp
(moraine/shp.py:308-323).  Both run in the HIP kernels of csrc/ks.hip; numpy
inputs are computed on the GPU and returned as numpy (the reference's
select_shp is CPU-only; here it runs on the GPU for both kinds of input).
"""
from __future__ import annotations

import numpy as np
import torch

from . import _dev, _lib

_KS = {torch.float32: "mr_ks_test_f32", torch.float64: "mr_ks_test_f64", torch.complex64: "mr_ks_test_rslc_f32"}


def _ks(rmli, az_half_win: int, r_half_win: int, want_p: bool, want_dist: bool, p_max, want_mask: bool):
    as_torch = _dev.is_torch(rmli)
    x = _dev.to_dev(rmli)
    if x.dtype not in _KS:
        raise TypeError(f"ks_test: rmli must be float32, float64 or complex64, got {x.dtype}")
    if x.ndim != 3:
        raise ValueError("ks_test: rmli must be (nlines, width, nimages)")
    x = x.contiguous()
    nl, w, n = x.shape
    aw, rw = 2 * az_half_win + 1, 2 * r_half_win + 1
    shape = (nl, w, aw, rw)
    rdt = torch.float32 if x.dtype == torch.complex64 else x.dtype
    p = torch.empty(shape, dtype=rdt, device=x.device) if want_p else None
    d = torch.empty(shape, dtype=rdt, device=x.device) if want_dist else None
    m = torch.empty(shape, dtype=torch.bool, device=x.device) if want_mask else None
    cnt = torch.empty((nl, w), dtype=torch.int32, device=x.device) if want_mask else None
    wsb = _lib.query("mr_ks_workspace_bytes", nl, w, n, az_half_win, r_half_win, 8 if x.dtype == torch.float64 else 4)
    ws = _dev.workspace(wsb, x.device)
    _lib.call(_KS[x.dtype], _dev.ptr(x), nl, w, n, az_half_win, r_half_win,
              float(p_max if p_max is not None else 0.0), _dev.ptr(p), _dev.ptr(d), _dev.ptr(m), _dev.ptr(cnt),
              _dev.ptr(ws), ws.numel(), _dev.stream())
    conv = (lambda t: _dev.back(t, as_torch) if t is not None else None)
    return conv(p), conv(d), conv(m), conv(cnt)


def ks_test(rmli, az_half_win: int, r_half_win: int, block_size: int = 128, return_dist: bool = False):
    """Two-sample KS test of every pixel against every pixel of its search window.

    rmli: (nlines, width, nimages) float32/float64 (numpy or torch).  Returns p,
    or (dist, p) with ``return_dist``; shape (nlines, width, 2*az+1, 2*r+1), dtype
    of rmli.  NaN where the secondary is outside the image or either stack has a
    NaN.  ``block_size`` is accepted for signature compatibility and unused.
    A complex64 stack (not accepted by moraine) is tested on its amplitudes:
    the result equals ``ks_test(rslc2amp(rslc), ...)`` (float32) bit for bit,
    without writing the amplitude array.
    """
    p, d, _, _ = _ks(rmli, az_half_win, r_half_win, True, return_dist, None, False)
    return (d, p) if return_dist else p


def ks_test_select_shp(rmli, az_half_win: int, r_half_win: int, p_max: float, return_p: bool = True):
    """Fused ks_test + select_shp in one pass: returns (p or None, is_shp, shp_num).

    is_shp and shp_num equal ``select_shp(ks_test(rmli, ...), p_max)`` bit for bit.
    """
    p, _, m, cnt = _ks(rmli, az_half_win, r_half_win, return_p, False, p_max, True)
    return p, m, cnt


def select_shp(p, p_max: float):
    """is_shp = p < p_max (NaN -> False) and shp_num = count per pixel (int32).
    float64 p is compared in float64 (other dtypes as float32)."""
    as_torch = _dev.is_torch(p)
    x = _dev.to_dev(p)
    if x.dtype != torch.float64:
        x = x.to(torch.float32)
    if x.ndim != 4:
        raise ValueError("select_shp: p must be 4-D (n_az, n_r, az_win, r_win)")
    npix = x.shape[0] * x.shape[1]
    win = x.shape[2] * x.shape[3]
    m = torch.empty(x.shape, dtype=torch.bool, device=x.device)
    cnt = torch.empty(x.shape[:2], dtype=torch.int32, device=x.device)
    _lib.call("mr_select_shp_f64" if x.dtype == torch.float64 else "mr_select_shp_f32", _dev.ptr(x), npix, win, float(p_max), _dev.ptr(m), _dev.ptr(cnt), _dev.stream())
    return _dev.back(m, as_torch), _dev.back(cnt, as_torch)
