"""Covariance / coherence matrix estimation on the GPU.

Drop-in for moraine.emperical_co (moraine/co.py:117-137),
moraine.emperical_co_pc (moraine/co.py:230-266) and moraine.uncompress_coh
(moraine/co.py:281-299).  Kernels: csrc/co.hip.
"""
from __future__ import annotations

import numpy as np
import torch

from . import _dev, _lib
from .tnet import TempNet, nimage_from_npair


def emperical_co(rslc, is_shp, block_size: int = 128):
    """Full-raster (cov, coh), each (nlines, width, nimages, nimages) complex64.

    cov = sum/n_shp, coh = sum/sqrt(P_m P_j) over the in-bounds SHPs of each
    pixel's window; half windows are (az_win-1)//2, (r_win-1)//2.  The outputs
    have rslc's dtype (moraine/co.py:132-133): complex128 input runs in FP64.
    """
    as_torch = _dev.is_torch(rslc)
    z = _dev.to_dev_complex(rslc)
    m = _dev.to_dev(is_shp, torch.bool)
    nl, w, n = z.shape
    aw, rw = m.shape[-2:]
    if tuple(m.shape[:2]) != (nl, w):
        raise ValueError("emperical_co: is_shp must be (nlines, width, az_win, r_win)")
    cov = torch.empty((nl, w, n, n), dtype=z.dtype, device=z.device)
    coh = torch.empty_like(cov)
    fn = "mr_emperical_co_c128" if _dev.is_c128(z) else "mr_emperical_co"
    _lib.call(fn, _dev.ptr(z), _dev.ptr(m), nl, w, n, aw, rw, 0, nl, _dev.ptr(cov), _dev.ptr(coh), _dev.stream())
    return _dev.back(cov, as_torch), _dev.back(coh, as_torch)


def _pairs(image_pairs, n):
    """(P, 2) int32 pair array; indices follow NumPy's rules for an axis of n
    images: negative ones count from the end, others outside [-n, n) raise."""
    if image_pairs is None:
        image_pairs = TempNet.from_bandwidth(n).image_pairs
    if _dev.is_torch(image_pairs):
        image_pairs = image_pairs.cpu().numpy()
    pr = np.asarray(image_pairs)
    if pr.size and (pr.ndim != 2 or pr.shape[1] != 2):
        raise ValueError(f"image_pairs must be (n_pairs, 2), got shape {pr.shape}")
    pr = pr.reshape(-1, 2).astype(np.int64)
    if np.any(pr < -n) or np.any(pr >= n):
        raise IndexError(f"image pair index out of range for {n} images")
    return np.ascontiguousarray(np.where(pr < 0, pr + n, pr).astype(np.int32))


_FULL_PAIRS: dict = {}


def _full_pairs_dev(n: int, device) -> torch.Tensor:
    """The full pair network of n images in triu order, uploaded once per (n, device):
    an upload from pageable memory per call would wait for the stream and stop the
    host from queueing ahead of the GPU."""
    key = (n, str(device))
    if key not in _FULL_PAIRS:
        _FULL_PAIRS[key] = _dev.to_dev(_pairs(None, n)).to(device)
    return _FULL_PAIRS[key]


def emperical_co_pc(rslc, idx, pc_is_shp, block_size: int = 128, image_pairs=None):
    """Coherence of the selected image pairs for a point cloud: (n_pc, n_pairs),
    rslc's dtype (complex128 input runs in FP64)."""
    as_torch = _dev.is_torch(rslc)
    z = _dev.to_dev_complex(rslc)
    nl, w, n = z.shape
    ix = _dev.to_dev(idx, torch.int32)
    m = _dev.to_dev(pc_is_shp, torch.bool)
    aw, rw = m.shape[-2:]
    pairs = _full_pairs_dev(n, z.device) if image_pairs is None else _dev.to_dev(_pairs(image_pairs, n))
    n_pc = ix.shape[0]
    coh = torch.empty((n_pc, pairs.shape[0]), dtype=z.dtype, device=z.device)
    fn = "mr_emperical_co_pc_c128" if _dev.is_c128(z) else "mr_emperical_co_pc"
    _lib.call(fn, _dev.ptr(z), nl, w, n, _dev.ptr(ix), _dev.ptr(m), n_pc, aw, rw, _dev.ptr(pairs),
              pairs.shape[0], _dev.ptr(coh), _dev.stream())
    return _dev.back(coh, as_torch)


def ad_intf_pc(ref_rslc, sec_rslc, idx, pc_is_shp, block_size: int = 128):
    """Adaptively multilooked interferogram of one image pair per point,
    (n_pc,) complex64, or complex128 for complex128 input (moraine/co.py:376-406);
    ``block_size`` is unused."""
    as_torch = _dev.is_torch(ref_rslc)
    a = _dev.to_dev_complex(ref_rslc)
    b = _dev.to_dev_complex(sec_rslc)
    if a.dtype != b.dtype:
        a, b = a.to(torch.complex128), b.to(torch.complex128)
    if a.ndim != 2 or a.shape != b.shape:
        raise ValueError("ad_intf_pc: ref_rslc and sec_rslc must be (nlines, width) of one shape")
    nl, w = a.shape
    ix = _dev.to_dev(idx, torch.int32).contiguous()
    m = _dev.to_dev(pc_is_shp, torch.bool).contiguous()
    aw, rw = m.shape[-2:]
    n_pc = ix.shape[0]
    out = torch.empty(n_pc, dtype=a.dtype, device=a.device)
    _lib.call("mr_ad_intf_pc_c128" if _dev.is_c128(a) else "mr_ad_intf_pc", _dev.ptr(a), _dev.ptr(b), nl, w, _dev.ptr(ix), _dev.ptr(m), n_pc, aw, rw,
              _dev.ptr(out), _dev.stream())
    return _dev.back(out, as_torch)


def regularize_spectral(coh, beta):
    """(1 - beta) coh + beta I for a stack (..., N, N); beta a number or an array
    of the stack's leading shape.  FP64 arithmetic; the output has coh's
    precision (complex64, or complex128 for complex128 input)."""
    as_torch = _dev.is_torch(coh)
    c = _dev.to_dev_complex(coh)
    n = c.shape[-1]
    if c.ndim < 2 or c.shape[-2] != n:
        raise ValueError("regularize_spectral: coh must be (..., N, N)")
    lead = tuple(c.shape[:-2])
    nmat = int(np.prod(lead)) if lead else 1
    b = beta.detach().to(torch.float64) if torch.is_tensor(beta) else torch.as_tensor(np.asarray(beta, np.float64))
    if b.numel() == 1:
        b, per = b.reshape(1), 0
    else:
        b, per = torch.broadcast_to(b.reshape(b.shape), lead).reshape(-1), 1
    b = b.to(c.device).contiguous()
    out = torch.empty_like(c)
    _lib.call("mr_regularize_spectral_c128" if _dev.is_c128(c) else "mr_regularize_spectral", _dev.ptr(c), nmat, n, _dev.ptr(b), per, _dev.ptr(out), _dev.stream())
    return _dev.back(out, as_torch)


def _pd_call(co, nearest: bool):
    as_torch = _dev.is_torch(co)
    a = _dev.to_dev(co)
    if a.is_complex():
        raise TypeError("isPD/nearestPD: real matrices expected (the reference passes |coherence|)")
    if a.dtype != torch.float64:
        a = a.to(torch.float32)
    if a.ndim < 2 or a.shape[-1] != a.shape[-2]:
        raise ValueError("isPD/nearestPD: co must be a stack of square matrices (..., N, N)")
    a = a.contiguous()
    n = a.shape[-1]
    lead = tuple(a.shape[:-2])
    nmat = int(np.prod(lead)) if lead else 1
    ws = _dev.workspace(_lib.query("mr_pd_workspace_bytes", nmat, n), a.device)
    sfx = "f64" if a.dtype == torch.float64 else "f32"
    if nearest:
        out = torch.empty_like(a)
        _lib.call(f"mr_nearest_pd_{sfx}", _dev.ptr(a), nmat, n, _dev.ptr(out), _dev.ptr(ws), ws.numel(), _dev.stream())
    else:
        out = torch.empty(lead, dtype=torch.bool, device=a.device)
        _lib.call(f"mr_is_pd_{sfx}", _dev.ptr(a), nmat, n, _dev.ptr(out), _dev.ptr(ws), ws.numel(), _dev.stream())
    return _dev.back(out, as_torch)


def isPD(co):
    """Whether each matrix of a real stack (..., N, N) is positive definite (its
    Cholesky factorisation succeeds, FP64): bool (...) (moraine/co.py:409-414)."""
    return _pd_call(co, False)


def nearestPD(co):
    """Nearest positive-definite matrix of each matrix of a real stack (..., N, N)
    by Higham's method as moraine/co.py:429-463 writes it (its undefined
    ``wherePD`` read as isPD); FP64 arithmetic, the input's dtype out."""
    return _pd_call(co, True)


def uncompress_coh(coh, image_pairs=None):
    """(..., n_pairs) compressed -> (..., n, n) Hermitian with unit diagonal, coh's dtype."""
    as_torch = _dev.is_torch(coh)
    c = _dev.to_dev_complex(coh)
    if image_pairs is None:
        n = nimage_from_npair(c.shape[-1])
        pairs_np = TempNet.from_bandwidth(n).image_pairs
    else:
        pairs_np = np.asarray(image_pairs.cpu() if _dev.is_torch(image_pairs) else image_pairs, np.int32)
        n = int(pairs_np[-1, -1]) + 1
        # NumPy's indexing rules: negative indices count from the end, others out of range raise
        if np.any(pairs_np < -n) or np.any(pairs_np >= n):
            raise IndexError(f"image pair index out of range for {n} images")
        pairs_np = np.where(pairs_np < 0, pairs_np + n, pairs_np).astype(np.int32)
    P = pairs_np.shape[0]
    if c.shape[-1] != P:
        if c.shape[-1] != 1:  # NumPy broadcasting: a last axis of 1 serves every pair
            raise ValueError(f"coh has {c.shape[-1]} pairs, image_pairs {P}")
        c = c.expand(tuple(c.shape[:-1]) + (P,))
    c = c.contiguous()
    lead = tuple(c.shape[:-1])
    npts = int(np.prod(lead)) if lead else 1
    out = torch.empty(lead + (n, n), dtype=c.dtype, device=c.device)
    pairs = _dev.to_dev(np.ascontiguousarray(pairs_np))
    _lib.call("mr_uncompress_coh_c128" if _dev.is_c128(c) else "mr_uncompress_coh", _dev.ptr(c), npts, n, _dev.ptr(pairs), pairs_np.shape[0], _dev.ptr(out),
              _dev.stream())
    return _dev.back(out, as_torch)
