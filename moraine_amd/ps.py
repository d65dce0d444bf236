"""Amplitude dispersion on the GPU.

Drop-in for moraine.amp_disp (moraine/ps.py:17-72), the PS-candidate measure
beside the DS path.  Kernel: csrc/ks.hip (amp_disp_kernel).
"""
from __future__ import annotations

import torch

from . import _dev, _lib


def amp_disp(rslc):
    """std(|rslc|) / mean(|rslc|) over the image axis: (nlines, width, nimages)
    complex64 -> (nlines, width) float32 (numpy or torch, same kind out)."""
    as_torch = _dev.is_torch(rslc)
    z = _dev.to_dev(rslc, torch.complex64).contiguous()
    if z.ndim != 3:
        raise ValueError("amp_disp: rslc must be (nlines, width, nimages)")
    nl, w, n = z.shape
    out = torch.empty((nl, w), dtype=torch.float32, device=z.device)
    _lib.call("mr_amp_disp", _dev.ptr(z), nl, w, n, _dev.ptr(out), _dev.stream())
    return _dev.back(out, as_torch)
