"""SLC amplitude on the GPU.

Drop-in for moraine.rslc2amp (moraine/calamp.py:33-40), the step the SHP CLI
runs before ks_test (moraine/cli/shp.py:98).  Kernel: csrc/ks.hip
(rslc2amp_kernel); the amplitude rounds as the C library's hypotf, which the
reference's CPU path calls.
"""
from __future__ import annotations

import torch

from . import _dev, _lib


def rslc2amp(rslc):
    """|rslc| as float32, same shape (numpy or torch complex64 in, same kind out)."""
    as_torch = _dev.is_torch(rslc)
    z = _dev.to_dev(rslc, torch.complex64).contiguous()
    amp = torch.empty(z.shape, dtype=torch.float32, device=z.device)
    _lib.call("mr_rslc2amp", _dev.ptr(z), z.numel(), _dev.ptr(amp), _dev.stream())
    return _dev.back(amp, as_torch)
