"""Array plumbing between the moraine-style API and the C ABI.

Inputs may be numpy arrays (host) or torch tensors (ROCm device).  Mirroring
moraine's numpy/cupy dispatch (moraine/utils_.py:33-41), results come back in
the input's kind: torch in -> torch (same device) out, numpy in -> numpy out.
Either way the computation runs in the HIP kernels; numpy inputs are copied to
the GPU and the results copied back.
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch


def is_torch(x) -> bool:
    return isinstance(x, torch.Tensor)


def device() -> torch.device:
    if not torch.cuda.is_available():
        raise RuntimeError("moraine_amd runs on a ROCm GPU; torch.cuda.is_available() is False")
    return torch.device("cuda", torch.cuda.current_device())


def to_dev(x, dtype: torch.dtype | None = None) -> torch.Tensor:
    """Contiguous device tensor view/copy of ``x`` (never aliases a host array)."""
    t = x if is_torch(x) else torch.from_numpy(np.ascontiguousarray(x))
    if dtype is not None and t.dtype != dtype:
        t = t.to(dtype)
    if not t.is_cuda:
        t = t.to(device())
    return t.contiguous()


def to_dev_complex(x) -> torch.Tensor:
    """Contiguous complex device tensor of ``x``, keeping the reference's dtype
    rule (outputs have the input's precision): complex128 (or float64) input
    becomes complex128 and runs the FP64 kernels; anything else becomes
    complex64 (float32/complex64 exactly)."""
    t = x if is_torch(x) else torch.from_numpy(np.ascontiguousarray(x))
    dt = torch.complex128 if t.dtype in (torch.complex128, torch.float64) else torch.complex64
    return to_dev(t, dt)


def is_c128(t: torch.Tensor) -> bool:
    return t.dtype == torch.complex128


def ptr(t: torch.Tensor | None):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def back(t: torch.Tensor, as_torch: bool):
    return t if as_torch else t.cpu().numpy()


def workspace(nbytes: int, dev: torch.device) -> torch.Tensor:
    return torch.empty(max(int(nbytes), 256), dtype=torch.uint8, device=dev)
