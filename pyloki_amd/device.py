"""Device plumbing: torch tensors hold HBM buffers and supply the HIP stream.

PyTorch-ROCm is used only for device memory and streams; every kernel of the search path
is launched by the HIP library through the C ABI.
"""

from __future__ import annotations

import ctypes

import numpy as np
import torch


def require_gpu() -> torch.device:
    if not torch.cuda.is_available():
        msg = "pyloki_amd needs a ROCm GPU (MI355X/gfx950); none is visible"
        raise RuntimeError(msg)
    return torch.device("cuda", torch.cuda.current_device())


def stream_ptr(stream: torch.cuda.Stream | None = None) -> ctypes.c_void_p:
    s = stream if stream is not None else torch.cuda.current_stream()
    return ctypes.c_void_p(s.cuda_stream)


def ptr(t: torch.Tensor | None) -> ctypes.c_void_p:
    if t is None:
        return ctypes.c_void_p(0)
    if not t.is_cuda:
        msg = "expected a device tensor"
        raise ValueError(msg)
    if not t.is_contiguous():
        msg = "expected a contiguous tensor"
        raise ValueError(msg)
    return ctypes.c_void_p(t.data_ptr())


def to_device(arr: np.ndarray, dtype: torch.dtype | None = None) -> torch.Tensor:
    dev = require_gpu()
    t = torch.from_numpy(np.ascontiguousarray(arr))
    if dtype is not None:
        t = t.to(dtype)
    return t.to(dev)
