"""Device buffers for structured records, held in torch tensors (plumbing only).

The library takes raw device pointers; torch provides the allocation, the
stream and host<->device copies so tests and the benchmark need no HIP
bindings of their own.
"""
from __future__ import annotations

import numpy as np


def _torch():
    import torch
    return torch


def to_device(arr: np.ndarray, device: str = "cuda"):
    """Copy a numpy structured/plain array to a uint8 device tensor of the same bytes."""
    torch = _torch()
    raw = np.ascontiguousarray(arr).view(np.uint8).reshape(-1)
    return torch.from_numpy(raw.copy()).to(device)


def empty(n: int, dtype: np.dtype, device: str = "cuda"):
    torch = _torch()
    return torch.empty(int(n) * np.dtype(dtype).itemsize, dtype=torch.uint8, device=device)


def zeros(n: int, dtype: np.dtype, device: str = "cuda"):
    torch = _torch()
    return torch.zeros(int(n) * np.dtype(dtype).itemsize, dtype=torch.uint8, device=device)


def to_host(t, dtype: np.dtype) -> np.ndarray:
    """Device uint8 tensor -> numpy array of `dtype` (synchronous copy); a
    subarray dtype such as (float32, 4) gives shape (n, 4)."""
    dtype = np.dtype(dtype)
    raw = t.cpu().numpy()
    if dtype.shape:
        return raw.view(dtype.base).reshape(-1, *dtype.shape).copy()
    return raw.view(dtype).copy()


def ptr(t) -> int:
    return t.data_ptr()


def synchronize() -> None:
    """Wait for torch's work on the current device (copies into buffers the library reads)."""
    _torch().cuda.synchronize()
