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
    """Device uint8 tensor -> numpy array of `dtype` (synchronous copy)."""
    return t.cpu().numpy().view(dtype).copy()


def ptr(t) -> int:
    return t.data_ptr()
