import os
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs on the GPU box)")


def hits_equal(a: np.ndarray, b: np.ndarray) -> bool:
    """Bitwise equality of two hit (or any record) arrays."""
    a = np.ascontiguousarray(a)
    b = np.ascontiguousarray(b)
    return a.shape == b.shape and np.array_equal(a.view(np.uint8), b.view(np.uint8))


def first_diff(a: np.ndarray, b: np.ndarray) -> str:
    a = np.ascontiguousarray(a)
    b = np.ascontiguousarray(b)
    ra = a.view(np.uint8).reshape(len(a), -1)
    rb = b.view(np.uint8).reshape(len(b), -1)
    bad = np.nonzero((ra != rb).any(axis=1))[0]
    if len(bad) == 0:
        return "no difference"
    i = bad[0]
    return f"{len(bad)} records differ; first at {i}: {a[i]} vs {b[i]}"


@pytest.fixture(scope="session")
def rtq_ctx():
    """One library context on cuda:0 for the whole GPU session."""
    from realtimeraytracing_amd import Context
    import torch
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    ctx = Context(0)
    yield ctx
    ctx.close()


@pytest.fixture(scope="session")
def torch_mod():
    import torch
    return torch
