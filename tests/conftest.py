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


def far_origin_rays(v9: np.ndarray, n: int, distance: float, aim: str, seed: int) -> np.ndarray:
    """Rays aimed at points on triangle edges ("edge") or at vertices
    ("vertex") of random triangles, from `distance` away along a random
    direction: the origin is the aim point minus distance * d, rounded to
    float32, so |o| grows with the distance while the ray still grazes the
    edge or vertex."""
    from realtimeraytracing_amd import RAY_DTYPE
    rng = np.random.default_rng(seed)
    tri = v9[rng.integers(0, len(v9), n)].reshape(n, 3, 3).astype(np.float64)
    a = rng.integers(0, 3, n)
    pa = tri[np.arange(n), a]
    pb = tri[np.arange(n), (a + 1) % 3]
    s = rng.uniform(0.0, 1.0, (n, 1)) if aim == "edge" else np.zeros((n, 1))
    target = pa + s * (pb - pa)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    r = np.zeros(n, RAY_DTYPE)
    r["o"] = (target - distance * d).astype(np.float32)
    r["d"] = d.astype(np.float32)
    r["tmin"] = 0.0
    r["tmax"] = np.inf
    return r


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
