"""Host logic of bench.py, no GPU: the one-GPU guard, the reader of the
committed HBM-traffic summary, and the comparison behind the `exact` flags
(closest-hit records bit for bit, any-hit records by occlusion only)."""
import json
import subprocess
import sys
from pathlib import Path

import numpy as np

import bench
from realtimeraytracing_amd import HIT_DTYPE

ROOT = Path(__file__).resolve().parents[1]


def test_more_than_one_gpu_exits_with_a_message():
    r = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 2
    msg = json.loads(r.stdout.strip().splitlines()[-1])
    assert "one GPU" in msg["error"]


def test_read_traffic(tmp_path):
    f = tmp_path / "t.txt"
    f.write_text("# comment\ntraffic_bytes_per_launch k_a: 123\ntraffic_bytes_per_launch k_b: 4.5e9\n")
    assert bench.read_traffic(str(f), "k_a") == 123.0
    assert bench.read_traffic(str(f), "k_b") == 4.5e9
    assert bench.read_traffic(str(f), "k_c") is None
    assert bench.read_traffic(str(tmp_path / "missing.txt"), "k_a") is None


def test_committed_traffic_covers_every_reported_kernel():
    path = str(ROOT / "profiles" / "traffic_terrain_trace.txt")
    names = [f"k_trace<watertight,{m},{l}>" for m in ("any", "closest") for l in ("binary", "wide")]
    names += [f"k_trace<watertight,per-item,{l},batch>" for l in ("binary", "wide", "learned")]
    for n in names:
        v = bench.read_traffic(path, n)
        assert v is not None and v > 0, n


def _hits(ids, ts):
    h = np.zeros(len(ids), HIT_DTYPE)
    h["id"], h["t"] = ids, ts
    return h


def test_same_results_rule():
    d = _hits([3, -1, 7], [0.5, np.inf, 2.0])
    s = _hits([1, -1, 4, 9], [0.1, np.inf, 0.3, 0.2])
    assert bench.same_results((d, s), (d.copy(), s.copy()))
    # any hit: another occluder is the same result, a lost occlusion is not
    s2 = s.copy()
    s2["id"][0], s2["t"][0] = 5, 0.7
    assert bench.same_results((d, s), (d, s2))
    s3 = s.copy()
    s3["id"][3] = -1
    assert not bench.same_results((d, s), (d, s3))
    # closest hit: any bit counts
    d2 = d.copy()
    d2["t"][0] = np.nextafter(np.float32(0.5), np.float32(1.0))
    assert not bench.same_results((d, s), (d2, s))
