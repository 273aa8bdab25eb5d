"""Host logic of bench.py, no GPU: the one-GPU guard, the reader of the
committed traffic summary and the roofline built from it, the rocprof kernel
labels, and the comparison behind the `exact` flags (closest-hit records bit
for bit, any-hit records by occlusion only)."""
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


def test_read_profile_and_roofline(tmp_path):
    f = tmp_path / "t.txt"
    f.write_text(f"# comment\nkernel_sources_sha256: {bench.kernel_digest()}\ntree_sha256: 0123abcd\n"
                 "hbm_bytes_per_launch k_a: 2000000000\nl2_bytes_per_launch k_a: 9.8304e9\n")
    p = bench.read_profile(str(f))
    assert p["fresh"] and p["hbm"]["k_a"] == 2e9 and p["l2"]["k_a"] == 9.8304e9 and p["tree"] == "0123abcd"
    # measured on another tree: stale, with the reason
    other = bench.check_tree(bench.read_profile(str(f)), "ffff0000")
    assert not other["fresh"] and "ffff0000" in bench.roofline_of(other, "k_a", 1.0, 1, 96)["traffic_source"]
    assert bench.check_tree(p, "0123abcd")["fresh"]
    r = bench.roofline_of(p, "k_a", 1.0, 17_000_000_000, 96)
    assert r["achieved"] == 2000.0 and r["frac"] == 0.25 and r["traffic"] == 2e9
    assert r["algorithmic"]["GB_s"] == 17000.0 and r["l2"]["GB_s"] == 9830.4
    assert 0 < r["l2"]["frac"] < 1
    # a kernel the summary does not hold, and a summary of other kernel sources
    assert bench.roofline_of(p, "k_b", 1.0, 1, 96)["frac"] is None
    f.write_text("kernel_sources_sha256: 0000\nhbm_bytes_per_launch k_a: 2e9\n")
    stale = bench.read_profile(str(f))
    assert not stale["fresh"]
    r = bench.roofline_of(stale, "k_a", 1.0, 1, 96)
    assert r["frac"] is None and r["traffic_source"].startswith("stale")
    assert not bench.read_profile(str(tmp_path / "missing.txt"))["fresh"]


def test_vmem_ops():
    sets = [{"rays": 10, "node_visits": 100, "tri_tests": 20}, {"rays": 5, "node_visits": 40, "tri_tests": 0}]
    v = bench.vmem_of(sets, "wide", 1.0, {"l1_miss": 1e6, "l1_hit": 2e6})
    ops = 6 * 140 + 3 * 20 + 3 * 15
    assert v["lane_ops_per_launch"] == ops
    assert v["frac_of_l1_miss_rate"] == round(ops * 1e3 / 1e6, 3)
    assert v["frac_of_l1_hit_rate"] == round(ops * 1e3 / 2e6, 3)
    assert bench.vmem_of(sets, "learned_wide", 1.0, None)["lane_ops_per_launch"] == ops + 140
    assert "probe_G_per_s" not in bench.vmem_of(sets, "binary", 1.0, None)


def test_kernel_labels():
    assert bench.kernel_label("void rtq::(anonymous namespace)::k_trace<0, 1, false, 0, 0>(rtq::x)") == \
        "k_trace<watertight,any,binary>"
    assert bench.kernel_label("rtq::(anonymous namespace)::k_trace_occ<0, 2, false, 1, 2, 6>(x)") == \
        "k_trace<watertight,per-item,wide,batch>"
    assert bench.kernel_label("k_trace_occ<0, 0, false, 1, 0, 7>") == "k_trace<watertight,closest,wide>"
    assert bench.kernel_label("k_trace_occ<0, 2, false, 3, 2, 6>") == "k_trace<watertight,per-item,learned_wide,batch>"
    assert bench.kernel_label("k_trace_occ<1, 1, false, 2, 0, 7>") == "k_trace<refmt,any,learned>"
    assert bench.kernel_label("k_trace<0, 0, true, 0, 0>") is None   # counter instantiation
    assert bench.kernel_label("__amd_rocclr_copyBuffer") is None


def test_tree_of_log(tmp_path):
    f = tmp_path / "run.log"
    f.write_text("rocprofv3 noise\n" + json.dumps({"metric": "m", "detail": {"tree_sha256": "abc123"}}) +
                 "\n" + json.dumps({"copy_calibration_bytes": 1}) + "\n")
    assert bench.tree_of_log(str(f)) == "abc123"
    f.write_text("no bench line\n")
    assert bench.tree_of_log(str(f)) is None


def test_committed_traffic_covers_every_reported_kernel():
    """The committed summary was measured on the current kernel sources (the
    bench checks the terrain tree it records at run time) and holds HBM and L2
    bytes for every kernel the bench reports."""
    prof = bench.read_profile(str(ROOT / "profiles" / "traffic_terrain_trace.txt"))
    assert prof["fresh"], f"re-profile: summary digest {prof['digest']} != sources {bench.kernel_digest()}"
    assert prof["tree"], "the summary records the terrain tree it was measured on"
    names = [f"k_trace<watertight,{m},{l}>" for m in ("any", "closest") for l in ("binary", "wide")]
    names += [f"k_trace<watertight,per-item,{l},batch>" for l in ("binary", "wide")]
    for n in names:
        assert prof["hbm"].get(n, 0) > 0 and prof["l2"].get(n, 0) > 0, n


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
