#!/usr/bin/env python3
"""Benchmark: Mrays/s on the 2M-triangle terrain, diffuse + shadow rays.

Workload (one GPU): a 1024 x 1024 height-field terrain (2,093,058 triangles,
6-octave value noise), BVH built on the GPU (PLOC, radius 16, <= 4 triangles
per traversal leaf).  1920 x 1080 primary rays are traced once to find the
surface points; from each hit one cosine-weighted diffuse ray (closest hit)
and four shadow rays to point lights at height 2 (any hit) are generated on
the GPU.  One step = trace all diffuse rays + trace all shadow rays, inputs
resident in HBM.  Mrays/s counts only rays with a non-empty interval (every
primary ray hits, so all 10,368,000 are live), M = 10^6.

Forms measured on the same scene and rays (hit records are bit-identical
between them; closest-hit records and any-hit occlusion flags are compared
here too, `exact`):
  plain       binary 64-B nodes, one launch per ray set, default settings
  compressed  8-wide 128-B nodes with 8-bit outward-rounded child boxes
  batched     both ray sets in one persistent launch replayed from a
              hipGraph, settings (refill, chunk, node layout) from the
              autotuner
  learned     the batched form with the learned child order (model trained
              on this scene's own traversal logs before timing): on 8-wide
              nodes a learned first child per node, ray octant and mode, on
              binary nodes a learned child order
`value` is the fastest form; `detail.forms` has every form.

Further workloads in `detail` (not the headline):
`detail.compressed_8m`: four tiled terrains (8,372,232 triangles), 2048 x 2048
primary + 4M diffuse rays, 8-wide compressed against binary nodes, hit
records compared whole, t error against an FP64 ray/plane distance.
`detail.uneven_batch`: 64 scenes of 1e3-4e6 triangles with 64k-2M rays each,
one graph-replayed launch against 64 launches, results compared.
`detail.occluded` measures the learned-ordering workload: the same terrain
plus 2,097,152 random-soup triangles over the unit cube, 2048 x 2048 primary
hits -> 4M diffuse + 16M shadow rays (nearly all occluded): plain, 8-wide
compressed nodes, learned order on binary nodes and on 8-wide nodes (two
launches each) and learned order batched in one graph-replayed launch, with
the same exactness check.

Prints one JSON line with the metric, a roofline object for the dominant
trace kernel of the best form (HBM bytes per launch from the committed,
calibrated rocprofv3 counters over the launch's HIP-event duration, with the
algorithmic bytes and the L2 rate beside them) and the CPU baseline (the
reference traversal restated in C, run on this host's cores over the same
rays; 1 warm-up, median of 5).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
# L2 peak: 2 KiB per clock per XCD (16 channels x 128 B), 8 XCDs, 2.4 GHz
L2_PEAK_GBS = 2048 * 8 * 2.4
TRI_BYTES = 48       # leaf triangle record
RAY_BYTES = 32
HIT_BYTES = 16
BINARY_NODE_BYTES = 64   # one binary node per visit
WIDE_NODE_BYTES = 96     # an 8-wide visit reads 96 B of its 128-B slot
LEARNED_WIDE_NODE_BYTES = 104   # with the learned first-child table (bytes 96-103)
NODE_BYTES = {"binary": BINARY_NODE_BYTES, "wide": WIDE_NODE_BYTES, "learned": BINARY_NODE_BYTES,
              "learned_wide": LEARNED_WIDE_NODE_BYTES}
# sources of the traversal kernels; the committed traffic summary records
# their digest and the fingerprint of the terrain tree it was measured on
# (tree_digest), and is ignored when either differs
KERNEL_SOURCES = ("trace.hip", "batch.hip", "trace_layout.hpp")


def kernel_label(rocprof_name: str):
    """bench's name of a trace kernel from its rocprof name (template arguments
    leaf test, mode, counters, layout, batch[, occupancy]), or None."""
    import re
    m = re.search(r"k_trace(?:_occ)?<(\d+), (\d+), (false|true), (\d+), (\d+)", rocprof_name)
    if not m or m.group(3) == "true":
        return None
    leaf = "watertight" if m.group(1) == "0" else "refmt"
    mode = {"0": "closest", "1": "any", "2": "per-item"}[m.group(2)]
    layout = {"0": "binary", "1": "wide", "2": "learned", "3": "learned_wide"}[m.group(4)]
    return f"k_trace<{leaf},{mode},{layout}" + (",batch>" if m.group(5) != "0" else ">")


def summarize_profile(out_path: str, dirs: list[str], commands: list[str], last: int = 5, tree=None) -> str:
    """Writes the traffic summary read by read_profile from rocprofv3 --pmc
    CSVs (FETCH_SIZE, WRITE_SIZE, TCC_READ_SECTORS_sum + TCC_WRITE_SECTORS_sum
    passes of the same runs, each ending with a 512 MiB device copy): per
    trace kernel the median over its last `last` dispatches in the run that
    dispatches it most, each counter scaled by what it reads for the copy
    (512 MiB read, 512 MiB written)."""
    import collections
    import csv
    copy_bytes = 512 << 20
    # counter -> label -> (dispatches in the run, median of its last ones); the
    # run that dispatches a kernel most holds its timed launches (every run
    # also traces its primary rays once with the binary closest-hit kernel)
    vals = collections.defaultdict(dict)
    scale = {}
    for d in dirs:
        per = collections.defaultdict(lambda: collections.defaultdict(float))   # dispatch -> counter -> sum
        names = {}
        for f in Path(d).glob("*counter_collection.csv"):
            for r in csv.DictReader(open(f)):
                key = int(r["Dispatch_Id"])
                per[key][r["Counter_Name"]] += float(r["Counter_Value"])
                names[key] = r["Kernel_Name"]
        by_label = collections.defaultdict(list)
        copies = []
        for key in sorted(per):
            if "copyBuffer" in names[key]:
                copies.append(key)
            lab = kernel_label(names[key])
            if lab:
                by_label[lab].append(key)
        for counter in {c for k in per for c in per[k]}:
            if copies:
                best = max(per[k].get(counter, 0.0) for k in copies[-3:])
                if counter in ("FETCH_SIZE", "WRITE_SIZE"):
                    scale[counter] = copy_bytes / (best * 1024.0) if best else None
                else:
                    scale[counter] = (copy_bytes / 32.0) / best if best else None
            for lab, keys in by_label.items():
                cand = (len(keys), float(np.median([per[k].get(counter, 0.0) for k in keys[-last:]])))
                if cand[0] > vals[counter].get(lab, (0, 0.0))[0]:
                    vals[counter][lab] = cand
    lines = ["# HBM and L2 bytes per launch of the terrain trace kernels (rocprofv3 --pmc, calibrated).",
             "# A 512 MiB device copy at the end of each run gives the scale of every counter: bytes =",
             "# FETCH_SIZE x 1024 x s_F + WRITE_SIZE x 1024 x s_W (HBM) and 32 B x (TCC_READ_SECTORS x s_R",
             "# + TCC_WRITE_SECTORS x s_W2) (L2); medians over each kernel's last %d dispatches in the run" % last,
             "# that dispatches it most (its timed launches; batched forms run last in their pass).",
             "# scales: " + ", ".join(f"{c} {v:.4f}" for c, v in sorted(scale.items()) if v),
             "# commands (repository root $R, from /tmp with TMPDIR=/tmp):"]
    lines += [f"#   {c}" for c in commands]
    lines.append(f"kernel_sources_sha256: {kernel_digest()}")
    if tree:
        lines.append(f"tree_sha256: {tree}")
    labels = sorted({lab for c in vals for lab in vals[c]})
    def get(counter, lab):
        v = vals[counter].get(lab)
        return v[1] if v else None

    for lab in labels:
        f_, w_ = get("FETCH_SIZE", lab), get("WRITE_SIZE", lab)
        if f_ is not None and w_ is not None and scale.get("FETCH_SIZE") and scale.get("WRITE_SIZE"):
            hbm = f_ * 1024 * scale["FETCH_SIZE"] + w_ * 1024 * scale["WRITE_SIZE"]
            lines.append(f"hbm_bytes_per_launch {lab}: {hbm:.0f}")
        r_, w2 = get("TCC_READ_SECTORS_sum", lab), get("TCC_WRITE_SECTORS_sum", lab)
        if r_ is not None and w2 is not None and scale.get("TCC_READ_SECTORS_sum") and \
                scale.get("TCC_WRITE_SECTORS_sum"):
            l2 = 32.0 * (r_ * scale["TCC_READ_SECTORS_sum"] + w2 * scale["TCC_WRITE_SECTORS_sum"])
            lines.append(f"l2_bytes_per_launch {lab}: {l2:.0f}")
    text = "\n".join(lines) + "\n"
    Path(out_path).write_text(text)
    return text


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--forms", default="plain,compressed,batched,learned")
    ap.add_argument("--max-leaf", type=int, default=4, help="triangles per traversal leaf (1..8)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-occluded", action="store_true", help="skip the terrain + soup learned-order workload")
    ap.add_argument("--occluded-res", type=int, default=2048, help="primary image side of the occluded workload")
    ap.add_argument("--no-compressed-8m", action="store_true", help="skip the 4-terrain 8-wide workload")
    ap.add_argument("--no-uneven-batch", action="store_true", help="skip the 64-scene uneven batch")
    ap.add_argument("--no-shade", action="store_true", help="skip shading / BVH-view timing")
    ap.add_argument("--no-refit", action="store_true", help="skip the scene-edit refit timing")
    ap.add_argument("--no-build-timing", action="store_true", help="skip the fused / split build timing")
    ap.add_argument("--batched-layout", choices=("tuned", "binary", "wide"), default="tuned",
                    help="node layout of the batched form: the autotuner's choice, or pinned (profiling one kernel)")
    ap.add_argument("--no-cpu-reference-case", action="store_true",
                    help="skip the 65,536-triangle soup against the CPU reference traversal")
    ap.add_argument("--traffic-file", default=str(ROOT / "profiles" / "traffic_terrain_trace.txt"))
    ap.add_argument("--cpu-repeats", type=int, default=5, help="CPU baseline: timed runs after one warm-up (median)")
    ap.add_argument("--summarize-profile", nargs="+", metavar=("OUT", "PMC_DIR"),
                    help="write the traffic summary OUT from rocprofv3 --pmc output directories and exit")
    ap.add_argument("--profile-commands", nargs="*", default=[], help="commands recorded in the summary")
    ap.add_argument("--profile-tree-log", default=None,
                    help="output of one of the profiled bench runs: the summary records its terrain tree")
    ap.add_argument("--copy-calibration-mb", type=int, default=0,
                    help="after the run, copy a buffer of this many MiB device-to-device 3 times: a kernel of "
                         "known traffic to scale rocprof FETCH_SIZE / WRITE_SIZE readings")
    return ap.parse_args()


def kernel_digest() -> str:
    """SHA-256 (first 16 hex digits) of the traversal kernels' sources."""
    import hashlib
    h = hashlib.sha256()
    for name in KERNEL_SOURCES:
        h.update((ROOT / "realtimeraytracing_amd" / "csrc" / name).read_bytes())
    return h.hexdigest()[:16]


def tree_digest(sc) -> str:
    """SHA-256 (first 16 hex digits) of a built scene's traversal layouts: the
    binary nodes with their leaf ranges, and the 8-wide nodes if built."""
    import hashlib
    h = hashlib.sha256()
    nodes, ranges = sc.export_trav()
    h.update(nodes.tobytes())
    h.update(ranges.tobytes())
    if sc.info().n_wide_nodes:
        h.update(sc.export_wide().tobytes())
    return h.hexdigest()[:16]


def tree_of_log(path: str):
    """The tree fingerprint a bench run printed (detail.tree_sha256 of its JSON
    line), or None."""
    for line in Path(path).read_text(errors="replace").splitlines():
        if line.startswith("{") and "tree_sha256" in line:
            try:
                return json.loads(line)["detail"]["tree_sha256"]
            except (ValueError, KeyError, TypeError):
                continue
    return None


def read_profile(path: str) -> dict:
    """Per-kernel bytes per launch from the committed rocprof counter summary:
    `hbm_bytes_per_launch <kernel>: v` (calibrated FETCH_SIZE + WRITE_SIZE) and
    `l2_bytes_per_launch <kernel>: v` (TCC read + write sectors x 32 B), with
    the kernel-source digest and terrain tree they were measured at
    (`kernel_sources_sha256: d`, `tree_sha256: t`).  `fresh` is False when the
    digest differs from the current sources; the caller compares the tree."""
    out = {"hbm": {}, "l2": {}, "digest": None, "tree": None, "fresh": False, "path": path, "stale": None}
    p = Path(path)
    if not p.exists():
        return out
    for line in p.read_text().splitlines():
        for key, field in (("hbm_bytes_per_launch ", "hbm"), ("l2_bytes_per_launch ", "l2")):
            if line.startswith(key):
                k, v = line[len(key):].rsplit(":", 1)
                out[field][k.strip()] = float(v)
        if line.startswith("kernel_sources_sha256:"):
            out["digest"] = line.split(":", 1)[1].strip()
        if line.startswith("tree_sha256:"):
            out["tree"] = line.split(":", 1)[1].strip()
    out["fresh"] = out["digest"] == kernel_digest()
    if not out["fresh"]:
        out["stale"] = (f"{p.name} was measured at kernel sources {out['digest']}, current {kernel_digest()}")
    return out


def check_tree(prof: dict, tree: str) -> dict:
    """Marks the summary stale when it was measured on another terrain tree."""
    if prof["fresh"] and prof["tree"] != tree:
        prof["fresh"] = False
        prof["stale"] = f"{Path(prof['path']).name} was measured on terrain tree {prof['tree']}, current {tree}"
    return prof


# per-lane vector-memory operations of one visit, by node layout: 16-B loads
# of an 8-wide node (96 B; + one 8-B load of the learned table), 16-B loads of
# a 64-B binary node (visits of the top-of-tree table read LDS: an upper bound)
VMEM_OPS_PER_VISIT = {"wide": 6, "learned_wide": 7, "binary": 4, "learned": 4}
VMEM_OPS_PER_TRI = 3        # three 16-B vertex loads
VMEM_OPS_PER_RAY = 3        # two 16-B ray loads, one 16-B hit store


def probe_vmem(slots_miss: int = 219_000, slots_hit: int = 64) -> dict | None:
    """Per-lane 16-B load rates of dependent chains (librtqprobe.so, six loads
    per 128-B slot as an 8-wide visit, 6 workgroups of 256 lanes per CU):
    over 28 MB of slots (loads miss L1) and over 64 slots (loads hit L1), with
    all 64 lanes of each wave loading and (`_32`) with lanes 0-31 only."""
    import ctypes
    lib_path = ROOT / "realtimeraytracing_amd" / "librtqprobe.so"
    if not lib_path.exists():
        return None
    lib = ctypes.CDLL(str(lib_path))
    fn = lib.rtqprobe_vmem_rate_lanes
    fn.argtypes = [ctypes.c_int, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
                   ctypes.POINTER(ctypes.c_double)]
    fn.restype = ctypes.c_int
    out = {}
    for lanes, suffix in ((64, ""), (32, "_32")):
        for name, slots in (("l1_miss", slots_miss), ("l1_hit", slots_hit)):
            r = ctypes.c_double(0.0)
            rc = fn(0, 6, slots, 6, lanes, ctypes.byref(r))
            if rc != 0:
                raise RuntimeError(f"vector-memory probe failed: HIP error {rc}")
            out[name + suffix] = r.value
    return out


def vmem_of(sets: list, layout: str, launch_ms: float, probe) -> dict:
    """Per-lane vector-memory operations of one launch (from the traversal
    counters of its ray sets) per second, against the probe's rates."""
    ops = sum(VMEM_OPS_PER_VISIT[layout] * c["node_visits"] + VMEM_OPS_PER_TRI * c["tri_tests"] +
              VMEM_OPS_PER_RAY * c["rays"] for c in sets)
    rate = ops / (launch_ms * 1e-3)
    out = {"lane_ops_per_launch": int(ops), "achieved_G_per_s": round(rate / 1e9, 1),
           "ops": f"{VMEM_OPS_PER_VISIT[layout]} x node visits + {VMEM_OPS_PER_TRI} x triangle tests + "
                  f"{VMEM_OPS_PER_RAY} x rays (16-B lane loads / stores)"}
    if probe:
        out.update({"probe_G_per_s": {k: round(v / 1e9, 1) for k, v in probe.items()},
                    "frac_of_l1_miss_rate": round(rate / probe["l1_miss"], 3),
                    "frac_of_l1_hit_rate": round(rate / probe["l1_hit"], 3)})
    return out


def roofline_of(prof: dict, kernel: str, launch_ms: float, alg_bytes: int, node_bytes: int) -> dict:
    """HBM roofline of one launch: calibrated DRAM bytes over its HIP-event
    time against the 8 TB/s peak; beside it the algorithmic bytes (ray + hit +
    node and triangle fetches, most of them served by L2 and the Infinity
    Cache) and the L2 request bytes against the L2 peak."""
    s = launch_ms * 1e-3
    hbm = prof["hbm"].get(kernel) if prof["fresh"] else None
    l2 = prof["l2"].get(kernel) if prof["fresh"] else None
    achieved = hbm / s / 1e9 if hbm else None
    l2_gbs = l2 / s / 1e9 if l2 else None
    src = (f"{Path(prof['path']).name}: rocprofv3 FETCH_SIZE / WRITE_SIZE (calibrated) and TCC sectors, kernel "
           f"sources {prof['digest']}, terrain tree {prof['tree']}") if prof["fresh"] else \
        f"stale: {prof['stale'] or Path(prof['path']).name}"
    return {"bound": "hbm", "achieved": round(achieved, 1) if achieved else None, "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4) if achieved else None,
            "traffic": hbm, "kernel": kernel, "launch_ms": round(launch_ms, 4), "traffic_source": src,
            "algorithmic": {"bytes_per_launch": alg_bytes, "GB_s": round(alg_bytes / s / 1e9, 1),
                            "bytes_per_ray": f"32 ray + 16 hit + {node_bytes} x node visits + 48 x triangle tests",
                            "note": "node and triangle fetches mostly hit L2 and the 256 MB Infinity Cache, so "
                                    "this rate is not an HBM rate"},
            "l2": {"bytes_per_launch": l2, "GB_s": round(l2_gbs, 1) if l2_gbs else None, "peak_GB_s": L2_PEAK_GBS,
                   "frac": round(l2_gbs / L2_PEAK_GBS, 4) if l2_gbs else None}}


class Workload:
    """A scene with primary hits and the diffuse + shadow rays built from them."""

    def __init__(self, ctx, tris, models, cam, stream, max_leaf, node_width):
        from realtimeraytracing_amd import HIT_DTYPE, RAY_DTYPE, RTQ_TRACE_BINARY, devmem, scenes
        import torch
        self.ctx, self.sp = ctx, stream
        self.sc = ctx.upload(tris, models)
        self.info = self.sc.build(16, max_leaf, 0, node_width=node_width)
        self.tree = tree_digest(self.sc)
        n = cam.width * cam.height
        self.prim, self.phit = devmem.empty(n, RAY_DTYPE), devmem.empty(n, HIT_DTYPE)
        ctx.raygen_pinhole(cam, self.prim.data_ptr(), stream)
        self.sc.trace(self.prim.data_ptr(), n, self.phit.data_ptr(), RTQ_TRACE_BINARY, stream)
        self.diff, self.shad = devmem.empty(n, RAY_DTYPE), devmem.empty(4 * n, RAY_DTYPE)
        self.sc.raygen_diffuse(self.prim.data_ptr(), self.phit.data_ptr(), n, scenes.DIFFUSE_SEED,
                               scenes.SECONDARY_EPS, self.diff.data_ptr(), stream)
        self.sc.raygen_shadow(self.prim.data_ptr(), self.phit.data_ptr(), n, scenes.LIGHTS, scenes.SECONDARY_EPS,
                              self.shad.data_ptr(), stream)
        self.hd, self.hs = devmem.empty(n, HIT_DTYPE), devmem.empty(4 * n, HIT_DTYPE)
        torch.cuda.synchronize()
        phit = devmem.to_host(self.phit, HIT_DTYPE)
        self.n_diff, self.n_shad = n, 4 * n
        n_live = int((phit["id"] >= 0).sum())
        self.live_rays = 5 * n_live

    def sets(self):
        from realtimeraytracing_amd import RTQ_TRACE_ANY
        return (("diffuse", self.diff, self.n_diff, self.hd, 0),
                ("shadow", self.shad, self.n_shad, self.hs, RTQ_TRACE_ANY))

    def counters(self, layout_flags, node_bytes):
        """Per-set traversal counters of a layout (outside any timed region)."""
        from realtimeraytracing_amd import devmem
        import torch
        out = {}
        for name, d_rays, cnt, d_hits, flags in self.sets():
            dv, dt = devmem.empty(cnt, np.uint32), devmem.empty(cnt, np.uint32)
            self.sc.trace_stats(d_rays.data_ptr(), cnt, d_hits.data_ptr(), dv.data_ptr(), dt.data_ptr(),
                                flags | layout_flags, self.sp)
            torch.cuda.synchronize()
            v = int(devmem.to_host(dv, np.uint32).astype(np.int64).sum())
            t = int(devmem.to_host(dt, np.uint32).astype(np.int64).sum())
            out[name] = dict(rays=cnt, node_visits=v, tri_tests=t,
                             bytes=cnt * (RAY_BYTES + HIT_BYTES) + v * node_bytes + t * TRI_BYTES)
        return out

    def hits(self):
        from realtimeraytracing_amd import HIT_DTYPE, devmem
        import torch
        torch.cuda.synchronize()
        return devmem.to_host(self.hd, HIT_DTYPE), devmem.to_host(self.hs, HIT_DTYPE)

    def release(self):
        self.sc.release()


def same_results(ref, got) -> bool:
    """Closest-hit records bit-identical, any-hit occlusion flags identical."""
    (rd, rs), (gd, gs) = ref, got
    return bool(np.array_equal(np.ascontiguousarray(rd).view(np.uint8), np.ascontiguousarray(gd).view(np.uint8))
                and np.array_equal(rs["id"] >= 0, gs["id"] >= 0))


# lights and diffuse seed of the rays the order model trains on, not the timed
# rays: each light 0.05-0.07 from one of the timed lights (the same layout,
# moved), as a renderer would train on earlier frames
TRAIN_LIGHTS = np.array([[0.3, 2.0, 0.2], [0.7, 2.0, 0.3], [0.2, 2.0, 0.7], [0.8, 2.0, 0.8]], np.float32)
TRAIN_DIFFUSE_SEED = 101


def train_order(ctx, wl, stream):
    """Order model from this scene's own traversal logs, applied to the scene's
    nodes; untimed setup.  The logs come from held-out rays of the same primary
    hits (other lights, another diffuse seed; about 200k shadow and 100k diffuse
    rays sampled), so the timed rays are not the training rays."""
    import torch
    from realtimeraytracing_amd import RAY_DTYPE, devmem, scenes
    from realtimeraytracing_amd.order import OrderModel
    t0 = time.perf_counter()
    model = OrderModel(ctx, seed=1)
    n = wl.n_diff
    tdiff, tshad = devmem.empty(n, RAY_DTYPE), devmem.empty(4 * n, RAY_DTYPE)
    wl.sc.raygen_diffuse(wl.prim.data_ptr(), wl.phit.data_ptr(), n, TRAIN_DIFFUSE_SEED, scenes.SECONDARY_EPS,
                         tdiff.data_ptr(), stream)
    wl.sc.raygen_shadow(wl.prim.data_ptr(), wl.phit.data_ptr(), n, TRAIN_LIGHTS, scenes.SECONDARY_EPS,
                        tshad.data_ptr(), stream)
    torch.cuda.synchronize()
    ss, ks = model.log(wl.sc, tshad, 4 * n, True, stride=max(1, 4 * n // 200_000), capacity=1 << 22)
    sd, kd = model.log(wl.sc, tdiff, n, False, stride=max(1, n // 100_000), capacity=1 << 22)
    del tdiff, tshad
    samples = torch.cat([ss, sd]).contiguous()
    losses = model.train(samples, epochs=6, batch=1024, lr=2e-3) if len(samples) else []
    # decisions only for any-hit rays, and only when the logs held any-hit samples
    model.apply(wl.sc, confidence_any=0.5 if ks else 1.0, confidence_closest=1.0)
    torch.cuda.synchronize()
    return {"samples_any": ks, "samples_closest": kd, "epochs": 6,
            "trained_on": "held-out rays: shadow rays to 4 other lights (each 0.05-0.07 from a timed one), "
                          "diffuse seed 101",
            "loss_first": round(float(losses[0]), 4) if len(losses) else None,
            "loss_last": round(float(losses[-1]), 4) if len(losses) else None,
            "setup_s": round(time.perf_counter() - t0, 2)}


def t_error_stats(tris, rays, hits, n_sample=20_000, seed=0):
    """t of the hits against the FP64 plane distance of the hit triangle, in
    ulps of the problem's scale max(|t|, |o|, |p0|): median, 99th percentile,
    max, and max of the error times |cos(incidence)|."""
    hit_idx = np.nonzero(hits["id"] >= 0)[0]
    idx = np.random.default_rng(seed).choice(hit_idx, size=min(n_sample, len(hit_idx)), replace=False)
    h, r = hits[idx], rays[idx]
    tri = tris[h["id"]]
    p0, p1, p2 = (tri[k][:, :3].astype(np.float64) for k in ("p0", "p1", "p2"))
    o, d = r["o"].astype(np.float64), r["d"].astype(np.float64)
    nrm = np.cross(p1 - p0, p2 - p0)
    t64 = np.einsum("ij,ij->i", nrm, p0 - o) / np.einsum("ij,ij->i", nrm, d)
    scale = np.maximum(np.abs(t64), np.maximum(np.abs(o).max(axis=1), np.abs(p0).max(axis=1)))
    err = np.abs(h["t"].astype(np.float64) - t64) / np.spacing(scale.astype(np.float32)).astype(np.float64)
    cos = np.abs(np.einsum("ij,ij->i", nrm, d)) / (np.linalg.norm(nrm, axis=1) * np.linalg.norm(d, axis=1))
    return {"hits_sampled": int(len(err)), "median_ulp": round(float(np.median(err)), 3),
            "p99_ulp": round(float(np.percentile(err, 99)), 3), "max_ulp": round(float(err.max()), 3),
            "max_ulp_times_cos": round(float((err * cos).max()), 3)}


def main() -> int:
    args = parse()
    if args.summarize_profile:
        tree = tree_of_log(args.profile_tree_log) if args.profile_tree_log else None
        print(summarize_profile(args.summarize_profile[0], args.summarize_profile[1:], args.profile_commands,
                                tree=tree))
        return 0
    if args.gpus != 1:
        print(json.dumps({"error": f"--gpus {args.gpus}: this build runs on one GPU only; the trace path "
                                   "would shard by ray ranges over a BVH replicated on every GPU"}))
        return 2
    import torch
    from realtimeraytracing_amd import (RTQ_BUILD_PLOC_SPLIT, RTQ_TRACE_ANY, RTQ_TRACE_BINARY, RTQ_TRACE_LEARNED, Context, devmem,
                                        scenes)

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    torch.cuda.set_device(0)
    # an explicit stream: the library launches on it and the events below see
    # exactly those launches (a null stream would mean the library's own stream)
    stream = torch.cuda.Stream()
    torch.cuda.set_stream(stream)
    sp = stream.cuda_stream
    assert sp != 0

    def time_steps(step):
        for _ in range(args.warmup):
            step()
        torch.cuda.synchronize()
        evs = [[torch.cuda.Event(enable_timing=True) for _ in range(3)] for _ in range(args.steps)]
        start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        start.record(stream)
        for k in range(args.steps):
            step(evs[k])
        stop.record(stream)
        torch.cuda.synchronize()
        return start.elapsed_time(stop) / args.steps, evs

    def two_launches(wl, layout):
        def step(ev=None):
            if ev is not None:
                ev[0].record(stream)
            wl.sc.trace(wl.diff.data_ptr(), wl.n_diff, wl.hd.data_ptr(), layout, sp)
            if ev is not None:
                ev[1].record(stream)
            wl.sc.trace(wl.shad.data_ptr(), wl.n_shad, wl.hs.data_ptr(), RTQ_TRACE_ANY | layout, sp)
            if ev is not None:
                ev[2].record(stream)
        return step

    def graph_step(graph):
        def step(ev=None):
            if ev is not None:
                ev[0].record(stream)
            graph.launch(sp)
            if ev is not None:
                ev[1].record(stream)
                ev[2].record(stream)
        return step

    t0 = time.time()
    ctx = Context(0)
    tris, models = scenes.terrain_scene(1024, seed=0)
    cam = scenes.pinhole(**scenes.TERRAIN_CAMERA)
    wl = Workload(ctx, tris, models, cam, sp, args.max_leaf, node_width=8)   # binary + 8-wide layouts
    info = wl.info
    setup_s = time.time() - t0

    default_params = ctx.launch_params()
    counters_by_layout = {"binary": wl.counters(RTQ_TRACE_BINARY, BINARY_NODE_BYTES),
                          "wide": wl.counters(0, WIDE_NODE_BYTES)}
    profile = check_tree(read_profile(args.traffic_file), wl.tree)
    vmem_probe = probe_vmem()
    reference = None
    results = {}
    learned_setup = None
    for form in args.forms.split(","):
        ctx.set_launch_params(default_params.refill, default_params.chunk, bool(default_params.prefer_wide),
                              default_params.spec_slack, default_params.leaf_slack,
                              default_params.stack_lds)
        tuned = None
        if form in ("plain", "compressed"):
            layout = RTQ_TRACE_BINARY if form == "plain" else 0
            lname = "binary" if form == "plain" else "wide"
            ms_per_step, evs = time_steps(two_launches(wl, layout))
            diff_ms = float(np.mean([e[0].elapsed_time(e[1]) for e in evs]))
            shad_ms = float(np.mean([e[1].elapsed_time(e[2]) for e in evs]))
            counters = counters_by_layout[lname]
            dom = "shadow" if shad_ms >= diff_ms else "diffuse"
            dom_ms = shad_ms if dom == "shadow" else diff_ms
            dom_bytes = counters[dom]["bytes"]
            dom_sets = [counters[dom]]
            kernel = f"k_trace<watertight,{'any' if dom == 'shadow' else 'closest'},{lname}>"
        elif form in ("batched", "learned"):
            # both ray sets in one persistent launch, replayed from a hipGraph,
            # launch settings (incl. node layout) from the autotuner
            extra = 0
            if form == "learned":
                learned_setup = train_order(ctx, wl, sp)
                extra = RTQ_TRACE_LEARNED
            items = ctx.items([(wl.sc, wl.diff.data_ptr(), wl.n_diff, wl.hd.data_ptr(), extra),
                               (wl.sc, wl.shad.data_ptr(), wl.n_shad, wl.hs.data_ptr(), RTQ_TRACE_ANY | extra)])
            tr = ctx.autotune(items, reps=2)
            if form == "batched" and args.batched_layout != "tuned":
                tr.best.prefer_wide = int(args.batched_layout == "wide")
                ctx.set_launch_params(tr.best.refill, tr.best.chunk, bool(tr.best.prefer_wide), tr.best.spec_slack,
                                      tr.best.leaf_slack, tr.best.stack_lds)
            tuned = {"refill": tr.best.refill, "chunk": tr.best.chunk, "prefer_wide": tr.best.prefer_wide,
                     "spec_slack": tr.best.spec_slack, "leaf_slack": tr.best.leaf_slack,
                     "stack_lds": tr.best.stack_lds, "candidates": tr.n_tried, "best_ms": round(tr.best_ms, 4),
                     "default_ms": round(tr.default_ms, 4)}
            if form == "learned" and wl.sc.info().order_modes:
                # learned decisions: the batch traces 8-wide nodes with their
                # learned first child, or binary nodes with the learned order
                lname = "learned_wide" if tr.best.prefer_wide else "learned"
                lflags = RTQ_TRACE_LEARNED | (0 if tr.best.prefer_wide else RTQ_TRACE_BINARY)
                counters_by_layout[lname] = wl.counters(lflags, NODE_BYTES[lname])
            else:
                # the batched form, or learned order without decisions (the plain
                # order), on the layout the autotuner chose
                lname = "wide" if tr.best.prefer_wide else "binary"
            graph = ctx.batch(items)
            ms_per_step, evs = time_steps(graph_step(graph))
            graph.destroy()
            diff_ms = shad_ms = None
            counters = counters_by_layout[lname]
            dom = "batch"
            dom_ms = float(np.mean([e[0].elapsed_time(e[1]) for e in evs]))
            dom_bytes = counters["diffuse"]["bytes"] + counters["shadow"]["bytes"]
            dom_sets = [counters["diffuse"], counters["shadow"]]
            kernel = f"k_trace<watertight,per-item,{lname},batch>"
        else:
            raise SystemExit(f"unknown form {form}")
        got = wl.hits()
        if reference is None:
            reference = got
        node_bytes = NODE_BYTES[lname]
        results[form] = {
            "Mrays_s": wl.live_rays / (ms_per_step * 1e-3) / 1e6,
            "ms_per_step": ms_per_step, "diffuse_ms": diff_ms, "shadow_ms": shad_ms,
            "layout": lname, "tuned": tuned, "counters": counters, "exact": same_results(reference, got),
            "roofline": roofline_of(profile, kernel, dom_ms, dom_bytes, node_bytes),
        }
        results[form]["roofline"]["vmem"] = vmem_of(dom_sets, lname, dom_ms, vmem_probe)
    ctx.set_launch_params(default_params.refill, default_params.chunk, bool(default_params.prefer_wide),
                              default_params.spec_slack, default_params.leaf_slack,
                              default_params.stack_lds)

    best = max(results, key=lambda f: results[f]["Mrays_s"])
    res = results[best]

    # CPU baseline: the reference traversal (raytracer.glsl:246-295 restated in
    # C: 48-B pre-order nodes, RefMT leaf test, no pruning) on this host's cores
    cpu = None
    if not args.no_cpu_baseline:
        import oracle
        from realtimeraytracing_amd import RAY_DTYPE
        nodes = wl.sc.export_ref()
        v9 = oracle.world_vertices(tris, models)
        diff = devmem.to_host(wl.diff, RAY_DTYPE)
        shad = devmem.to_host(wl.shad, RAY_DTYPE)
        cores = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
        cores = int(os.environ.get("OMP_NUM_THREADS", cores))
        def median_s(fn):
            fn()                                  # warm-up
            ts = []
            for _ in range(max(1, args.cpu_repeats)):
                t1 = time.perf_counter()
                fn()
                ts.append(time.perf_counter() - t1)
            return float(np.median(ts)), ts

        cpu_s, cpu_ts = median_s(lambda: (oracle.trace_ref(nodes, v9, diff, leaf=oracle.LEAF_REFMT, threads=cores),
                                          oracle.trace_ref(nodes, v9, shad, leaf=oracle.LEAF_REFMT, threads=cores)))
        cpu2_s, _ = median_s(lambda: (oracle.trace_bvh(nodes, v9, diff, threads=cores),
                                      oracle.trace_bvh(nodes, v9, shad, mode=oracle.MODE_ANY, threads=cores)))
        cpu = {"value": round(wl.live_rays / cpu_s / 1e6, 3), "unit": "Mrays/s", "cores": cores, "kind": "port",
               "sample": f"all {wl.n_diff} diffuse + {wl.n_shad} shadow rays of one step, reference traversal "
                         f"semantics (RefMT, no pruning); 1 warm-up, median of {len(cpu_ts)} runs "
                         f"({min(cpu_ts):.2f}-{max(cpu_ts):.2f} s)",
               "watertight_pruned_Mrays_s": round(wl.live_rays / cpu2_s / 1e6, 3)}

    # caller side of the path: shading the primary hits into an RGBA32F image
    # (getColor with wireframe) and the BVH view at depth 8, on the same frame
    shading = None
    if not args.no_shade:
        from realtimeraytracing_amd import HIT_DTYPE, RTQ_SHADE_WIREFRAME
        from realtimeraytracing_amd.render import materials_array
        n = wl.n_diff
        f4 = np.dtype((np.float32, 4))
        mats = devmem.to_device(materials_array([[0.2, 0.3, 0.1, 1.0]]))
        rgba, ov = devmem.empty(n, f4), devmem.empty(n, f4)
        torch.cuda.synchronize()

        def shade_step(ev=None):
            if ev is not None:
                ev[0].record(stream)
            wl.sc.shade(wl.phit.data_ptr(), n, mats.data_ptr(), 1, rgba.data_ptr(), None, RTQ_SHADE_WIREFRAME, sp)
            if ev is not None:
                ev[1].record(stream)
                ev[2].record(stream)

        def overlay_step(ev=None):
            if ev is not None:
                ev[0].record(stream)
            wl.sc.bvh_overlay(wl.prim.data_ptr(), n, 8, ov.data_ptr(), sp)
            if ev is not None:
                ev[1].record(stream)
                ev[2].record(stream)
        shade_ms, _ = time_steps(shade_step)
        ov_ms, _ = time_steps(overlay_step)
        n_hit = int((devmem.to_host(wl.phit, HIT_DTYPE)["id"] >= 0).sum())
        # per pixel: 16-B hit in, 16-B pixel out; per hit: the triangle's model id,
        # the model's material id and the 16-B colour (4 + 4 + 16 B)
        sbytes = 32 * n + 24 * n_hit
        shading = {"pixels": n, "hits": n_hit, "shade_ms": round(shade_ms, 5),
                   "shade_GB_s": round(sbytes / (shade_ms * 1e-3) / 1e9, 1),
                   "shade_hbm_frac": round(sbytes / (shade_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                   "shade_algorithmic_bytes": sbytes,
                   "overlay_depth": 8, "overlay_ms": round(ov_ms, 4),
                   "overlay_Mrays_s": round(n / (ov_ms * 1e-3) / 1e6, 1)}
        del rgba, ov, mats

    # the reference's own CPU case: 65,536-triangle soup, 256 x 256 camera rays,
    # RefMT closest hit on the GPU against the reference traversal on the CPU
    cfg0 = None
    if not args.no_cpu_reference_case:
        import oracle
        from realtimeraytracing_amd import HIT_DTYPE, RAY_DTYPE, RTQ_TRACE_LEAF_REFMT
        stris, smodels = scenes.soup_scene(65536, seed=0)
        ssc = ctx.upload(stris, smodels)
        sinfo = ssc.build(16, args.max_leaf, 0)
        scam = scenes.pinhole(**scenes.SOUP_CAMERA)
        sn = scam.width * scam.height
        srays, shits = devmem.empty(sn, RAY_DTYPE), devmem.empty(sn, HIT_DTYPE)
        ctx.raygen_pinhole(scam, srays.data_ptr(), sp)
        torch.cuda.synchronize()

        def soup_step(ev=None):
            if ev is not None:
                ev[0].record(stream)
            ssc.trace(srays.data_ptr(), sn, shits.data_ptr(), RTQ_TRACE_LEAF_REFMT, sp)
            if ev is not None:
                ev[1].record(stream)
                ev[2].record(stream)
        s_ms, _ = time_steps(soup_step)
        got = devmem.to_host(shits, HIT_DTYPE)
        host_rays = devmem.to_host(srays, RAY_DTYPE)
        cores = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
        cores = int(os.environ.get("OMP_NUM_THREADS", cores))
        t_c = time.perf_counter()
        want = oracle.trace_ref(ssc.export_ref(), oracle.world_vertices(stris, smodels), host_rays,
                                leaf=oracle.LEAF_REFMT, threads=cores)
        c_s = time.perf_counter() - t_c
        cfg0 = {"workload": "65,536-triangle soup (seed 0), 256x256 pinhole rays from (0.5, 0.5, -1), RefMT closest hit",
                "gpu_Mrays_s": round(sn / (s_ms * 1e-3) / 1e6, 1), "gpu_ms": round(s_ms, 4),
                "cpu_reference_Mrays_s": round(sn / c_s / 1e6, 2), "cpu_cores": cores,
                "exact_vs_reference_traversal": bool(np.array_equal(got.view(np.uint8), want.view(np.uint8))),
                "hit_fraction": round(float((got["id"] >= 0).mean()), 4), "build_ms": round(sinfo.build_ms, 3)}
        ssc.release()

    # build schedules: fused PLOC passes (default) against one launch per step
    build = None
    if not args.no_build_timing:
        times, outs = {}, {}
        for name, flags in (("split", RTQ_BUILD_PLOC_SPLIT), ("fused", 0)):
            ms = []
            for _ in range(3):
                binfo = wl.sc.build(16, args.max_leaf, flags, node_width=8)
                ms.append(binfo.build_ms)
            times[name] = min(ms)
            outs[name] = (wl.sc.export_clusters(), binfo.ploc_iterations)
        (cf, pf), (cs, ps) = outs["fused"], outs["split"]
        same = pf == ps and all(np.array_equal(cf[k], cs[k]) for k in ("triangle_id", "left", "right", "parent")) \
            and np.array_equal(cf["boxes"].view(np.uint8), cs["boxes"].view(np.uint8))
        build = {"workload": "terrain 1024x1024 (2,093,058 tris): bounds, Morton codes, sort, PLOC (radius 16), "
                             "pre-order placement, leaf-ordered triangles, binary and 8-wide traversal layouts "
                             "(the reference node array is written on export, outside this time); best of 3",
                 "fused_ms": round(times["fused"], 3), "split_ms": round(times["split"], 3),
                 "fused_vs_split": round(times["split"] / times["fused"], 3), "ploc_passes": pf,
                 "same_clusters": bool(same)}
        if not args.no_cpu_baseline:
            # the reference's build restated in C (bounds, Morton codes, pair sort,
            # PLOC with its nearest-neighbour search in parallel, pre-order flatten)
            import oracle
            cores = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
            cores = int(os.environ.get("OMP_NUM_THREADS", cores))
            oracle.build_ref_nodes(tris, models)          # warm-up
            cts = []
            for _ in range(3):
                t_b = time.perf_counter()
                _, cpasses = oracle.build_ref_nodes(tris, models)
                cts.append((time.perf_counter() - t_b) * 1e3)
            build["cpu_reference_build_ms"] = round(float(np.median(cts)), 1)
            build["cpu_reference_build"] = (f"orc_build_ref_nodes, {cores} threads, 1 warm-up + median of 3, "
                                            f"{cpasses} PLOC passes (same as the GPU: {cpasses == pf})")
            build["gpu_vs_cpu_build"] = round(build["cpu_reference_build_ms"] / times["fused"], 1)

    # scene edit: every terrain vertex moved (heights rippled by up to 0.02),
    # then a refit of the kept tree against a full rebuild
    refit = None
    if not args.no_refit:
        tris2 = tris.copy()
        for name in ("p0", "p1", "p2"):
            x = tris2[name][:, 0].astype(np.float64)
            z = tris2[name][:, 2].astype(np.float64)
            tris2[name][:, 1] = (tris2[name][:, 1] + 0.02 * np.sin(12.0 * x) * np.cos(9.0 * z)).astype(np.float32)
        torch.cuda.synchronize()
        t_u = time.perf_counter()
        wl.sc.update(tris2)
        update_ms = (time.perf_counter() - t_u) * 1e3
        # the same vertices from device memory (as an animation kernel would
        # leave them): best of 3, each followed by nothing else
        d_tris2 = devmem.to_device(tris2)
        torch.cuda.synchronize()
        update_dev_ms = []
        for _ in range(3):
            t_u = time.perf_counter()
            wl.sc.update_device(d_tris2.data_ptr(), len(tris2))
            update_dev_ms.append((time.perf_counter() - t_u) * 1e3)
        del d_tris2
        refit_ms = []
        for _ in range(3):
            t_r = time.perf_counter()
            wl.sc.refit()
            refit_ms.append((time.perf_counter() - t_r) * 1e3)
        ms_refit, _ = time_steps(two_launches(wl, RTQ_TRACE_BINARY))
        h_refit = wl.hits()
        rinfo = wl.sc.build(16, args.max_leaf, 0, node_width=8)
        ms_rebuild, _ = time_steps(two_launches(wl, RTQ_TRACE_BINARY))
        h_rebuild = wl.hits()
        refit = {"workload": "terrain vertices moved (y += 0.02 sin(12x) cos(9z)), plain binary diffuse + shadow",
                 "update_ms": round(update_ms, 2), "update_device_ms": round(min(update_dev_ms), 3),
                 "refit_ms": round(min(refit_ms), 3),
                 "rebuild_ms": round(rinfo.build_ms, 3),
                 "Mrays_s_after_refit": round(wl.live_rays / (ms_refit * 1e-3) / 1e6, 1),
                 "Mrays_s_after_rebuild": round(wl.live_rays / (ms_rebuild * 1e-3) / 1e6, 1),
                 "exact_refit_vs_rebuild": same_results(h_rebuild, h_refit)}
    wl.release()

    # learned-ordering workload: terrain + soup, plain against learned order
    occluded = None
    if not args.no_occluded:
        t3 = time.time()
        otris, omodels = scenes.occluded_scene(1024, 2_097_152, seed=0)
        ocam = scenes.pinhole(**{**scenes.TERRAIN_CAMERA, "width": args.occluded_res,
                                 "height": args.occluded_res})
        ctx.set_launch_params(default_params.refill, default_params.chunk, bool(default_params.prefer_wide),
                              default_params.spec_slack, default_params.leaf_slack,
                              default_params.stack_lds)
        owl = Workload(ctx, otris, omodels, ocam, sp, args.max_leaf, node_width=8)
        n_otris = len(otris)
        del otris
        oforms = {}
        oref = None
        for form in ("plain", "compressed", "learned_binary", "learned", "learned_batched"):
            flags = 0 if form == "compressed" else RTQ_TRACE_BINARY
            setup = None
            if form == "learned_binary":
                setup = train_order(ctx, owl, sp)
            if form.startswith("learned"):
                flags = RTQ_TRACE_LEARNED | (RTQ_TRACE_BINARY if form == "learned_binary" else 0)
            if form == "learned_batched":
                ctx.set_launch_params(default_params.refill, default_params.chunk, bool(default_params.prefer_wide),
                              default_params.spec_slack, default_params.leaf_slack,
                              default_params.stack_lds)
                items = ctx.items([(owl.sc, owl.diff.data_ptr(), owl.n_diff, owl.hd.data_ptr(), flags),
                                   (owl.sc, owl.shad.data_ptr(), owl.n_shad, owl.hs.data_ptr(),
                                    RTQ_TRACE_ANY | flags)])
                otr = ctx.autotune(items, reps=2)
                otuned = {"refill": otr.best.refill, "chunk": otr.best.chunk, "prefer_wide": otr.best.prefer_wide,
                          "spec_slack": otr.best.spec_slack, "leaf_slack": otr.best.leaf_slack,
                          "stack_lds": otr.best.stack_lds, "candidates": otr.n_tried}
                graph = ctx.batch(items)
                ms_per_step, evs = time_steps(graph_step(graph))
                graph.destroy()
            else:
                otuned = None
                ms_per_step, evs = time_steps(two_launches(owl, flags))
            got = owl.hits()
            if oref is None:
                oref = got
            c = owl.counters(flags, {"compressed": WIDE_NODE_BYTES, "learned": LEARNED_WIDE_NODE_BYTES}.get(
                form, BINARY_NODE_BYTES))
            split = form != "learned_batched"
            oforms[form] = {
                "Mrays_s": round(owl.live_rays / (ms_per_step * 1e-3) / 1e6, 2),
                "ms_per_step": round(ms_per_step, 4),
                "diffuse_ms": round(float(np.mean([e[0].elapsed_time(e[1]) for e in evs])), 4) if split else None,
                "shadow_ms": round(float(np.mean([e[1].elapsed_time(e[2]) for e in evs])), 4) if split else None,
                "shadow_node_visits": c["shadow"]["node_visits"], "shadow_tri_tests": c["shadow"]["tri_tests"],
                "occluded_fraction": round(float((got[1]["id"] >= 0).mean()), 4),
                "exact": same_results(oref, got), "order_model": setup, "tuned": otuned}
        occluded = {
            "workload": f"terrain 1024x1024 + 2,097,152 soup triangles ({n_otris:,} tris), "
                        f"{args.occluded_res}x{args.occluded_res} primary hits -> 1 diffuse + 4 shadow rays each",
            "rays_per_step": owl.live_rays, "forms": oforms,
            "learned_vs_plain": round(oforms["learned"]["Mrays_s"] / oforms["plain"]["Mrays_s"], 3),
            "learned_vs_compressed": round(oforms["learned"]["Mrays_s"] / oforms["compressed"]["Mrays_s"], 3),
            "learned_binary_vs_plain": round(oforms["learned_binary"]["Mrays_s"] / oforms["plain"]["Mrays_s"], 3),
            "compressed_vs_plain": round(oforms["compressed"]["Mrays_s"] / oforms["plain"]["Mrays_s"], 3),
            "learned_batched_vs_plain": round(oforms["learned_batched"]["Mrays_s"] / oforms["plain"]["Mrays_s"], 3),
            "shadow_speedup": round(oforms["plain"]["shadow_ms"] / oforms["learned"]["shadow_ms"], 3),
            "setup_s": round(time.time() - t3, 1)}
        owl.release()

    # config 3: four tiled terrains, compressed 8-wide nodes against binary
    compressed_8m = None
    if not args.no_compressed_8m:
        from realtimeraytracing_amd import HIT_DTYPE, RAY_DTYPE
        t4 = time.time()
        # an earlier block's autotune may have left binary nodes preferred
        ctx.set_launch_params(default_params.refill, default_params.chunk, bool(default_params.prefer_wide),
                              default_params.spec_slack, default_params.leaf_slack,
                              default_params.stack_lds)
        ttris = scenes.tiled_terrains(1024, 2, seed=0)
        tsc = ctx.upload(ttris, scenes.identity_models())
        tinfo = tsc.build(16, args.max_leaf, 0, node_width=8)
        tcam = scenes.pinhole(**scenes.TILED_CAMERA)
        n = tcam.width * tcam.height
        prim, diff = devmem.empty(n, RAY_DTYPE), devmem.empty(n, RAY_DTYPE)
        hit = {k: devmem.empty(n, HIT_DTYPE) for k in ("pb", "pw", "db", "dw")}
        ctx.raygen_pinhole(tcam, prim.data_ptr(), sp)
        tsc.trace(prim.data_ptr(), n, hit["pb"].data_ptr(), RTQ_TRACE_BINARY, sp)
        tsc.raygen_diffuse(prim.data_ptr(), hit["pb"].data_ptr(), n, scenes.DIFFUSE_SEED, scenes.SECONDARY_EPS,
                           diff.data_ptr(), sp)
        torch.cuda.synchronize()
        forms3 = {}
        for lname, layout in (("binary", RTQ_TRACE_BINARY), ("wide", 0)):
            hp, hd = hit["p" + lname[0]], hit["d" + lname[0]]

            def step(ev=None, layout=layout, hp=hp, hd=hd):
                if ev is not None:
                    ev[0].record(stream)
                tsc.trace(prim.data_ptr(), n, hp.data_ptr(), layout, sp)
                if ev is not None:
                    ev[1].record(stream)
                tsc.trace(diff.data_ptr(), n, hd.data_ptr(), layout, sp)
                if ev is not None:
                    ev[2].record(stream)
            ms_per_step, evs = time_steps(step)
            forms3[lname] = {"Mrays_s": round(2 * n / (ms_per_step * 1e-3) / 1e6, 2),
                             "ms_per_step": round(ms_per_step, 4),
                             "primary_ms": round(float(np.mean([e[0].elapsed_time(e[1]) for e in evs])), 4),
                             "diffuse_ms": round(float(np.mean([e[1].elapsed_time(e[2]) for e in evs])), 4)}
        torch.cuda.synchronize()
        exact = bool(torch.equal(hit["pb"], hit["pw"]) and torch.equal(hit["db"], hit["dw"]))
        prim_h, diff_h = devmem.to_host(prim, RAY_DTYPE), devmem.to_host(diff, RAY_DTYPE)
        compressed_8m = {
            "workload": f"4 tiled terrains 1024x1024 ({tinfo.n_tris:,} tris), {tcam.width}x{tcam.height} primary "
                        "rays + 1 diffuse ray per hit, closest hit",
            "rays_per_step": 2 * n, "forms": forms3, "ids_exact_wide_vs_binary": exact,
            "wide_vs_binary": round(forms3["wide"]["Mrays_s"] / forms3["binary"]["Mrays_s"], 3),
            "t_error_vs_fp64": {"primary": t_error_stats(ttris, prim_h, devmem.to_host(hit["pw"], HIT_DTYPE)),
                                "diffuse": t_error_stats(ttris, diff_h, devmem.to_host(hit["dw"], HIT_DTYPE))},
            "build_ms": round(tinfo.build_ms, 2), "trav_nodes": tinfo.n_trav_nodes, "wide_nodes": tinfo.n_wide_nodes,
            "setup_s": round(time.time() - t4, 1)}
        del ttris, prim_h, diff_h
        tsc.release()

    # config 2: 64 uneven scenes (soups and spheres), one graph-replayed launch
    # against one launch per scene
    uneven = None
    if not args.no_uneven_batch:
        from realtimeraytracing_amd import HIT_DTYPE, RAY_DTYPE
        t5 = time.time()
        ctx.set_launch_params(default_params.refill, default_params.chunk, bool(default_params.prefer_wide),
                              default_params.spec_slack, default_params.leaf_slack,
                              default_params.stack_lds)
        bsc, brays, bsep, bbat, bn = [], [], [], [], []
        n_tris_total = 0
        for s_ in range(64):
            btris, rays = scenes.batched_scene(s_, seed=0)
            sc_ = ctx.upload(btris, scenes.identity_models())
            sc_.build(16, args.max_leaf, 0, node_width=8)
            n_tris_total += len(btris)
            bsc.append(sc_)
            brays.append(devmem.to_device(rays))
            bn.append(len(rays))
            bsep.append(devmem.empty(len(rays), HIT_DTYPE))
            bbat.append(devmem.empty(len(rays), HIT_DTYPE))
            del btris, rays
        torch.cuda.synchronize()
        gen_s = time.time() - t5
        total_rays = sum(bn)

        def separate(ev=None):
            if ev is not None:
                ev[0].record(stream)
            for sc_, r_, h_, n_ in zip(bsc, brays, bsep, bn):
                sc_.trace(r_.data_ptr(), n_, h_.data_ptr(), 0, sp)
            if ev is not None:
                ev[1].record(stream)
                ev[2].record(stream)
        sep_ms, _ = time_steps(separate)
        items = ctx.items([(sc_, r_.data_ptr(), n_, h_.data_ptr(), 0) for sc_, r_, h_, n_ in zip(bsc, brays, bbat, bn)])
        tr = ctx.autotune(items, reps=2)
        graph = ctx.batch(items)
        bat_ms, _ = time_steps(graph_step(graph))
        graph.destroy()
        torch.cuda.synchronize()
        exact = all(torch.equal(a, b) for a, b in zip(bsep, bbat))
        uneven = {
            "workload": f"64 scenes, triangles log-uniform in [1e3, 4e6] ({n_tris_total:,} in all), soups and "
                        f"jittered spheres, {total_rays:,} random-direction rays (log-uniform 64k-2M per scene), "
                        "closest hit",
            "separate_launches": {"Mrays_s": round(total_rays / (sep_ms * 1e-3) / 1e6, 2), "ms": round(sep_ms, 4)},
            "batched_graph": {"Mrays_s": round(total_rays / (bat_ms * 1e-3) / 1e6, 2), "ms": round(bat_ms, 4),
                              "tuned": {"refill": tr.best.refill, "chunk": tr.best.chunk,
                                        "prefer_wide": tr.best.prefer_wide, "spec_slack": tr.best.spec_slack,
                                        "leaf_slack": tr.best.leaf_slack, "stack_lds": tr.best.stack_lds,
                                        "candidates": tr.n_tried}},
            "batched_vs_separate": round(sep_ms / bat_ms, 3), "exact": exact,
            "generation_s": round(gen_s, 1), "setup_s": round(time.time() - t5, 1)}
        for sc_ in bsc:
            sc_.release()
        del brays, bsep, bbat
        ctx.set_launch_params(default_params.refill, default_params.chunk, bool(default_params.prefer_wide),
                              default_params.spec_slack, default_params.leaf_slack,
                              default_params.stack_lds)

    def rounded(d):
        return {k: (round(v, 4) if isinstance(v, float) else v) for k, v in d.items()}

    out = {
        "metric": "Mrays/s (2M-tri terrain, diffuse+shadow) learned vs plain, hit IDs exact",
        "value": round(res["Mrays_s"], 2),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(res["ms_per_step"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "terrain 1024x1024 (2,093,058 tris), 1920x1080 primary hits -> "
                               "1 diffuse (closest) + 4 shadow (any) rays each",
                   "form": best, "rays_per_step": wl.live_rays, "max_leaf_size": args.max_leaf,
                   "search_radius": 16},
        "roofline": res["roofline"],
        "cpu_baseline": cpu,
        "detail": {"forms": {f: rounded({k: v for k, v in r.items() if k != "roofline"}) | {
                        "roofline_frac": r["roofline"]["frac"]} for f, r in results.items()},
                   "speedup_vs_plain": round(res["Mrays_s"] / results["plain"]["Mrays_s"], 3)
                   if "plain" in results else None,
                   "learned_vs_plain": round(results["learned"]["Mrays_s"] / results["plain"]["Mrays_s"], 3)
                   if "learned" in results and "plain" in results else None,
                   "order_model": learned_setup,
                   "occluded": occluded,
                   "compressed_8m": compressed_8m,
                   "uneven_batch": uneven,
                   "shading": shading,
                   "build": build,
                   "refit": refit,
                   "cpu_reference_case": cfg0,
                   "build_ms": round(info.build_ms, 2), "ploc_passes": info.ploc_iterations,
                   "depth": info.depth, "trav_nodes": info.n_trav_nodes, "wide_nodes": info.n_wide_nodes,
                   "wide_depth": info.wide_depth, "tree_sha256": wl.tree, "setup_s": round(setup_s, 1)},
    }
    print(json.dumps(out))
    ctx.close()
    if args.copy_calibration_mb > 0:
        nbytes = args.copy_calibration_mb << 20
        src = torch.ones(nbytes, dtype=torch.uint8, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes": nbytes}))
    return 0


if __name__ == "__main__":
    sys.exit(main())
