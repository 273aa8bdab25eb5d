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

Prints one JSON line with the metric, a roofline object for the dominant
trace kernel (algorithmic bytes per launch from the traversal counters over
its HIP-event duration) and the CPU baseline (the reference traversal
restated in C, run on this host's cores over the same rays).
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
NODE_BYTES = 64      # two-child traversal node
TRI_BYTES = 48       # leaf triangle record
RAY_BYTES = 32
HIT_BYTES = 16


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--traffic-file", default=str(ROOT / "profiles" / "traffic_terrain_shadow_anyhit.txt"))
    return ap.parse_args()


def read_traffic(path: str):
    """HBM bytes per launch of the dominant kernel, from the committed rocprof
    counter summary (line `traffic_bytes_per_launch: <value>`), else None."""
    p = Path(path)
    if not p.exists():
        return None
    for line in p.read_text().splitlines():
        if line.startswith("traffic_bytes_per_launch:"):
            return float(line.split(":", 1)[1])
    return None


def main() -> int:
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": f"--gpus {args.gpus}: this build runs on one GPU only; "
                                   "the trace path shards by ray ranges with a replicated BVH (see DESIGN.md)"}))
        return 2
    import torch
    from realtimeraytracing_amd import HIT_DTYPE, RAY_DTYPE, RTQ_TRACE_ANY, Context, devmem, scenes

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    torch.cuda.set_device(0)
    # an explicit stream: the library launches on it and the events below see
    # exactly those launches (a null stream would mean the library's own stream)
    stream = torch.cuda.Stream()
    torch.cuda.set_stream(stream)
    sp = stream.cuda_stream
    assert sp != 0

    t0 = time.time()
    tris, models = scenes.terrain_scene(1024, seed=0)
    ctx = Context(0)
    sc = ctx.upload(tris, models)
    info = sc.build(16, 4, 0)
    cam = scenes.pinhole(**scenes.TERRAIN_CAMERA)
    n = cam.width * cam.height
    d_prim, d_phit = devmem.empty(n, RAY_DTYPE), devmem.empty(n, HIT_DTYPE)
    ctx.raygen_pinhole(cam, d_prim.data_ptr(), sp)
    sc.trace(d_prim.data_ptr(), n, d_phit.data_ptr(), 0, sp)
    d_diff, d_shad = devmem.empty(n, RAY_DTYPE), devmem.empty(4 * n, RAY_DTYPE)
    sc.raygen_diffuse(d_prim.data_ptr(), d_phit.data_ptr(), n, scenes.DIFFUSE_SEED, scenes.SECONDARY_EPS,
                      d_diff.data_ptr(), sp)
    sc.raygen_shadow(d_prim.data_ptr(), d_phit.data_ptr(), n, scenes.LIGHTS, scenes.SECONDARY_EPS,
                     d_shad.data_ptr(), sp)
    d_hd, d_hs = devmem.empty(n, HIT_DTYPE), devmem.empty(4 * n, HIT_DTYPE)
    torch.cuda.synchronize()
    phit = devmem.to_host(d_phit, HIT_DTYPE)
    n_live = int((phit["id"] >= 0).sum())
    n_diff, n_shad = n, 4 * n
    live_rays = n_live + 4 * n_live
    setup_s = time.time() - t0

    # traversal counters (outside the timed region): algorithmic bytes per launch
    counters = {}
    for name, d_rays, cnt, d_hits, flags in (("diffuse", d_diff, n_diff, d_hd, 0),
                                              ("shadow", d_shad, n_shad, d_hs, RTQ_TRACE_ANY)):
        dv, dt = devmem.empty(cnt, np.uint32), devmem.empty(cnt, np.uint32)
        sc.trace_stats(d_rays.data_ptr(), cnt, d_hits.data_ptr(), dv.data_ptr(), dt.data_ptr(), flags, sp)
        torch.cuda.synchronize()
        v = devmem.to_host(dv, np.uint32).astype(np.int64)
        t = devmem.to_host(dt, np.uint32).astype(np.int64)
        counters[name] = dict(rays=cnt, node_visits=int(v.sum()), tri_tests=int(t.sum()),
                              bytes=cnt * (RAY_BYTES + HIT_BYTES) + int(v.sum()) * NODE_BYTES + int(t.sum()) * TRI_BYTES)

    def step(ev=None):
        if ev is not None:
            ev[0].record(stream)
        sc.trace(d_diff.data_ptr(), n_diff, d_hd.data_ptr(), 0, sp)
        if ev is not None:
            ev[1].record(stream)
        sc.trace(d_shad.data_ptr(), n_shad, d_hs.data_ptr(), RTQ_TRACE_ANY, sp)
        if ev is not None:
            ev[2].record(stream)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    evs = [[torch.cuda.Event(enable_timing=True) for _ in range(3)] for _ in range(args.steps)]
    start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    start.record(stream)
    for k in range(args.steps):
        step(evs[k])
    stop.record(stream)
    torch.cuda.synchronize()
    total_ms = start.elapsed_time(stop)
    diff_ms = float(np.mean([e[0].elapsed_time(e[1]) for e in evs]))
    shad_ms = float(np.mean([e[1].elapsed_time(e[2]) for e in evs]))
    ms_per_step = total_ms / args.steps
    mrays = live_rays / (ms_per_step * 1e-3) / 1e6

    dom = "shadow" if shad_ms >= diff_ms else "diffuse"
    dom_ms = shad_ms if dom == "shadow" else diff_ms
    achieved = counters[dom]["bytes"] / (dom_ms * 1e-3) / 1e9
    traffic = read_traffic(args.traffic_file)
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                "kernel": f"k_trace<watertight,{'any' if dom == 'shadow' else 'closest'}>",
                "launch_ms": round(dom_ms, 4), "algorithmic_bytes_per_launch": counters[dom]["bytes"]}

    # CPU baseline: the reference traversal (raytracer.glsl:246-295 restated in
    # C: 48-B pre-order nodes, RefMT leaf test, no pruning) on this host's cores
    cpu = None
    if not args.no_cpu_baseline:
        import oracle
        nodes = sc.export_ref()
        v9 = oracle.world_vertices(tris, models)
        diff = devmem.to_host(d_diff, RAY_DTYPE)
        shad = devmem.to_host(d_shad, RAY_DTYPE)
        cores = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
        cores = int(os.environ.get("OMP_NUM_THREADS", cores))
        t1 = time.perf_counter()
        oracle.trace_ref(nodes, v9, diff, leaf=oracle.LEAF_REFMT, threads=cores)
        oracle.trace_ref(nodes, v9, shad, leaf=oracle.LEAF_REFMT, threads=cores)
        cpu_s = time.perf_counter() - t1
        t2 = time.perf_counter()
        oracle.trace_bvh(nodes, v9, diff, threads=cores)
        oracle.trace_bvh(nodes, v9, shad, mode=oracle.MODE_ANY, threads=cores)
        cpu2_s = time.perf_counter() - t2
        cpu = {"value": round(live_rays / cpu_s / 1e6, 3), "unit": "Mrays/s", "cores": cores, "kind": "port",
               "sample": f"all {n_diff} diffuse + {n_shad} shadow rays of one step, reference traversal "
                         f"semantics (RefMT, no pruning), {cpu_s:.2f} s",
               "watertight_pruned_Mrays_s": round(live_rays / cpu2_s / 1e6, 3)}

    out = {
        "metric": "Mrays/s (2M-tri terrain, diffuse+shadow) learned vs plain, hit IDs exact",
        "value": round(mrays, 2),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "terrain 1024x1024 (2,093,058 tris), 1920x1080 primary hits -> "
                               "1 diffuse (closest) + 4 shadow (any) rays each",
                   "form": "plain", "rays_per_step": live_rays, "max_leaf_size": 4, "search_radius": 16},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "detail": {"diffuse_ms": round(diff_ms, 4), "shadow_ms": round(shad_ms, 4),
                   "diffuse_Mrays_s": round(n_live / (diff_ms * 1e-3) / 1e6, 1),
                   "shadow_Mrays_s": round(4 * n_live / (shad_ms * 1e-3) / 1e6, 1),
                   "build_ms": round(info.build_ms, 2), "ploc_passes": info.ploc_iterations,
                   "depth": info.depth, "trav_nodes": info.n_trav_nodes,
                   "counters": counters, "setup_s": round(setup_s, 1)},
    }
    print(json.dumps(out))
    sc.release()
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
