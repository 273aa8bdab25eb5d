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
between them, checked in tests/test_gpu_wide.py and tests/test_gpu_batch.py):
  plain       binary 64-B nodes, one launch per ray set, default settings
  compressed  8-wide 128-B nodes with 8-bit outward-rounded child boxes
  batched     both ray sets in one persistent launch replayed from a
              hipGraph, settings (refill, chunk, node layout) from the
              autotuner
`value` is the fastest form; `detail.forms` has every form.

Prints one JSON line with the metric, a roofline object for the dominant
trace kernel of that form (algorithmic bytes per launch from the traversal
counters over its HIP-event duration) and the CPU baseline (the reference
traversal restated in C, run on this host's cores over the same rays).
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
TRI_BYTES = 48       # leaf triangle record
RAY_BYTES = 32
HIT_BYTES = 16


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--forms", default="plain,compressed,batched")
    ap.add_argument("--max-leaf", type=int, default=4, help="triangles per traversal leaf (1..8)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--traffic-file", default=str(ROOT / "profiles" / "traffic_terrain_trace.txt"))
    ap.add_argument("--copy-calibration-mb", type=int, default=0,
                    help="after the run, copy a buffer of this many MiB device-to-device 3 times: a kernel of "
                         "known traffic to scale rocprof FETCH_SIZE / WRITE_SIZE readings")
    return ap.parse_args()


def read_traffic(path: str, kernel: str):
    """HBM bytes per launch of a kernel from the committed rocprof counter
    summary (lines `traffic_bytes_per_launch <kernel>: <value>`), else None."""
    p = Path(path)
    if not p.exists():
        return None
    for line in p.read_text().splitlines():
        if line.startswith(f"traffic_bytes_per_launch {kernel}:"):
            return float(line.split(":", 1)[1])
    return None


def main() -> int:
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": f"--gpus {args.gpus}: this build runs on one GPU only; "
                                   "the trace path shards by ray ranges with a replicated BVH (see DESIGN.md)"}))
        return 2
    import torch
    from realtimeraytracing_amd import (HIT_DTYPE, RAY_DTYPE, RTQ_TRACE_ANY, RTQ_TRACE_BINARY, Context, devmem,
                                        scenes)

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
    info = sc.build(16, args.max_leaf, 0, node_width=8)     # builds the binary and the 8-wide layouts
    cam = scenes.pinhole(**scenes.TERRAIN_CAMERA)
    n = cam.width * cam.height
    d_prim, d_phit = devmem.empty(n, RAY_DTYPE), devmem.empty(n, HIT_DTYPE)
    ctx.raygen_pinhole(cam, d_prim.data_ptr(), sp)
    sc.trace(d_prim.data_ptr(), n, d_phit.data_ptr(), RTQ_TRACE_BINARY, sp)
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

    # traversal counters of each layout (outside the timed region): algorithmic bytes per launch
    def layout_counters(layout, node_bytes):
        counters = {}
        for name, d_rays, cnt, d_hits, flags in (("diffuse", d_diff, n_diff, d_hd, 0),
                                                  ("shadow", d_shad, n_shad, d_hs, RTQ_TRACE_ANY)):
            dv, dt = devmem.empty(cnt, np.uint32), devmem.empty(cnt, np.uint32)
            sc.trace_stats(d_rays.data_ptr(), cnt, d_hits.data_ptr(), dv.data_ptr(), dt.data_ptr(),
                           flags | layout, sp)
            torch.cuda.synchronize()
            v = int(devmem.to_host(dv, np.uint32).astype(np.int64).sum())
            t = int(devmem.to_host(dt, np.uint32).astype(np.int64).sum())
            counters[name] = dict(rays=cnt, node_visits=v, tri_tests=t,
                                  bytes=cnt * (RAY_BYTES + HIT_BYTES) + v * node_bytes + t * TRI_BYTES)
        return counters

    default_params = ctx.launch_params()
    counters_by_layout = {"binary": layout_counters(RTQ_TRACE_BINARY, 64), "wide": layout_counters(0, 128)}

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

    results = {}
    for form in args.forms.split(","):
        ctx.set_launch_params(default_params.refill, default_params.chunk, bool(default_params.prefer_wide))
        tuned = None
        if form in ("plain", "compressed"):
            layout = RTQ_TRACE_BINARY if form == "plain" else 0
            lname = "binary" if form == "plain" else "wide"

            def step(ev=None):
                if ev is not None:
                    ev[0].record(stream)
                sc.trace(d_diff.data_ptr(), n_diff, d_hd.data_ptr(), layout, sp)
                if ev is not None:
                    ev[1].record(stream)
                sc.trace(d_shad.data_ptr(), n_shad, d_hs.data_ptr(), RTQ_TRACE_ANY | layout, sp)
                if ev is not None:
                    ev[2].record(stream)

            ms_per_step, evs = time_steps(step)
            diff_ms = float(np.mean([e[0].elapsed_time(e[1]) for e in evs]))
            shad_ms = float(np.mean([e[1].elapsed_time(e[2]) for e in evs]))
            counters = counters_by_layout[lname]
            dom = "shadow" if shad_ms >= diff_ms else "diffuse"
            dom_ms = shad_ms if dom == "shadow" else diff_ms
            dom_bytes = counters[dom]["bytes"]
            kernel = f"k_trace<watertight,{'any' if dom == 'shadow' else 'closest'},{lname}>"
        elif form == "batched":
            # both ray sets in one persistent launch, replayed from a hipGraph,
            # launch settings (incl. node layout) from the autotuner
            items = ctx.items([(sc, d_diff.data_ptr(), n_diff, d_hd.data_ptr(), 0),
                               (sc, d_shad.data_ptr(), n_shad, d_hs.data_ptr(), RTQ_TRACE_ANY)])
            tr = ctx.autotune(items, reps=2)
            tuned = {"refill": tr.best.refill, "chunk": tr.best.chunk, "prefer_wide": tr.best.prefer_wide,
                     "candidates": tr.n_tried, "best_ms": round(tr.best_ms, 4), "default_ms": round(tr.default_ms, 4)}
            lname = "wide" if tr.best.prefer_wide else "binary"
            graph = ctx.batch(items)

            def step(ev=None):
                if ev is not None:
                    ev[0].record(stream)
                graph.launch(sp)
                if ev is not None:
                    ev[1].record(stream)
                    ev[2].record(stream)

            ms_per_step, evs = time_steps(step)
            graph.destroy()
            diff_ms = shad_ms = None
            counters = counters_by_layout[lname]
            dom = "batch"
            dom_ms = float(np.mean([e[0].elapsed_time(e[1]) for e in evs]))
            dom_bytes = counters["diffuse"]["bytes"] + counters["shadow"]["bytes"]
            kernel = f"k_trace<watertight,per-item,{lname},batch>"
        else:
            raise SystemExit(f"unknown form {form}")
        achieved = dom_bytes / (dom_ms * 1e-3) / 1e9
        node_bytes = 64 if lname == "binary" else 128
        results[form] = {
            "Mrays_s": live_rays / (ms_per_step * 1e-3) / 1e6,
            "ms_per_step": ms_per_step, "diffuse_ms": diff_ms, "shadow_ms": shad_ms,
            "layout": lname, "tuned": tuned, "counters": counters,
            "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": read_traffic(args.traffic_file, kernel),
                         "kernel": kernel, "launch_ms": round(dom_ms, 4),
                         "algorithmic_bytes_per_launch": dom_bytes,
                         "bytes_per_ray": "32 ray + 16 hit + %d x node visits + 48 x triangle tests" % node_bytes},
        }
    ctx.set_launch_params(default_params.refill, default_params.chunk, bool(default_params.prefer_wide))

    best = max(results, key=lambda f: results[f]["Mrays_s"])
    res = results[best]

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
                   "form": best, "rays_per_step": live_rays, "max_leaf_size": args.max_leaf, "search_radius": 16},
        "roofline": res["roofline"],
        "cpu_baseline": cpu,
        "detail": {"forms": {f: rounded({k: v for k, v in r.items() if k != "roofline"}) | {
                        "roofline_frac": r["roofline"]["frac"]} for f, r in results.items()},
                   "speedup_vs_plain": round(res["Mrays_s"] / results["plain"]["Mrays_s"], 3)
                   if "plain" in results else None,
                   "build_ms": round(info.build_ms, 2), "ploc_passes": info.ploc_iterations,
                   "depth": info.depth, "trav_nodes": info.n_trav_nodes, "wide_nodes": info.n_wide_nodes,
                   "wide_depth": info.wide_depth, "setup_s": round(setup_s, 1)},
    }
    print(json.dumps(out))
    sc.release()
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
