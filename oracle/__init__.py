"""ORACLE — test infrastructure only.

ctypes front end of oracle/build/liboracle.so, the CPU restatement of the
reference hot path (see oracle.h for what each function follows and how it is
pinned).  Only the tests, __graft_entry__.smoke() and the cpu_baseline leg of
bench.py use it, as the checker; the product library never does.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "build" / "liboracle.so"

LEAF_WATERTIGHT, LEAF_REFMT = 0, 1
MODE_CLOSEST, MODE_ANY = 0, 1

_lib = None


def _dtypes():
    from realtimeraytracing_amd.rtq import (AABB_DTYPE, HIT_DTYPE, MODEL_DTYPE, NODE_DTYPE, RAY_DTYPE,
                                            TRIANGLE_DTYPE)
    return TRIANGLE_DTYPE, MODEL_DTYPE, AABB_DTYPE, NODE_DTYPE, RAY_DTYPE, HIT_DTYPE


def build() -> Path:
    res = subprocess.run(["make", "-s", "-C", str(HERE)], capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"oracle build failed:\n{res.stdout}\n{res.stderr}")
    return LIB_PATH


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        L = ctypes.CDLL(str(LIB_PATH))
        vp, u32, u64, i32, f32 = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int, ctypes.c_float
        sigs = {
            "orc_expand_bits": (u32, [u32]),
            "orc_morton3d": (u32, [f32, f32, f32]),
            "orc_morton_codes": (None, [vp, u32, vp, i32, vp]),
            "orc_sort_pairs": (None, [vp, vp, u32]),
            "orc_ploc_build": (i32, [vp, u32, vp, i32, u32, vp, vp, vp, vp, vp, vp, vp, vp]),
            "orc_flatten_preorder": (i32, [u32, vp, vp, vp, vp, vp]),
            "orc_build_ref_nodes": (i32, [vp, u32, vp, i32, u32, vp]),
            "orc_tree_depth": (u32, [vp, u32]),
            "orc_world_vertices": (None, [vp, u32, vp, vp]),
            "orc_refmt": (i32, [vp, vp, f32, vp, vp, vp]),
            "orc_watertight": (i32, [vp, vp, f32, vp, vp, vp]),
            "orc_t_f64": (ctypes.c_double, [vp, vp]),
            "orc_trace_brute": (None, [vp, u32, vp, u64, i32, i32, vp, i32]),
            "orc_trace_ref": (None, [vp, vp, vp, u64, i32, vp, i32]),
            "orc_trace_bvh": (None, [vp, vp, vp, u64, i32, i32, f32, vp, i32, vp, vp]),
            "orc_gpu_box_check": (None, [vp, u32, vp, vp, u64, f32, f32, vp, vp]),
            "orc_rng_hash": (u64, [u64, u64, u64]),
            "orc_u01f": (f32, [u64, u64, u64]),
            "orc_raygen_pinhole": (None, [vp, vp, vp, vp, u32, u32, vp]),
            "orc_raygen_diffuse": (None, [vp, vp, vp, u64, u64, f32, vp]),
            "orc_raygen_shadow": (None, [vp, vp, vp, u64, vp, u32, f32, vp]),
            "orc_scan_exclusive": (None, [vp, vp, u64]),
            "orc_surface_area": (f32, [vp]),
            "orc_mlp_param_count": (u32, []),
            "orc_bf16_round": (f32, [f32]),
            "orc_mlp_logits": (None, [vp, vp, u64, u32, i32, vp]),
            "orc_mlp_loss": (ctypes.c_double, [vp, vp, u64]),
            "orc_mlp_grad": (ctypes.c_double, [vp, vp, u64, vp]),
            "orc_order_features": (None, [vp, vp, u32, u32, vp]),
            "orc_ref_box_hit": (i32, [vp, vp, f32]),
            "orc_refit_boxes": (None, [vp, vp, u32, vp, vp, vp, vp]),
            "orc_bvh_overlay": (None, [vp, u32, vp, u64, u32, i32, vp]),
            "orc_shade": (None, [vp, u32, vp, vp, u32, vp, u64, vp, u32, vp]),
        }
        for name, (res, args) in sigs.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _p(a):
    if a is None:
        return None
    if not a.flags["C_CONTIGUOUS"]:
        raise ValueError("oracle arrays must be C-contiguous")
    return a.ctypes.data


def _c(a, dtype=None):
    return None if a is None else np.ascontiguousarray(a, dtype=dtype)


def _threads(n):
    return int(n if n else os.environ.get("ORACLE_THREADS", os.cpu_count() or 1))


def expand_bits(v: int) -> int:
    return lib().orc_expand_bits(v)


def morton3d(x: float, y: float, z: float) -> int:
    return lib().orc_morton3d(x, y, z)


def morton_codes(tris, models, pad_origin=False) -> np.ndarray:
    out = np.zeros(len(tris), np.uint32)
    lib().orc_morton_codes(_p(tris), len(tris), _p(models), int(pad_origin), _p(out))
    return out


def sort_pairs(codes, idx):
    c = np.ascontiguousarray(codes, np.uint32).copy()
    i = np.ascontiguousarray(idx, np.uint32).copy()
    lib().orc_sort_pairs(_p(c), _p(i), len(c))
    return c, i


def ploc_build(tris, models, pad_origin=False, radius=16) -> dict:
    TRI, MOD, AABB, NODE, RAY, HIT = _dtypes()
    n = len(tris)
    nc = 2 * n - 1
    out = {"boxes": np.zeros(nc, AABB)}
    for k in ("triangle_id", "left", "right", "parent", "leaf_count"):
        out[k] = np.zeros(nc, np.uint32)
    out["sorted_indices"] = np.zeros(n, np.uint32)
    out["sorted_codes"] = np.zeros(n, np.uint32)
    passes = lib().orc_ploc_build(_p(tris), n, _p(models), int(pad_origin), radius, _p(out["boxes"]),
                                  _p(out["triangle_id"]), _p(out["left"]), _p(out["right"]), _p(out["parent"]),
                                  _p(out["leaf_count"]), _p(out["sorted_indices"]), _p(out["sorted_codes"]))
    out["passes"] = passes
    return out


def flatten_preorder(n, clusters) -> np.ndarray:
    TRI, MOD, AABB, NODE, RAY, HIT = _dtypes()
    out = np.zeros(2 * n - 1, NODE)
    r = lib().orc_flatten_preorder(n, _p(clusters["boxes"]), _p(clusters["triangle_id"]), _p(clusters["left"]),
                                   _p(clusters["right"]), _p(out))
    if r != 0:
        raise RuntimeError("flatten failed")
    return out


def build_ref_nodes(tris, models, pad_origin=False, radius=16):
    TRI, MOD, AABB, NODE, RAY, HIT = _dtypes()
    out = np.zeros(2 * len(tris) - 1, NODE)
    passes = lib().orc_build_ref_nodes(_p(tris), len(tris), _p(models), int(pad_origin), radius, _p(out))
    if passes < 0:
        raise RuntimeError("oracle PLOC failed")
    return out, passes


def refit_boxes(tris, models, clusters) -> np.ndarray:
    """Boxes of the cluster tree `clusters` (triangle_id, left, right) for the
    vertices of `tris`: what rtq_bvh_refit computes."""
    TRI, MOD, AABB, NODE, RAY, HIT = _dtypes()
    tris = _c(tris, TRI)
    models = _c(models, MOD)
    n = len(tris)
    out = np.zeros(2 * n - 1, AABB)
    tid = _c(clusters["triangle_id"], np.uint32)
    left = _c(clusters["left"], np.uint32)
    right = _c(clusters["right"], np.uint32)
    lib().orc_refit_boxes(_p(tris), _p(models), n, _p(tid), _p(left), _p(right), _p(out))
    return out


def tree_depth(nodes) -> int:
    return lib().orc_tree_depth(_p(nodes), len(nodes))


def world_vertices(tris, models) -> np.ndarray:
    out = np.zeros((len(tris), 9), np.float32)
    lib().orc_world_vertices(_p(tris), len(tris), _p(models), _p(out))
    return out


def leaf_test(ray, v9, leaf=LEAF_WATERTIGHT, tmax=None):
    TRI, MOD, AABB, NODE, RAY, HIT = _dtypes()
    r = np.ascontiguousarray(np.asarray(ray, RAY).reshape(1))
    v = np.ascontiguousarray(np.asarray(v9, np.float32).reshape(9))
    t, u, w = ctypes.c_float(), ctypes.c_float(), ctypes.c_float()
    fn = lib().orc_watertight if leaf == LEAF_WATERTIGHT else lib().orc_refmt
    tm = float(r["tmax"][0]) if tmax is None else tmax
    h = fn(_p(r), _p(v), tm, ctypes.byref(t), ctypes.byref(u), ctypes.byref(w))
    return (bool(h), t.value, u.value, w.value)


def t_f64(ray, v9) -> float:
    TRI, MOD, AABB, NODE, RAY, HIT = _dtypes()
    r = np.ascontiguousarray(np.asarray(ray, RAY).reshape(1))
    v = np.ascontiguousarray(np.asarray(v9, np.float32).reshape(9))
    return lib().orc_t_f64(_p(r), _p(v))


def trace_brute(verts9, rays, leaf=LEAF_WATERTIGHT, mode=MODE_CLOSEST, threads=0):
    TRI, MOD, AABB, NODE, RAY, HIT = _dtypes()
    hits = np.zeros(len(rays), HIT)
    lib().orc_trace_brute(_p(verts9), len(verts9), _p(rays), len(rays), leaf, mode, _p(hits), _threads(threads))
    return hits


def trace_ref(nodes, verts9, rays, leaf=LEAF_REFMT, threads=0):
    TRI, MOD, AABB, NODE, RAY, HIT = _dtypes()
    hits = np.zeros(len(rays), HIT)
    lib().orc_trace_ref(_p(nodes), _p(verts9), _p(rays), len(rays), leaf, _p(hits), _threads(threads))
    return hits


def trace_bvh(nodes, verts9, rays, leaf=LEAF_WATERTIGHT, mode=MODE_CLOSEST, pad=None, threads=0):
    TRI, MOD, AABB, NODE, RAY, HIT = _dtypes()
    if pad is None:
        pad = scene_pad(nodes)
    hits = np.zeros(len(rays), HIT)
    visits, tests = ctypes.c_uint64(), ctypes.c_uint64()
    lib().orc_trace_bvh(_p(nodes), _p(verts9), _p(rays), len(rays), leaf, mode, pad, _p(hits), _threads(threads),
                        ctypes.byref(visits), ctypes.byref(tests))
    return hits, visits.value, tests.value


def scene_pad(nodes) -> float:
    """The build's box pad: 2^-18 of the largest |coordinate| of the root box."""
    ext = max(float(np.abs(nodes[0]["box"]["min"]).max()), float(np.abs(nodes[0]["box"]["max"]).max()))
    return max(ext * 2.0 ** -18, 1e-30)


def gpu_box_check(nodes, rays, hits, rel, pad=None):
    """(rays whose hit a binary box test would skip, the same for 8-wide)."""
    if pad is None:
        pad = scene_pad(nodes)
    lb, lw = ctypes.c_uint64(), ctypes.c_uint64()
    lib().orc_gpu_box_check(_p(nodes), len(nodes), _p(rays), _p(hits), len(rays), pad, rel, ctypes.byref(lb),
                            ctypes.byref(lw))
    return lb.value, lw.value


def rng_hash(seed, stream, index) -> int:
    return lib().orc_rng_hash(seed, stream, index)


def u01f(seed, stream, index) -> float:
    return lib().orc_u01f(seed, stream, index)


def raygen_pinhole(eye, right, up, fwd, width, height):
    TRI, MOD, AABB, NODE, RAY, HIT = _dtypes()
    arrs = [np.ascontiguousarray(a, np.float32) for a in (eye, right, up, fwd)]
    rays = np.zeros(width * height, RAY)
    lib().orc_raygen_pinhole(*(_p(a) for a in arrs), width, height, _p(rays))
    return rays


def raygen_diffuse(verts9, prim, hits, seed, eps):
    TRI, MOD, AABB, NODE, RAY, HIT = _dtypes()
    out = np.zeros(len(prim), RAY)
    lib().orc_raygen_diffuse(_p(verts9), _p(prim), _p(hits), len(prim), seed, eps, _p(out))
    return out


def raygen_shadow(verts9, prim, hits, lights, eps):
    TRI, MOD, AABB, NODE, RAY, HIT = _dtypes()
    L = np.ascontiguousarray(lights, np.float32).reshape(-1)
    out = np.zeros(len(prim) * (len(L) // 3), RAY)
    lib().orc_raygen_shadow(_p(verts9), _p(prim), _p(hits), len(prim), _p(L), len(L) // 3, eps, _p(out))
    return out


def scan_exclusive(a) -> np.ndarray:
    a = np.ascontiguousarray(a, np.uint32)
    out = np.zeros_like(a)
    lib().orc_scan_exclusive(_p(a), _p(out), len(a))
    return out


# ---- learned child ordering (order.c) ----------------------------------------

MLP_INPUTS, MLP_HIDDEN, MLP_SAMPLE = 32, 64, 34


def mlp_param_count() -> int:
    return lib().orc_mlp_param_count()


def bf16_round(x: float) -> float:
    return lib().orc_bf16_round(x)


def mlp_logits(params, x, bf16=False) -> np.ndarray:
    """FP64 logits of the rows of x (n x >= 32 floats; the first 32 are used)."""
    P = np.ascontiguousarray(params, np.float32)
    X = np.ascontiguousarray(x, np.float32)
    out = np.zeros(len(X), np.float64)
    lib().orc_mlp_logits(_p(P), _p(X), len(X), X.shape[1], int(bf16), _p(out))
    return out


def mlp_loss(params, samples) -> float:
    P = np.ascontiguousarray(params, np.float32)
    S = np.ascontiguousarray(samples, np.float32)
    assert S.shape[1] == MLP_SAMPLE
    return lib().orc_mlp_loss(_p(P), _p(S), len(S))


def mlp_grad(params, samples):
    """(mean weighted BCE, FP64 gradient)."""
    P = np.ascontiguousarray(params, np.float32)
    S = np.ascontiguousarray(samples, np.float32)
    assert S.shape[1] == MLP_SAMPLE
    g = np.zeros(mlp_param_count(), np.float64)
    loss = lib().orc_mlp_grad(_p(P), _p(S), len(S), _p(g))
    return loss, g


def order_features(node16, aux4, octant, any_hit) -> np.ndarray:
    nd = np.ascontiguousarray(node16, np.float32).reshape(16)
    ax = np.ascontiguousarray(aux4, np.uint32).reshape(4)
    f = np.zeros(32, np.float32)
    lib().orc_order_features(_p(nd), _p(ax), int(octant), int(any_hit), _p(f))
    return f


# ---- shading (shade.c) ----------------------------------------------------------

def ref_box_hit(ray, box, threshold: float) -> int:
    """intersectBVH (raytracer.glsl:182-237): 0 miss, 1 hit, 2 hit near two slab planes."""
    TRI, MOD, AABB, NODE, RAY, HIT = _dtypes()
    r = _c(ray, RAY)
    b = _c(box, AABB)
    return lib().orc_ref_box_hit(_p(r), _p(b), threshold)


def bvh_overlay(nodes, rays, depth: int, cut: bool = False) -> np.ndarray:
    """bvhColor per ray (raytracer.glsl:246-276), float32 [n, 4]."""
    TRI, MOD, AABB, NODE, RAY, HIT = _dtypes()
    nodes = _c(nodes, NODE)
    rays = _c(rays, RAY)
    out = np.zeros((len(rays), 4), np.float32)
    lib().orc_bvh_overlay(_p(nodes), len(nodes), _p(rays), len(rays), depth, int(cut), _p(out))
    return out


def shade(tris, models, materials, hits, overlay=None, wireframe=False) -> np.ndarray:
    """Pixel colour per hit record (raytracer.glsl:159-179, 299-331), float32 [n, 4]."""
    TRI, MOD, AABB, NODE, RAY, HIT = _dtypes()
    tris = _c(tris, TRI)
    models = _c(models, MOD)
    mats = _c(np.asarray(materials, np.float32).reshape(-1, 4))
    hits = _c(hits, HIT)
    ov = None if overlay is None else _c(np.asarray(overlay, np.float32).reshape(-1, 4))
    out = np.zeros((len(hits), 4), np.float32)
    lib().orc_shade(_p(tris), len(tris), _p(models), _p(mats), len(mats), _p(hits), len(hits), _p(ov),
                    1 if wireframe else 0, _p(out))
    return out
