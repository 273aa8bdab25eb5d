"""ctypes bindings of librtq (include/rtq/rtq.h).

Python is the test and benchmark harness; the product is the C ABI.  Loading
the library never touches the GPU; the first rtq_create does.  If librtq.so
is missing this module raises: there is no CPU fallback for any entry point.
"""
from __future__ import annotations

import ctypes
import os
from pathlib import Path

import numpy as np

PKG = Path(__file__).resolve().parent
LIB_PATH = PKG / "librtq.so"

# ---- record layouts (include/rtq/scene_types.h) -------------------------
TRIANGLE_DTYPE = np.dtype({"names": ["p0", "p1", "p2", "model_id", "pad"],
                           "formats": [("<f4", 4), ("<f4", 4), ("<f4", 4), "<u4", ("<u4", 3)],
                           "offsets": [0, 16, 32, 48, 52], "itemsize": 64})
MODEL_DTYPE = np.dtype({"names": ["m", "material_id"], "formats": [("<f4", 16), "<u4"],
                        "offsets": [0, 64], "itemsize": 68})
AABB_DTYPE = np.dtype({"names": ["min", "pad0", "max", "pad1"],
                       "formats": [("<f4", 3), "<f4", ("<f4", 3), "<f4"],
                       "offsets": [0, 12, 16, 28], "itemsize": 32})
NODE_DTYPE = np.dtype({"names": ["box", "triangle_id", "left_child", "right_child", "pad"],
                       "formats": [AABB_DTYPE, "<u4", "<u4", "<u4", "<u4"],
                       "offsets": [0, 32, 36, 40, 44], "itemsize": 48})
RAY_DTYPE = np.dtype({"names": ["o", "tmin", "d", "tmax"], "formats": [("<f4", 3), "<f4", ("<f4", 3), "<f4"],
                      "offsets": [0, 12, 16, 28], "itemsize": 32})
HIT_DTYPE = np.dtype({"names": ["id", "t", "u", "v"], "formats": ["<i4", "<f4", "<f4", "<f4"],
                      "offsets": [0, 4, 8, 12], "itemsize": 16})

RTQ_OK, RTQ_EINVAL, RTQ_ENOMEM, RTQ_EHIP, RTQ_ECAPACITY, RTQ_EINTERNAL, RTQ_ENOTBUILT = range(7)
RTQ_BUILD_REF_PADDED_BOUNDS = 0x1
RTQ_BUILD_PLOC_SPLIT = 0x2
RTQ_TRACE_CLOSEST = 0x0
RTQ_TRACE_ANY = 0x1
RTQ_TRACE_LEAF_REFMT = 0x2
RTQ_TRACE_BINARY = 0x4
RTQ_TRACE_LEARNED = 0x8
RTQ_SHADE_WIREFRAME = 0x1
MATERIAL_DTYPE = np.dtype({"names": ["color"], "formats": [("<f4", 4)], "offsets": [0], "itemsize": 16})
ORDER_INPUTS, ORDER_HIDDEN, ORDER_PARAMS, ORDER_SAMPLE = 32, 64, 6337, 34

EXPORTED = [
    "rtq_abi_version", "rtq_status_string", "rtq_create", "rtq_destroy", "rtq_last_error",
    "rtq_synchronize", "rtq_scene_upload", "rtq_scene_release", "rtq_build_params_default",
    "rtq_bvh_build", "rtq_bvh_info_get", "rtq_bvh_export_ref", "rtq_bvh_export_clusters",
    "rtq_morton_codes", "rtq_sort_pairs", "rtq_scan_exclusive_u32", "rtq_trace", "rtq_trace_stats",
    "rtq_trace_host", "rtq_raygen_pinhole", "rtq_raygen_diffuse", "rtq_raygen_shadow",
    "rtq_trace_batch", "rtq_batch_create", "rtq_batch_launch", "rtq_batch_destroy",
    "rtq_launch_params_get", "rtq_launch_params_set", "rtq_autotune", "rtq_tune_save", "rtq_tune_load",
    "rtq_bvh_export_trav", "rtq_order_log", "rtq_order_init", "rtq_order_grad", "rtq_order_train",
    "rtq_order_predict", "rtq_order_apply", "rtq_order_save", "rtq_order_load",
    "rtq_order_model_create", "rtq_order_model_destroy", "rtq_order_model_log", "rtq_order_model_clear_samples",
    "rtq_order_model_train", "rtq_order_model_apply", "rtq_order_model_save", "rtq_order_model_load",
    "rtq_order_model_info", "rtq_shade", "rtq_bvh_overlay", "rtq_obj_load", "rtq_obj_parse",
    "rtq_scene_update", "rtq_bvh_refit", "rtq_trace_noalloc", "rtq_bvh_export_wide", "rtq_scene_update_device",
]
MAX_BATCH_ITEMS = 64


class BatchItem(ctypes.Structure):
    _fields_ = [("scene_id", ctypes.c_uint32), ("flags", ctypes.c_uint32), ("rays_dev", ctypes.c_void_p),
                ("hits_dev", ctypes.c_void_p), ("n_rays", ctypes.c_uint64)]


class LaunchParams(ctypes.Structure):
    _fields_ = [("refill", ctypes.c_uint32), ("chunk", ctypes.c_uint32), ("prefer_wide", ctypes.c_uint32),
                ("spec_slack", ctypes.c_uint32), ("leaf_slack", ctypes.c_uint32), ("stack_lds", ctypes.c_uint32)]


class TuneResult(ctypes.Structure):
    _fields_ = [("best", LaunchParams), ("best_ms", ctypes.c_float), ("default_ms", ctypes.c_float),
                ("n_tried", ctypes.c_uint32), ("from_cache", ctypes.c_uint32)]


class BuildParams(ctypes.Structure):
    _fields_ = [("search_radius", ctypes.c_uint32), ("max_leaf_size", ctypes.c_uint32),
                ("flags", ctypes.c_uint32), ("node_width", ctypes.c_uint32)]


class BvhInfo(ctypes.Structure):
    _fields_ = [("n_tris", ctypes.c_uint32), ("n_ref_nodes", ctypes.c_uint32),
                ("n_trav_nodes", ctypes.c_uint32), ("depth", ctypes.c_uint32),
                ("ploc_iterations", ctypes.c_uint32), ("max_leaf_size", ctypes.c_uint32),
                ("box_pad", ctypes.c_float), ("build_ms", ctypes.c_float),
                ("node_width", ctypes.c_uint32), ("n_wide_nodes", ctypes.c_uint32),
                ("wide_depth", ctypes.c_uint32), ("order_modes", ctypes.c_uint32),
                ("scratch_spill_bytes", ctypes.c_uint64)]


class TrainParams(ctypes.Structure):
    _fields_ = [("epochs", ctypes.c_uint32), ("batch", ctypes.c_uint32), ("lr", ctypes.c_float),
                ("beta1", ctypes.c_float), ("beta2", ctypes.c_float), ("reserved", ctypes.c_uint32),
                ("seed", ctypes.c_uint64)]


class Pinhole(ctypes.Structure):
    _fields_ = [("eye", ctypes.c_float * 3), ("right", ctypes.c_float * 3), ("up", ctypes.c_float * 3),
                ("forward", ctypes.c_float * 3), ("width", ctypes.c_uint32), ("height", ctypes.c_uint32)]


class RtqError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"rtq status {status}: {message}")
        self.status = status


_lib: ctypes.CDLL | None = None


def load_library(path: str | os.PathLike | None = None) -> ctypes.CDLL:
    """Load librtq.so; raises if it has not been built (no fallback)."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = Path(path) if path else LIB_PATH
    if not p.exists():
        raise RuntimeError(f"{p} not found: build it with `python -c 'import __graft_entry__ as g; g.build()'`")
    lib = ctypes.CDLL(str(p))
    vp, u32, u64, i32 = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int
    st = ctypes.c_int
    sigs = {
        "rtq_abi_version": (i32, []),
        "rtq_status_string": (ctypes.c_char_p, [st]),
        "rtq_create": (st, [ctypes.POINTER(vp), i32]),
        "rtq_destroy": (None, [vp]),
        "rtq_last_error": (ctypes.c_char_p, [vp]),
        "rtq_synchronize": (st, [vp]),
        "rtq_scene_upload": (st, [vp, vp, u32, vp, u32, ctypes.POINTER(u32)]),
        "rtq_scene_release": (st, [vp, u32]),
        "rtq_scene_update": (st, [vp, u32, vp, u32, vp, u32]),
        "rtq_scene_update_device": (st, [vp, u32, vp, u32]),
        "rtq_bvh_refit": (st, [vp, u32]),
        "rtq_build_params_default": (None, [ctypes.POINTER(BuildParams)]),
        "rtq_bvh_build": (st, [vp, u32, ctypes.POINTER(BuildParams)]),
        "rtq_bvh_info_get": (st, [vp, u32, ctypes.POINTER(BvhInfo)]),
        "rtq_bvh_export_ref": (st, [vp, u32, vp, u64, ctypes.POINTER(u64)]),
        "rtq_bvh_export_wide": (st, [vp, u32, vp, u64, ctypes.POINTER(u64)]),
        "rtq_bvh_export_clusters": (st, [vp, u32, u64, vp, vp, vp, vp, vp, vp, vp]),
        "rtq_morton_codes": (st, [vp, u32, u32, vp]),
        "rtq_sort_pairs": (st, [vp, vp, vp, u64, u32, vp]),
        "rtq_scan_exclusive_u32": (st, [vp, vp, vp, u64, vp]),
        "rtq_trace": (st, [vp, u32, vp, u64, u32, vp, vp]),
        "rtq_trace_noalloc": (st, [vp, u32, vp, u64, u32, vp, vp]),
        "rtq_trace_stats": (st, [vp, u32, vp, u64, u32, vp, vp, vp, vp]),
        "rtq_trace_host": (st, [vp, u32, vp, u64, u32, vp]),
        "rtq_raygen_pinhole": (st, [vp, ctypes.POINTER(Pinhole), vp, vp]),
        "rtq_raygen_diffuse": (st, [vp, u32, vp, vp, u64, u64, ctypes.c_float, vp, vp]),
        "rtq_raygen_shadow": (st, [vp, u32, vp, vp, u64, vp, u32, ctypes.c_float, vp, vp]),
        "rtq_trace_batch": (st, [vp, ctypes.POINTER(BatchItem), u32, vp]),
        "rtq_batch_create": (st, [vp, ctypes.POINTER(BatchItem), u32, ctypes.POINTER(vp)]),
        "rtq_batch_launch": (st, [vp, vp]),
        "rtq_batch_destroy": (None, [vp]),
        "rtq_launch_params_get": (st, [vp, ctypes.POINTER(LaunchParams)]),
        "rtq_launch_params_set": (st, [vp, ctypes.POINTER(LaunchParams)]),
        "rtq_autotune": (st, [vp, ctypes.POINTER(BatchItem), u32, u32, ctypes.POINTER(TuneResult)]),
        "rtq_tune_save": (st, [vp, ctypes.c_char_p]),
        "rtq_tune_load": (st, [vp, ctypes.c_char_p]),
        "rtq_bvh_export_trav": (st, [vp, u32, vp, vp, u64, ctypes.POINTER(u64)]),
        "rtq_order_log": (st, [vp, u32, vp, u64, u32, u32, vp, u64, ctypes.POINTER(u64)]),
        "rtq_order_init": (st, [vp, u64, vp]),
        "rtq_order_grad": (st, [vp, vp, vp, u32, vp, ctypes.POINTER(ctypes.c_float)]),
        "rtq_order_train": (st, [vp, vp, u64, ctypes.POINTER(TrainParams), vp, vp, vp, ctypes.POINTER(u32), vp]),
        "rtq_order_predict": (st, [vp, vp, vp, u64, u32, vp]),
        "rtq_order_apply": (st, [vp, u32, vp, ctypes.c_float, ctypes.c_float]),
        "rtq_order_save": (st, [vp, ctypes.c_char_p, vp, vp, vp, u32]),
        "rtq_order_load": (st, [vp, ctypes.c_char_p, vp, vp, vp, ctypes.POINTER(u32)]),
        "rtq_order_model_create": (st, [vp, u64, u64, ctypes.POINTER(vp)]),
        "rtq_order_model_destroy": (None, [vp]),
        "rtq_order_model_log": (st, [vp, u32, vp, u64, u32, u32, ctypes.POINTER(u64)]),
        "rtq_order_model_clear_samples": (st, [vp]),
        "rtq_order_model_train": (st, [vp, ctypes.POINTER(TrainParams), vp]),
        "rtq_order_model_apply": (st, [vp, u32, ctypes.c_float, ctypes.c_float]),
        "rtq_order_model_save": (st, [vp, ctypes.c_char_p]),
        "rtq_order_model_load": (st, [vp, ctypes.c_char_p]),
        "rtq_order_model_info": (st, [vp, ctypes.POINTER(u64), ctypes.POINTER(u32)]),
        "rtq_shade": (st, [vp, u32, vp, u64, vp, u32, vp, u32, vp, vp]),
        "rtq_obj_load": (st, [ctypes.c_char_p, u32, vp, u64, ctypes.POINTER(u64), ctypes.c_char_p, u64]),
        "rtq_obj_parse": (st, [ctypes.c_char_p, u64, u32, vp, u64, ctypes.POINTER(u64), ctypes.c_char_p, u64]),
        "rtq_bvh_overlay": (st, [vp, u32, vp, u64, u32, vp, vp]),
    }
    for name, (res, args) in sigs.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _lib = lib
    return lib


def _ptr(a: np.ndarray | None) -> int | None:
    return None if a is None else a.ctypes.data


def _obj_call(fn, head: tuple, model_id: int) -> np.ndarray:
    n = ctypes.c_uint64()
    msg = ctypes.create_string_buffer(256)
    rc = fn(*head, model_id, None, 0, ctypes.byref(n), msg, len(msg))
    if rc != RTQ_OK:
        raise RtqError(rc, msg.value.decode(errors="replace"))
    out = np.zeros(n.value, TRIANGLE_DTYPE)
    rc = fn(*head, model_id, _ptr(out), n.value, ctypes.byref(n), msg, len(msg))
    if rc != RTQ_OK:
        raise RtqError(rc, msg.value.decode(errors="replace"))
    return out


def obj_load(path, model_id: int = 0) -> np.ndarray:
    """Triangles of a Wavefront OBJ file (rtq_obj_load: Mesh::load, mesh.cpp:186-263)."""
    return _obj_call(load_library().rtq_obj_load, (str(path).encode(),), model_id)


def obj_parse(text, model_id: int = 0) -> np.ndarray:
    """Triangles of OBJ text in memory (rtq_obj_parse)."""
    data = text.encode() if isinstance(text, str) else bytes(text)
    return _obj_call(load_library().rtq_obj_parse, (data, len(data)), model_id)


class Context:
    """One device + stream (rtq_ctx)."""

    def __init__(self, device: int = 0):
        self.lib = load_library()
        h = ctypes.c_void_p()
        s = self.lib.rtq_create(ctypes.byref(h), device)
        if s != RTQ_OK:
            raise RtqError(s, f"rtq_create(device={device}) failed: "
                              f"{self.lib.rtq_status_string(s).decode()}")
        self.h = h

    def check(self, s: int) -> None:
        if s != RTQ_OK:
            raise RtqError(s, self.lib.rtq_last_error(self.h).decode())

    def close(self) -> None:
        if self.h:
            self.lib.rtq_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def synchronize(self) -> None:
        self.check(self.lib.rtq_synchronize(self.h))

    def upload(self, tris: np.ndarray, models: np.ndarray) -> "Scene":
        tris = np.ascontiguousarray(tris, dtype=TRIANGLE_DTYPE)
        models = np.ascontiguousarray(models, dtype=MODEL_DTYPE)
        sid = ctypes.c_uint32()
        self.check(self.lib.rtq_scene_upload(self.h, _ptr(tris), len(tris), _ptr(models), len(models),
                                             ctypes.byref(sid)))
        return Scene(self, sid.value, len(tris))

    # device-pointer primitives
    def sort_pairs(self, keys_ptr: int, vals_ptr: int, n: int, key_bits: int = 32, stream: int | None = None):
        self.check(self.lib.rtq_sort_pairs(self.h, keys_ptr, vals_ptr, n, key_bits, stream))

    def scan_exclusive(self, in_ptr: int, out_ptr: int, n: int, stream: int | None = None):
        self.check(self.lib.rtq_scan_exclusive_u32(self.h, in_ptr, out_ptr, n, stream))

    def raygen_pinhole(self, cam: Pinhole, rays_ptr: int, stream: int | None = None):
        self.check(self.lib.rtq_raygen_pinhole(self.h, ctypes.byref(cam), rays_ptr, stream))

    # batched traversal
    @staticmethod
    def items(specs) -> "ctypes.Array":
        """specs: iterable of (scene, rays_ptr, n_rays, hits_ptr, flags)."""
        specs = list(specs)
        arr = (BatchItem * len(specs))()
        for i, (scene, rays_ptr, n, hits_ptr, flags) in enumerate(specs):
            arr[i] = BatchItem(scene.id, flags, rays_ptr, hits_ptr, n)
        return arr

    def trace_batch(self, items, stream: int | None = None) -> None:
        self.check(self.lib.rtq_trace_batch(self.h, items, len(items), stream))

    def batch(self, items) -> "Batch":
        h = ctypes.c_void_p()
        self.check(self.lib.rtq_batch_create(self.h, items, len(items), ctypes.byref(h)))
        return Batch(self, h, items)

    def launch_params(self) -> LaunchParams:
        p = LaunchParams()
        self.check(self.lib.rtq_launch_params_get(self.h, ctypes.byref(p)))
        return p

    def set_launch_params(self, refill: int, chunk: int, prefer_wide: bool = True, spec_slack: int = 16,
                          leaf_slack: int = 8, stack_lds: int = 0) -> None:
        """stack_lds: LDS stack entries of batched traces, 0 = the layout's default."""
        p = LaunchParams(refill, chunk, int(prefer_wide), spec_slack, leaf_slack, stack_lds)
        self.check(self.lib.rtq_launch_params_set(self.h, ctypes.byref(p)))

    def autotune(self, items, reps: int = 3) -> TuneResult:
        out = TuneResult()
        self.check(self.lib.rtq_autotune(self.h, items, len(items), reps, ctypes.byref(out)))
        return out

    def tune_save(self, path) -> None:
        self.check(self.lib.rtq_tune_save(self.h, str(path).encode()))

    def order_init(self, seed: int, params_ptr: int) -> None:
        self.check(self.lib.rtq_order_init(self.h, seed, params_ptr))

    def order_grad(self, params_ptr: int, samples_ptr: int, n: int, grad_ptr: int) -> float:
        loss = ctypes.c_float()
        self.check(self.lib.rtq_order_grad(self.h, params_ptr, samples_ptr, n, grad_ptr, ctypes.byref(loss)))
        return loss.value

    def order_train(self, samples_ptr: int, n: int, tp: "TrainParams", params_ptr: int, m_ptr: int, v_ptr: int,
                    step: int = 0) -> tuple[int, np.ndarray]:
        """Returns (step after training, mean loss of each epoch)."""
        st = ctypes.c_uint32(step)
        losses = np.zeros(max(tp.epochs, 1), np.float32)
        self.check(self.lib.rtq_order_train(self.h, samples_ptr, n, ctypes.byref(tp), params_ptr, m_ptr, v_ptr,
                                            ctypes.byref(st), _ptr(losses)))
        return st.value, losses[:tp.epochs]

    def order_predict(self, params_ptr: int, features_ptr: int, n: int, stride: int, prob_ptr: int) -> None:
        self.check(self.lib.rtq_order_predict(self.h, params_ptr, features_ptr, n, stride, prob_ptr))

    def order_save(self, path, params_ptr: int, m_ptr: int, v_ptr: int, step: int) -> None:
        self.check(self.lib.rtq_order_save(self.h, str(path).encode(), params_ptr, m_ptr, v_ptr, step))

    def order_load(self, path, params_ptr: int, m_ptr: int | None = None, v_ptr: int | None = None) -> int:
        st = ctypes.c_uint32()
        self.check(self.lib.rtq_order_load(self.h, str(path).encode(), params_ptr, m_ptr, v_ptr,
                                           ctypes.byref(st)))
        return st.value

    def tune_load(self, path) -> None:
        self.check(self.lib.rtq_tune_load(self.h, str(path).encode()))


class OrderModelHandle:
    """rtq_order_model: the order model owning its device memory (params,
    Adam state, sample store)."""

    def __init__(self, ctx: "Context", seed: int = 1, sample_capacity: int = 1 << 22):
        self.ctx = ctx
        h = ctypes.c_void_p()
        ctx.check(ctx.lib.rtq_order_model_create(ctx.h, seed, sample_capacity, ctypes.byref(h)))
        self.h = h

    def log(self, scene: "Scene", rays_ptr: int, n: int, flags: int = 0, stride: int = 1) -> int:
        k = ctypes.c_uint64()
        self.ctx.check(self.ctx.lib.rtq_order_model_log(self.h, scene.id, rays_ptr, n, flags, stride, ctypes.byref(k)))
        return k.value

    def clear_samples(self) -> None:
        self.ctx.check(self.ctx.lib.rtq_order_model_clear_samples(self.h))

    def train(self, tp: "TrainParams") -> np.ndarray:
        losses = np.zeros(max(tp.epochs, 1), np.float32)
        self.ctx.check(self.ctx.lib.rtq_order_model_train(self.h, ctypes.byref(tp), _ptr(losses)))
        return losses[:tp.epochs]

    def apply(self, scene: "Scene", confidence_closest: float = 1.0, confidence_any: float = 0.5) -> None:
        self.ctx.check(self.ctx.lib.rtq_order_model_apply(self.h, scene.id, ctypes.c_float(confidence_closest),
                                                          ctypes.c_float(confidence_any)))

    def save(self, path) -> None:
        self.ctx.check(self.ctx.lib.rtq_order_model_save(self.h, str(path).encode()))

    def load(self, path) -> None:
        self.ctx.check(self.ctx.lib.rtq_order_model_load(self.h, str(path).encode()))

    def info(self) -> tuple[int, int]:
        n, st = ctypes.c_uint64(), ctypes.c_uint32()
        self.ctx.check(self.ctx.lib.rtq_order_model_info(self.h, ctypes.byref(n), ctypes.byref(st)))
        return n.value, st.value

    def destroy(self) -> None:
        if self.h:
            self.ctx.lib.rtq_order_model_destroy(self.h)
            self.h = None


class Batch:
    """A batch captured in a hipGraph (rtq_batch)."""

    def __init__(self, ctx: Context, h, items):
        self.ctx, self.h, self.items = ctx, h, items

    def launch(self, stream: int | None = None) -> None:
        self.ctx.check(self.ctx.lib.rtq_batch_launch(self.h, stream))

    def destroy(self) -> None:
        if self.h:
            self.ctx.lib.rtq_batch_destroy(self.h)
            self.h = None


class Scene:
    def __init__(self, ctx: Context, sid: int, n_tris: int):
        self.ctx, self.id, self.n_tris = ctx, sid, n_tris

    def release(self) -> None:
        self.ctx.check(self.ctx.lib.rtq_scene_release(self.ctx.h, self.id))

    def build(self, search_radius: int = 16, max_leaf_size: int = 4, flags: int = 0,
              node_width: int = 2) -> "BvhInfo":
        p = BuildParams(search_radius, max_leaf_size, flags, node_width)
        self.ctx.check(self.ctx.lib.rtq_bvh_build(self.ctx.h, self.id, ctypes.byref(p)))
        return self.info()

    def update(self, tris: np.ndarray | None = None, models: np.ndarray | None = None) -> None:
        """New vertices and/or model matrices (same counts); the BVH is stale until refit or build."""
        t = None if tris is None else np.ascontiguousarray(tris, dtype=TRIANGLE_DTYPE)
        m = None if models is None else np.ascontiguousarray(models, dtype=MODEL_DTYPE)
        self.ctx.check(self.ctx.lib.rtq_scene_update(self.ctx.h, self.id, _ptr(t), 0 if t is None else len(t),
                                                     _ptr(m), 0 if m is None else len(m)))

    def update_device(self, tris_ptr: int, n: int) -> None:
        """New vertices already in device memory (n 64-B triangle records at tris_ptr); stale until refit or build."""
        self.ctx.check(self.ctx.lib.rtq_scene_update_device(self.ctx.h, self.id, tris_ptr, n))

    def refit(self) -> None:
        """New boxes for the kept tree (rtq_bvh_refit)."""
        self.ctx.check(self.ctx.lib.rtq_bvh_refit(self.ctx.h, self.id))

    def info(self) -> BvhInfo:
        out = BvhInfo()
        self.ctx.check(self.ctx.lib.rtq_bvh_info_get(self.ctx.h, self.id, ctypes.byref(out)))
        return out

    def export_ref(self) -> np.ndarray:
        n = ctypes.c_uint64()
        self.ctx.check(self.ctx.lib.rtq_bvh_export_ref(self.ctx.h, self.id, None, 0, ctypes.byref(n)))
        out = np.zeros(n.value, dtype=NODE_DTYPE)
        self.ctx.check(self.ctx.lib.rtq_bvh_export_ref(self.ctx.h, self.id, _ptr(out), n.value, ctypes.byref(n)))
        return out

    def export_clusters(self) -> dict:
        nc = 2 * self.n_tris - 1
        out = {
            "boxes": np.zeros(nc, dtype=AABB_DTYPE),
            "triangle_id": np.zeros(nc, np.uint32), "left": np.zeros(nc, np.uint32),
            "right": np.zeros(nc, np.uint32), "parent": np.zeros(nc, np.uint32),
            "sorted_indices": np.zeros(self.n_tris, np.uint32), "sorted_codes": np.zeros(self.n_tris, np.uint32),
        }
        self.ctx.check(self.ctx.lib.rtq_bvh_export_clusters(
            self.ctx.h, self.id, nc, *(_ptr(out[k]) for k in
                                       ("boxes", "triangle_id", "left", "right", "parent",
                                        "sorted_indices", "sorted_codes"))))
        return out

    def export_wide(self) -> np.ndarray:
        """The compressed 8-wide nodes, n x 32 uint32 (128 B each, level order)."""
        n = ctypes.c_uint64()
        self.ctx.check(self.ctx.lib.rtq_bvh_export_wide(self.ctx.h, self.id, None, 0, ctypes.byref(n)))
        nodes = np.zeros((n.value, 32), np.uint32)
        self.ctx.check(self.ctx.lib.rtq_bvh_export_wide(self.ctx.h, self.id, _ptr(nodes), n.value,
                                                        ctypes.byref(n)))
        return nodes

    def export_trav(self) -> tuple[np.ndarray, np.ndarray]:
        """Binary traversal nodes (n x 16 float32; words 12-14 hold u32 refs and the
        order word) and their child leaf ranges + depth (n x 4 uint32)."""
        n = ctypes.c_uint64()
        self.ctx.check(self.ctx.lib.rtq_bvh_export_trav(self.ctx.h, self.id, None, None, 0, ctypes.byref(n)))
        nodes = np.zeros((n.value, 16), np.float32)
        ranges = np.zeros((n.value, 4), np.uint32)
        self.ctx.check(self.ctx.lib.rtq_bvh_export_trav(self.ctx.h, self.id, _ptr(nodes), _ptr(ranges), n.value,
                                                        ctypes.byref(n)))
        return nodes, ranges

    def order_log(self, rays_ptr: int, n: int, samples_ptr: int, capacity: int, flags: int = 0,
                  stride: int = 1) -> int:
        """Traversal-log samples (ORDER_SAMPLE floats each) written to samples_ptr; returns their count."""
        out = ctypes.c_uint64()
        self.ctx.check(self.ctx.lib.rtq_order_log(self.ctx.h, self.id, rays_ptr, n, flags, stride, samples_ptr,
                                                  capacity, ctypes.byref(out)))
        return out.value

    def order_apply(self, params_ptr: int | None, confidence_any: float = 0.5,
                    confidence_closest: float = 0.5) -> None:
        """Order words from the model; a confidence >= 1 turns that mode's decisions off."""
        self.ctx.check(self.ctx.lib.rtq_order_apply(self.ctx.h, self.id, params_ptr,
                                                    ctypes.c_float(confidence_closest),
                                                    ctypes.c_float(confidence_any)))

    def morton_codes(self, flags: int = 0) -> np.ndarray:
        out = np.zeros(self.n_tris, np.uint32)
        self.ctx.check(self.ctx.lib.rtq_morton_codes(self.ctx.h, self.id, flags, _ptr(out)))
        return out

    def trace(self, rays_ptr: int, n: int, hits_ptr: int, flags: int = 0, stream: int | None = None) -> None:
        self.ctx.check(self.ctx.lib.rtq_trace(self.ctx.h, self.id, rays_ptr, n, flags, hits_ptr, stream))

    def trace_noalloc(self, rays_ptr: int, n: int, hits_ptr: int, flags: int = 0,
                      stream: int | None = None) -> None:
        """rtq_trace_noalloc: the same trace with no device allocation (fixed
        per-GPU scratch; overlapping calls on one GPU race)."""
        self.ctx.check(self.ctx.lib.rtq_trace_noalloc(self.ctx.h, self.id, rays_ptr, n, flags, hits_ptr, stream))

    def trace_stats(self, rays_ptr: int, n: int, hits_ptr: int, nodes_ptr: int | None, tris_ptr: int | None,
                    flags: int = 0, stream: int | None = None) -> None:
        self.ctx.check(self.ctx.lib.rtq_trace_stats(self.ctx.h, self.id, rays_ptr, n, flags, hits_ptr,
                                                    nodes_ptr, tris_ptr, stream))

    def trace_host(self, rays: np.ndarray, flags: int = 0) -> np.ndarray:
        rays = np.ascontiguousarray(rays, dtype=RAY_DTYPE)
        hits = np.zeros(len(rays), dtype=HIT_DTYPE)
        self.ctx.check(self.ctx.lib.rtq_trace_host(self.ctx.h, self.id, _ptr(rays), len(rays), flags, _ptr(hits)))
        return hits

    def raygen_diffuse(self, prim_ptr: int, hits_ptr: int, n: int, seed: int, eps: float, out_ptr: int,
                       stream: int | None = None) -> None:
        self.ctx.check(self.ctx.lib.rtq_raygen_diffuse(self.ctx.h, self.id, prim_ptr, hits_ptr, n, seed,
                                                       ctypes.c_float(eps), out_ptr, stream))

    def shade(self, hits_ptr: int, n: int, materials_ptr: int | None, n_materials: int, rgba_ptr: int,
              overlay_ptr: int | None = None, flags: int = 0, stream: int | None = None) -> None:
        """RGBA32F pixel per hit record (device buffers; materials: n_materials x 16 B)."""
        self.ctx.check(self.ctx.lib.rtq_shade(self.ctx.h, self.id, hits_ptr, n, materials_ptr, n_materials,
                                              overlay_ptr, flags, rgba_ptr, stream))

    def bvh_overlay(self, rays_ptr: int, n: int, depth: int, rgba_ptr: int, stream: int | None = None) -> None:
        """Colour of the reference's BVH view at `depth` per ray (device buffers)."""
        self.ctx.check(self.ctx.lib.rtq_bvh_overlay(self.ctx.h, self.id, rays_ptr, n, depth, rgba_ptr, stream))

    def raygen_shadow(self, prim_ptr: int, hits_ptr: int, n: int, lights: np.ndarray, eps: float, out_ptr: int,
                      stream: int | None = None) -> None:
        lights = np.ascontiguousarray(lights, dtype=np.float32).reshape(-1)
        self.ctx.check(self.ctx.lib.rtq_raygen_shadow(self.ctx.h, self.id, prim_ptr, hits_ptr, n, _ptr(lights),
                                                      len(lights) // 3, ctypes.c_float(eps), out_ptr, stream))
