"""The C ABI: the library loads, exports exactly what include/rtq/*.h declares,
the headers compile as C and C++ with the reference layouts, and argument
errors come back as status codes (no GPU needed for any of this)."""
import ctypes
import re
import shutil
import subprocess
from pathlib import Path

import numpy as np
import pytest

from realtimeraytracing_amd import (AABB_DTYPE, EXPORTED, HIT_DTYPE, MODEL_DTYPE, NODE_DTYPE, RAY_DTYPE,
                                    TRIANGLE_DTYPE, BatchItem, BuildParams, BvhInfo, LaunchParams, Pinhole,
                                    TrainParams, TuneResult, load_library)
from realtimeraytracing_amd.rtq import MATERIAL_DTYPE, RTQ_EINVAL

ROOT = Path(__file__).resolve().parents[1]
HEADERS = sorted((ROOT / "include" / "rtq").glob("*.h"))


def declared_functions():
    names = set()
    for h in HEADERS:
        text = h.read_text()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names |= set(re.findall(r"\b(rtq_\w+)\s*\(", text))
    return names


def test_exports_match_header():
    lib_path = ROOT / "realtimeraytracing_amd" / "librtq.so"
    assert lib_path.exists(), "build first"
    load_library()
    out = subprocess.run(["nm", "-D", "--defined-only", str(lib_path)], capture_output=True, text=True).stdout
    exported = {ln.split()[-1] for ln in out.splitlines() if " T " in ln}
    declared = declared_functions()
    assert declared == set(EXPORTED)
    assert declared <= exported, declared - exported
    assert {e for e in exported if not e.startswith("_")} == declared   # nothing else leaks


def test_probe_library_exports_its_entry_point():
    """The vector-memory probe bench.py measures with is its own library: it
    exports its two entry points, and librtq does not carry them."""
    probe = ROOT / "realtimeraytracing_amd" / "librtqprobe.so"
    assert probe.exists(), "build first"
    out = subprocess.run(["nm", "-D", "--defined-only", str(probe)], capture_output=True, text=True).stdout
    exported = {ln.split()[-1] for ln in out.splitlines() if " T " in ln}
    assert {e for e in exported if not e.startswith("_")} == {"rtqprobe_vmem_rate", "rtqprobe_vmem_rate_lanes"}
    lib = ctypes.CDLL(str(probe))
    fn = lib.rtqprobe_vmem_rate
    fn.argtypes = [ctypes.c_int, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, ctypes.POINTER(ctypes.c_double)]
    fn.restype = ctypes.c_int
    r = ctypes.c_double(0.0)
    assert fn(0, 9, 16, 6, ctypes.byref(r)) != 0    # 9 loads per step: refused before any device call
    assert fn(0, 6, 16, 6, None) != 0
    lanes = lib.rtqprobe_vmem_rate_lanes
    lanes.argtypes = [ctypes.c_int, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
                      ctypes.POINTER(ctypes.c_double)]
    lanes.restype = ctypes.c_int
    assert lanes(0, 6, 16, 6, 0, ctypes.byref(r)) != 0     # no active lane
    assert lanes(0, 6, 16, 6, 65, ctypes.byref(r)) != 0    # more lanes than a wave has


def test_numpy_layouts():
    assert TRIANGLE_DTYPE.itemsize == 64 and TRIANGLE_DTYPE.fields["model_id"][1] == 48
    assert MODEL_DTYPE.itemsize == 68
    assert AABB_DTYPE.itemsize == 32 and AABB_DTYPE.fields["max"][1] == 16
    assert NODE_DTYPE.itemsize == 48
    assert [NODE_DTYPE.fields[k][1] for k in ("triangle_id", "left_child", "right_child")] == [32, 36, 40]
    assert RAY_DTYPE.itemsize == 32 and HIT_DTYPE.itemsize == 16
    assert MATERIAL_DTYPE.itemsize == 16


SIZES_C = r"""
#include <stdio.h>
#include "rtq/rtq.h"
int main(void) {
    printf("%zu %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu\n", sizeof(rtq_triangle), sizeof(rtq_mesh_model),
           sizeof(rtq_aabb), sizeof(rtq_bvh_node), sizeof(rtq_ray), sizeof(rtq_hit), sizeof(rtq_material),
           sizeof(rtq_build_params), sizeof(rtq_bvh_info), sizeof(rtq_pinhole), sizeof(rtq_batch_item),
           sizeof(rtq_launch_params), sizeof(rtq_tune_result), sizeof(rtq_train_params));
    return 0;
}
"""


@pytest.mark.parametrize("compiler,std", [("gcc", "-std=c11"), ("g++", "-std=c++17")])
def test_headers_compile_and_sizes(tmp_path, compiler, std):
    if shutil.which(compiler) is None:
        pytest.skip(f"{compiler} missing")
    src = tmp_path / ("sizes.c" if compiler == "gcc" else "sizes.cpp")
    src.write_text(SIZES_C)
    exe = tmp_path / "sizes"
    subprocess.run([compiler, std, "-Wall", "-Werror", f"-I{ROOT / 'include'}", str(src), "-o", str(exe)],
                   check=True)
    sizes = [int(x) for x in subprocess.run([str(exe)], capture_output=True, text=True).stdout.split()]
    assert sizes == [64, 68, 32, 48, 32, 16, 16] + [ctypes.sizeof(c) for c in (
        BuildParams, BvhInfo, Pinhole, BatchItem, LaunchParams, TuneResult, TrainParams)]


def test_argument_errors_without_device():
    lib = load_library()
    assert lib.rtq_abi_version() == 2
    assert lib.rtq_status_string(RTQ_EINVAL) == b"invalid argument"
    assert lib.rtq_last_error(None) == b"null context"
    p = BuildParams()
    lib.rtq_build_params_default(ctypes.byref(p))
    assert (p.search_radius, p.max_leaf_size, p.flags) == (16, 4, 0)
    sid = ctypes.c_uint32()
    tris = np.zeros(1, TRIANGLE_DTYPE)
    models = np.zeros(1, MODEL_DTYPE)
    # a null context is rejected before any device work
    assert lib.rtq_scene_upload(None, tris.ctypes.data, 1, models.ctypes.data, 1, ctypes.byref(sid)) == RTQ_EINVAL
    assert lib.rtq_bvh_build(None, 1, None) == RTQ_EINVAL
    assert lib.rtq_trace(None, 1, None, 0, 0, None, None) == RTQ_EINVAL
    assert lib.rtq_create(None, 0) == RTQ_EINVAL
    # the batch, tuning and learned-ordering entry points too
    n = ctypes.c_uint64()
    assert lib.rtq_bvh_export_trav(None, 1, None, None, 0, ctypes.byref(n)) == RTQ_EINVAL
    assert lib.rtq_trace_batch(None, None, 0, None) == RTQ_EINVAL
    assert lib.rtq_autotune(None, None, 0, 1, None) == RTQ_EINVAL
    assert lib.rtq_order_log(None, 1, None, 0, 0, 1, None, 0, ctypes.byref(n)) == RTQ_EINVAL
    assert lib.rtq_order_init(None, 1, None) == RTQ_EINVAL
    assert lib.rtq_order_grad(None, None, None, 1, None, None) == RTQ_EINVAL
    assert lib.rtq_order_train(None, None, 1, None, None, None, None, None, None) == RTQ_EINVAL
    assert lib.rtq_order_apply(None, 1, None, 0.5, 0.5) == RTQ_EINVAL
    assert lib.rtq_order_model_create(None, 1, 16, None) == RTQ_EINVAL
    assert lib.rtq_order_model_log(None, 1, None, 0, 0, 1, ctypes.byref(n)) == RTQ_EINVAL
    assert lib.rtq_order_model_train(None, None, None) == RTQ_EINVAL
    assert lib.rtq_order_model_apply(None, 1, 1.0, 0.5) == RTQ_EINVAL
    assert lib.rtq_order_model_info(None, ctypes.byref(n), None) == RTQ_EINVAL
    lib.rtq_order_model_destroy(None)   # no-op
    assert lib.rtq_shade(None, 1, None, 0, None, 0, None, 0, None, None) == RTQ_EINVAL
    assert lib.rtq_bvh_overlay(None, 1, None, 0, 0, None, None) == RTQ_EINVAL
