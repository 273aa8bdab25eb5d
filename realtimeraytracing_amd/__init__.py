"""rtq: an MI355X (gfx950) ray-query engine for PLOC-built BVHs.

The product is the C ABI in include/rtq/rtq.h, implemented by the HIP library
realtimeraytracing_amd/librtq.so.  This package holds its sources (csrc/), the
build script, ctypes bindings used by tests and the benchmark, and the
synthetic scene generators.
"""
from .rtq import (  # noqa: F401
    AABB_DTYPE, EXPORTED, HIT_DTYPE, MATERIAL_DTYPE, MODEL_DTYPE, NODE_DTYPE, RAY_DTYPE, TRIANGLE_DTYPE,
    RTQ_BUILD_PLOC_SPLIT, RTQ_BUILD_REF_PADDED_BOUNDS, RTQ_TRACE_ANY, RTQ_TRACE_BINARY, RTQ_TRACE_CLOSEST, RTQ_TRACE_LEAF_REFMT,
    RTQ_TRACE_LEARNED, RTQ_SHADE_WIREFRAME, ORDER_INPUTS, ORDER_HIDDEN, ORDER_PARAMS, ORDER_SAMPLE,
    MAX_BATCH_ITEMS, Batch, BatchItem, BuildParams, BvhInfo, Context, LaunchParams, Pinhole, RtqError, Scene,
    OrderModelHandle, TrainParams, TuneResult, load_library, obj_load, obj_parse,
)
