"""One frame of the reference application's compute pass (application.cpp:220-246,
raytracer.glsl:299-331) through the C ABI: primary rays, closest-hit trace,
optional BVH view, shading into an RGBA32F image.

The reference traces with its Moller-Trumbore test (RTQ_TRACE_LEAF_REFMT, the
default here); `leaf_refmt=False` shades the watertight hits instead.
"""
from __future__ import annotations

import numpy as np

from . import devmem
from .rtq import HIT_DTYPE, MATERIAL_DTYPE, RAY_DTYPE, RTQ_SHADE_WIREFRAME, RTQ_TRACE_BINARY, RTQ_TRACE_LEAF_REFMT


def materials_array(colors) -> np.ndarray:
    """cr::MaterialGPU records from an (n, 4) array of RGBA colours."""
    c = np.asarray(colors, np.float32).reshape(-1, 4)
    out = np.zeros(len(c), MATERIAL_DTYPE)
    out["color"] = c
    return out


def render(ctx, scene, cam, materials, wireframe: bool = False, bvh_depth: int | None = None,
           leaf_refmt: bool = True):
    """Returns (rgba device tensor of width*height float4, hits device tensor).
    `cam` is an rtq_pinhole (e.g. ReferenceCamera.pinhole(width, height));
    `materials` an array of MATERIAL_DTYPE records."""
    n = cam.width * cam.height
    rays = devmem.empty(n, RAY_DTYPE)
    hits = devmem.empty(n, HIT_DTYPE)
    rgba = devmem.empty(n, np.dtype((np.float32, 4)))
    mats = devmem.to_device(np.ascontiguousarray(materials, MATERIAL_DTYPE)) if len(materials) else None
    devmem.synchronize()   # allocations and the material copy are on torch's stream
    ctx.raygen_pinhole(cam, rays.data_ptr())
    scene.trace(rays.data_ptr(), n, hits.data_ptr(), (RTQ_TRACE_LEAF_REFMT if leaf_refmt else 0) | RTQ_TRACE_BINARY)
    overlay = None
    if bvh_depth is not None:
        overlay = devmem.empty(n, np.dtype((np.float32, 4)))
        scene.bvh_overlay(rays.data_ptr(), n, bvh_depth, overlay.data_ptr())
    scene.shade(hits.data_ptr(), n, None if mats is None else mats.data_ptr(), len(materials), rgba.data_ptr(),
                None if overlay is None else overlay.data_ptr(), RTQ_SHADE_WIREFRAME if wireframe else 0)
    ctx.synchronize()
    return rgba, hits
