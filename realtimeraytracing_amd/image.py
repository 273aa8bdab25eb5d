"""RGBA32F pixels (rtq_shade output) to 8-bit images for visual checks.

The reference shows its rgba32f image on a full-screen quad whose top-left
vertex samples texture coordinate (0, 1) (srcOpenGL/application.cpp:114-128),
so texel row 0 is the bottom of the window, and its fragment shader drops
alpha (raytracer.frag).  A PPM written here has the window's orientation
(texel row height-1 first) and stores round(clamp(c, 0, 1) * 255) per colour,
the normalized fixed-point conversion of an 8-bit framebuffer (ties up; NaN
as 0).
"""
from __future__ import annotations

from pathlib import Path

import numpy as np


def to_rgb8(rgba: np.ndarray, width: int, height: int) -> np.ndarray:
    """height x width x 3 uint8, window orientation (top row first)."""
    px = np.asarray(rgba, np.float32).reshape(height, width, 4)[::-1, :, :3].astype(np.float64)
    px = np.nan_to_num(px, nan=0.0, posinf=1.0, neginf=0.0)     # NaN (undefined in GL) as 0
    return np.floor(np.clip(px, 0.0, 1.0) * 255.0 + 0.5).astype(np.uint8)


def ppm_bytes(rgba: np.ndarray, width: int, height: int) -> bytes:
    """Binary PPM (P6, maxval 255) of an RGBA32F texel array (texel j * width + i)."""
    return f"P6\n{width} {height}\n255\n".encode("ascii") + to_rgb8(rgba, width, height).tobytes()


def write_ppm(path, rgba: np.ndarray, width: int, height: int) -> Path:
    path = Path(path)
    path.write_bytes(ppm_bytes(rgba, width, height))
    return path


def read_ppm(path) -> np.ndarray:
    """height x width x 3 uint8 from a binary PPM written by write_ppm."""
    data = Path(path).read_bytes()
    parts = data.split(b"\n", 3)
    if parts[0] != b"P6" or parts[2] != b"255":
        raise ValueError("not a P6 / 255 PPM")
    w, h = (int(x) for x in parts[1].split())
    return np.frombuffer(parts[3], np.uint8, count=w * h * 3).reshape(h, w, 3)
