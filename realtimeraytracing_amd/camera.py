"""The reference's fly camera (srcCommon/scene/camera.hpp:32-61, camera.cpp) as
a source of rtq_pinhole bases.

The reference builds a primary ray per pixel (raytracer.glsl:92-100) as
    d = normalize(InvView * vec4((pos - 0.5) * (W_p, H_p), near, 1) - eye)
with InvView = inverse(lookAt(eye, eye + At, Up)).  lookAt's view space looks
down -z, so a point at view-space +near lies *behind* the camera: the rays
run along -At, and the image is mirrored left-right relative to a camera
looking along At (the default camera at (0, 0, -5) with yaw -90 therefore
faces +z, towards a model at the origin).  In basis form that is
    d = normalize(Right * W_p * s + Up * H_p * t - At * near),
    s = i / width - 0.5,  t = j / height - 0.5,
which is what `ReferenceCamera.pinhole` hands to rtq_raygen_pinhole.  The
camera vectors (At, Right, Up, lookAt's basis) and the state updates
(movement, mouse look) are float32 like glm's; the plane scaling is done in
float64 and rounded once.  The GPU then evaluates the basis form, not the
shader's matrix product, so directions agree with the shader's to rounding,
not bit for bit (GLSL leaves normalize's precision to the driver).
"""
from __future__ import annotations

import math

import numpy as np

from .rtq import Pinhole

FORWARD, BACKWARD, LEFT, RIGHT, UP, DOWN = range(6)  # cr::CameraMovement, camera.hpp:12-19

MOVEMENT_ACCELERATION = 5.0   # camera.hpp:51-53
MOVEMENT_SPEED = 20.0
MOUSE_SENSITIVITY = 0.1


def _normalize(v: np.ndarray) -> np.ndarray:
    """glm::normalize in float32: v * inversesqrt(dot(v, v))."""
    v = np.asarray(v, np.float32)
    d = np.float32(v[0] * v[0] + v[1] * v[1] + v[2] * v[2])
    return (v * np.float32(np.float32(1.0) / np.sqrt(d))).astype(np.float32)


def _cross(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    a = np.asarray(a, np.float32)
    b = np.asarray(b, np.float32)
    return np.array([a[1] * b[2] - b[1] * a[2], a[2] * b[0] - b[2] * a[0], a[0] * b[1] - b[0] * a[1]], np.float32)


_RADIANS = np.float32(0.017453292519943295)   # glm::radians for float


class ReferenceCamera:
    """cr::Camera: position, aspect ratio, vertical fov in degrees (45), near
    (0.1), far (200), world up (0, 1, 0); yaw -90 and pitch 0 degrees."""

    def __init__(self, position, aspect_ratio: float, fov: float = 45.0, near: float = 0.1, far: float = 200.0,
                 world_up=(0.0, 1.0, 0.0)):
        self.eye = np.asarray(position, np.float32).copy()
        self.aspect_ratio = np.float32(aspect_ratio)
        self.fov = np.float32(fov)
        self.near = np.float32(near)
        self.far = np.float32(far)
        self.world_up = np.asarray(world_up, np.float32).copy()
        self.yaw = np.float32(-90.0)
        self.pitch = np.float32(0.0)
        self.accelerate = False
        self.update_vectors()

    def update_vectors(self) -> None:
        """camera.cpp:110-120: At from yaw/pitch, then Right and Up (float32)."""
        yaw = np.float32(self.yaw * _RADIANS)
        pitch = np.float32(self.pitch * _RADIANS)
        cy, sy = np.cos(yaw), np.sin(yaw)
        cp, sp = np.cos(pitch), np.sin(pitch)
        self.at = _normalize(np.array([cy * cp, sp, sy * cp], np.float32))
        self.right = _normalize(_cross(self.at, self.world_up))
        self.up = _normalize(_cross(self.right, self.at))

    def plane_height(self) -> float:
        """camera.cpp:52-54: 2 near tan(fov / 2)."""
        return 2.0 * float(self.near) * math.tan(0.5 * math.radians(float(self.fov)))

    def plane_width(self) -> float:
        """camera.cpp:60-62 (getPlaneWidth(planeHeight), as getGpuData uses it)."""
        return self.plane_height() * float(self.aspect_ratio)

    def process_keyboard(self, direction: int, dt: float) -> None:
        """camera.cpp:68-92."""
        v = np.float32(MOVEMENT_SPEED * dt)
        if self.accelerate:
            v = np.float32(v * MOVEMENT_ACCELERATION)
        step = {FORWARD: -self.at, BACKWARD: self.at, LEFT: -self.right, RIGHT: self.right,
                UP: self.world_up, DOWN: -self.world_up}[direction]
        self.eye = (self.eye + step * v).astype(np.float32)

    def process_mouse(self, xoffset: float, yoffset: float, constrain_pitch: bool = True) -> None:
        """camera.cpp:94-108: sensitivity 0.1, pitch clamped to +-89 degrees."""
        self.yaw = np.float32(self.yaw + np.float32(xoffset * MOUSE_SENSITIVITY))
        self.pitch = np.float32(self.pitch + np.float32(yoffset * MOUSE_SENSITIVITY))
        if constrain_pitch:
            self.pitch = np.float32(min(max(float(self.pitch), -89.0), 89.0))
        self.update_vectors()

    def pinhole(self, width: int, height: int) -> Pinhole:
        """rtq_pinhole reproducing getRay (raytracer.glsl:92-100) for a width x
        height image: ray j * width + i is texel (i, j), row 0 at the bottom."""
        h = self.plane_height()
        w = h * float(self.aspect_ratio)
        near = float(self.near)
        # lookAt(eye, eye + At, Up)'s basis: f = normalize(center - eye), s = normalize(f x up), u = s x f
        f32 = _normalize((self.eye + self.at).astype(np.float32) - self.eye)
        s32 = _normalize(_cross(f32, self.up))
        u32 = _cross(s32, f32)
        f, s, u = (x.astype(np.float64) for x in (f32, s32, u32))
        cam = Pinhole()
        for k in range(3):
            cam.eye[k] = float(self.eye[k])
            cam.right[k] = s[k] * w
            cam.up[k] = u[k] * h
            cam.forward[k] = -f[k] * near
        cam.width, cam.height = width, height
        return cam
