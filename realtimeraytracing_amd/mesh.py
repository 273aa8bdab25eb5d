"""cr::Mesh (srcCommon/scene/geometry/mesh.hpp:17-46, mesh.cpp) on the host:
triangles plus a MeshModelGPU record, the reference's primitives and OBJ
loading, and the model-matrix setters with glm's float arithmetic.

Matrices are glm::mat4 values stored column-major (m[c * 4 + r]).  The
setters keep the reference's compositions: setScale and setRotation replace
M by transpose(X * transpose(M)) = M * X^T (mesh.cpp:30-60), and glm's
initializer lists are columns, so the rotation matrices written there are the
transposes of the textbook ones: for M = I the result rotates by Rx Ry Rz
(textbook, z first).  Products follow glm's order (column c of A * B is
((A0 B[c][0] + A1 B[c][1]) + A2 B[c][2]) + A3 B[c][3]); sin and cos are the
float results of the double functions, where glm calls the float ones (the
last bit can differ).

`scene_arrays` concatenates meshes the way Scene::getTriangleToGPUData does
(scene.cpp:24-38), numbering models by their position in the list; the
reference stores the mesh's global creation id (mesh.cpp:15-17), which is
the same number when every mesh created is added, in order.
"""
from __future__ import annotations

import math

import numpy as np

from .rtq import MODEL_DTYPE, TRIANGLE_DTYPE, obj_load


def _cols(m16) -> np.ndarray:
    """column-major 16 floats -> [column][row] float32"""
    return np.asarray(m16, np.float32).reshape(4, 4).copy()


def glm_mul(a16, b16) -> np.ndarray:
    """glm mat4 * mat4 in float32, glm's summation order."""
    A, B = _cols(a16), _cols(b16)
    out = np.zeros((4, 4), np.float32)
    for c in range(4):
        acc = A[0] * B[c][0]
        acc = (acc + A[1] * B[c][1]).astype(np.float32)
        acc = (acc + A[2] * B[c][2]).astype(np.float32)
        acc = (acc + A[3] * B[c][3]).astype(np.float32)
        out[c] = acc
    return out.reshape(16)


def glm_transpose(a16) -> np.ndarray:
    return _cols(a16).T.copy().reshape(16)


def glm_columns(*cols) -> np.ndarray:
    """glm::mat4({c0}, {c1}, {c2}, {c3}): each initializer list is a column."""
    return np.asarray(cols, np.float32).reshape(16)


IDENTITY = np.eye(4, dtype=np.float32).reshape(16)


class Mesh:
    def __init__(self, triangles: np.ndarray | None = None):
        self.triangles = np.zeros(0, TRIANGLE_DTYPE) if triangles is None else np.asarray(triangles, TRIANGLE_DTYPE)
        self.model = IDENTITY.copy()
        self.material_id = 0

    # ---- model matrix (mesh.cpp:19-61) ----
    def set_model(self, m16) -> None:
        self.model = np.asarray(m16, np.float32).reshape(16).copy()

    def set_position(self, position) -> None:
        self.model[12:15] = np.asarray(position, np.float32)

    def set_scale(self, scale: float) -> None:
        s = np.float32(scale)
        S = IDENTITY.copy()
        S[0] *= s
        S[5] *= s
        S[10] *= s
        self.model = glm_transpose(glm_mul(S, glm_transpose(self.model)))

    def set_rotation(self, theta_x: float, theta_y: float, theta_z: float) -> None:
        def cs(t):
            t = float(np.float32(t))
            return np.float32(math.cos(t)), np.float32(math.sin(t))
        cx, sx = cs(theta_x)
        cy, sy = cs(theta_y)
        cz, sz = cs(theta_z)
        rx = glm_columns([1, 0, 0, 0], [0, cx, -sx, 0], [0, sx, cx, 0], [0, 0, 0, 1])
        ry = glm_columns([cy, 0, sy, 0], [0, 1, 0, 0], [-sy, 0, cy, 0], [0, 0, 0, 1])
        rz = glm_columns([cz, -sz, 0, 0], [sz, cz, 0, 0], [0, 0, 1, 0], [0, 0, 0, 1])
        r = glm_mul(glm_mul(glm_mul(rz, ry), rx), glm_transpose(self.model))
        self.model = glm_transpose(r)

    def set_material(self, material_id: int) -> None:
        self.material_id = int(material_id)

    def model_record(self) -> np.ndarray:
        rec = np.zeros(1, MODEL_DTYPE)
        rec["m"][0] = self.model
        rec["material_id"] = self.material_id
        return rec

    # ---- primitives and loading (mesh.cpp:63-263) ----
    @staticmethod
    def _from_vertices(vs) -> "Mesh":
        v = np.asarray(vs, np.float32).reshape(-1, 3, 3)
        t = np.zeros(len(v), TRIANGLE_DTYPE)
        for k, name in enumerate(("p0", "p1", "p2")):
            t[name][:, :3] = v[:, k]
            t[name][:, 3] = 1.0
        return Mesh(t)

    @staticmethod
    def primitive_triangle() -> "Mesh":
        return Mesh._from_vertices([[0, 1, 0], [-1, -1, 0], [1, -1, 0]])

    @staticmethod
    def primitive_square() -> "Mesh":
        return Mesh._from_vertices([[-1, 1, 0], [-1, -1, 0], [1, -1, 0],
                                    [-1, 1, 0], [1, -1, 0], [1, 1, 0]])

    @staticmethod
    def primitive_cube() -> "Mesh":
        return Mesh._from_vertices([
            [-1, 1, -1], [-1, -1, -1], [1, -1, -1], [-1, 1, -1], [1, -1, -1], [1, 1, -1],   # front
            [-1, -1, 1], [-1, 1, 1], [1, -1, 1], [1, -1, 1], [-1, 1, 1], [1, 1, 1],          # back
            [1, -1, 1], [1, 1, 1], [1, 1, -1], [1, -1, 1], [1, 1, -1], [1, -1, -1],          # right
            [-1, 1, 1], [-1, -1, 1], [-1, 1, -1], [-1, 1, -1], [-1, -1, 1], [-1, -1, -1],    # left
            [-1, 1, -1], [1, 1, -1], [1, 1, 1], [-1, 1, -1], [1, 1, 1], [-1, 1, 1],          # top
            [-1, -1, -1], [1, -1, 1], [1, -1, -1], [-1, -1, -1], [-1, -1, 1], [1, -1, 1],    # bottom
        ])

    @staticmethod
    def load(path) -> "Mesh":
        return Mesh(obj_load(path))


def scene_arrays(meshes) -> tuple[np.ndarray, np.ndarray]:
    """(triangles, models) for rtq_scene_upload: mesh i's triangles get model id i."""
    tris = []
    models = np.zeros(len(meshes), MODEL_DTYPE)
    for i, m in enumerate(meshes):
        t = m.triangles.copy()
        t["model_id"] = i
        tris.append(t)
        models[i] = m.model_record()[0]
    return (np.concatenate(tris) if tris else np.zeros(0, TRIANGLE_DTYPE)), models
