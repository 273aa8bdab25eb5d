"""OBJ ingest (rtq_obj_load / rtq_obj_parse, Mesh::load mesh.cpp:186-263) and
the host mesh helpers (mesh.cpp:19-180): hand files with known triangles, an
independent numpy reading of the reference's bundled models when the reference
checkout is present, and closed forms for the model-matrix setters."""
import ctypes
import math
from pathlib import Path

import numpy as np
import pytest

import oracle
from realtimeraytracing_amd import TRIANGLE_DTYPE, RtqError, obj_load, obj_parse, scenes
from realtimeraytracing_amd import load_library
from realtimeraytracing_amd.mesh import Mesh, scene_arrays
from realtimeraytracing_amd.rtq import RTQ_ECAPACITY, RTQ_EINVAL

MODELS = Path("/root/reference/resources/models")


def verts(t):
    return np.stack([t["p0"][:, :3], t["p1"][:, :3], t["p2"][:, :3]], axis=1)


def test_parse_forms_and_indices():
    text = ("# comment\n"
            "o thing\nv 0 0 0\nv 1 0 0 1.0\nv 1 1 0 0.5 0.5 0.5\nvt 0 0\nvn 0 0 1\n"
            "g g1\ns off\nusemtl m\n"
            "f 1 2 3\r\n"
            "f 1/1 2/1 3/1\n"
            "f 1//1 2//1 3//1\n"
            "f\t1/1/1   2/1/1 3/1/1  \n"
            "f -3 -2 -1\n"
            "f 1 2\n"                       # fewer than 3: skipped
            "l 1 2\n")
    t = obj_parse(text, model_id=7)
    assert len(t) == 5
    for k in range(5):
        assert verts(t)[k].tolist() == [[0, 0, 0], [1, 0, 0], [1, 1, 0]]
    assert (t["p0"][:, 3] == 1).all() and (t["p2"][:, 3] == 1).all() and (t["model_id"] == 7).all()


def test_quads_and_polygons():
    # diagonal 0-2 shorter: (0,1,2) + (0,2,3)
    t = obj_parse("v 0 0 0\nv 2 0 0\nv 1 0.5 0\nv 0 1 0\nf 1 2 3 4\n")
    assert verts(t).tolist() == [[[0, 0, 0], [2, 0, 0], [1, 0.5, 0]], [[0, 0, 0], [1, 0.5, 0], [0, 1, 0]]]
    # diagonal 1-3 shorter or equal: (0,1,3) + (1,2,3)
    t = obj_parse("v 0 0 0\nv 1 0 0\nv 1 1 0\nv 0 1 0\nf 1 2 3 4\n")
    assert verts(t).tolist() == [[[0, 0, 0], [1, 0, 0], [0, 1, 0]], [[1, 0, 0], [1, 1, 0], [0, 1, 0]]]
    # a face may name vertices defined after it
    t = obj_parse("f 1 2 3 4\nv 0 0 0\nv 1 0 0\nv 1 1 0\nv 0 1 0\n")
    assert len(t) == 2
    # pentagon: fan from the first vertex
    t = obj_parse("v 0 0 0\nv 1 0 0\nv 2 1 0\nv 1 2 0\nv 0 1 0\nf 1 2 3 4 5\n")
    assert verts(t)[:, 0].tolist() == [[0, 0, 0]] * 3
    assert verts(t)[:, 2].tolist() == [[2, 1, 0], [1, 2, 0], [0, 1, 0]]


@pytest.mark.parametrize("text,what", [("v 0 0 0\nf 1 2 3\n", "past the last vertex"),
                                       ("v 0 0 0\nv 1 0 0\nv 0 1 0\nf 0 1 2\n", "index 0"),
                                       ("v 0 0 0\nf -2 1 1\n", "before the first vertex"),
                                       ("v 0 zero 0\n", "bad vertex"),
                                       ("v 0 0 0\nf 1 a 1\n", "bad face index")])
def test_parse_errors(text, what):
    with pytest.raises(RtqError) as e:
        obj_parse(text)
    assert e.value.status == RTQ_EINVAL and what in str(e.value)


def test_load_errors_and_capacity(tmp_path):
    with pytest.raises(RtqError) as e:
        obj_load(tmp_path / "missing.obj")
    assert e.value.status == RTQ_EINVAL and "cannot open" in str(e.value)
    lib = load_library()
    data = b"v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 3\nf 1 2 3\n"
    n = ctypes.c_uint64()
    buf = np.zeros(1, TRIANGLE_DTYPE)
    assert lib.rtq_obj_parse(data, len(data), 0, buf.ctypes.data, 1, ctypes.byref(n), None, 0) == RTQ_ECAPACITY
    assert n.value == 2
    assert lib.rtq_obj_parse(None, 0, 0, None, 0, ctypes.byref(n), None, 0) == 0 and n.value == 0
    assert lib.rtq_obj_parse(data, len(data), 0, None, 0, None, None, 0) == RTQ_EINVAL


def numpy_obj(path):
    """Independent reading of an OBJ file: float() per coordinate rounded to
    float32, 1-based / negative indices, quads split on the shorter diagonal."""
    vs, faces = [], []
    for line in Path(path).read_text().splitlines():
        p = line.split()
        if not p:
            continue
        if p[0] == "v":
            vs.append([float(x) for x in (p[1:4] + ["0"] * 3)[:3]])
        elif p[0] == "f":
            idx = [int(tok.split("/")[0]) for tok in p[1:]]
            faces.append([i - 1 if i > 0 else len(vs) + i for i in idx])
    v = np.asarray(vs, np.float64).astype(np.float32)
    out = []
    for f in faces:
        if len(f) == 3:
            out.append(f)
        elif len(f) == 4:
            e02 = v[f[2]] - v[f[0]]
            e13 = v[f[3]] - v[f[1]]
            s02 = np.float32(np.float32(e02[0] * e02[0] + e02[1] * e02[1]) + e02[2] * e02[2])
            s13 = np.float32(np.float32(e13[0] * e13[0] + e13[1] * e13[1]) + e13[2] * e13[2])
            if s02 < s13:
                out += [[f[0], f[1], f[2]], [f[0], f[2], f[3]]]
            else:
                out += [[f[0], f[1], f[3]], [f[1], f[2], f[3]]]
        elif len(f) > 4:
            out += [[f[0], f[k], f[k + 1]] for k in range(1, len(f) - 1)]
    return v[np.asarray(out)]


@pytest.mark.parametrize("name,count", [("teapot.obj", 6320), ("stanford-bunny.obj", 69451),
                                        ("suzanne.obj", 32 + 2 * 468)])
def test_reference_models(name, count):
    path = MODELS / name
    if not path.exists():
        pytest.skip("reference checkout not present")
    t = obj_load(path)
    assert len(t) == count
    assert np.array_equal(verts(t), numpy_obj(path))


def test_mesh_primitives():
    assert len(Mesh.primitive_triangle().triangles) == 1 and len(Mesh.primitive_square().triangles) == 2
    cube = verts(Mesh.primitive_cube().triangles).astype(np.float64)
    assert cube.shape == (12, 3, 3) and set(np.unique(cube)) == {-1.0, 1.0}
    # two triangles of area 2 on each of the 6 faces
    area = 0.5 * np.linalg.norm(np.cross(cube[:, 1] - cube[:, 0], cube[:, 2] - cube[:, 0]), axis=1)
    assert np.allclose(area, 2.0)
    faces = [(k, s) for t in cube for k in range(3) for s in (-1.0, 1.0) if (t[:, k] == s).all()]
    assert sorted(set(faces)) == sorted({(k, s) for k in range(3) for s in (-1.0, 1.0)}) and len(faces) == 12


def textbook(axis, a):
    c, s = math.cos(a), math.sin(a)
    if axis == 0:
        return np.array([[1, 0, 0], [0, c, -s], [0, s, c]])
    if axis == 1:
        return np.array([[c, 0, s], [0, 1, 0], [-s, 0, c]])
    return np.array([[c, -s, 0], [s, c, 0], [0, 0, 1]])


def test_model_setters_closed_form():
    m = Mesh.primitive_cube()
    m.set_scale(2.0)
    m.set_position((1.0, -2.0, 3.0))
    tris, models = scene_arrays([m])
    w = oracle.world_vertices(tris, models).reshape(-1, 3, 3)
    assert np.array_equal(w, verts(m.triangles) * 2.0 + np.array([1.0, -2.0, 3.0], np.float32))
    # setRotation on an identity model: textbook Rx Ry Rz (z applied first)
    a, b, c = 0.3, -1.1, 2.0
    r = Mesh.primitive_cube()
    r.set_rotation(a, b, c)
    tris, models = scene_arrays([r])
    w = oracle.world_vertices(tris, models).reshape(-1, 3, 3).astype(np.float64)
    R = textbook(0, a) @ textbook(1, b) @ textbook(2, c)
    expect = verts(r.triangles).astype(np.float64) @ R.T
    assert np.abs(w - expect).max() < 1e-6
    # a later setScale scales in object space: M * S
    r.set_scale(0.5)
    tris, models = scene_arrays([r])
    w2 = oracle.world_vertices(tris, models).reshape(-1, 3, 3).astype(np.float64)
    assert np.abs(w2 - 0.5 * expect).max() < 1e-6


@pytest.mark.gpu
def test_meshes_trace_on_gpu(rtq_ctx, torch_mod):
    """Meshes from OBJ text and primitives, placed by the setters, uploaded as
    one scene: GPU hits equal brute force over the oracle's world vertices."""
    grid = ["v %d %d 0" % (x, y) for y in range(5) for x in range(5)]
    quads = ["f %d %d %d %d" % (y * 5 + x + 1, y * 5 + x + 2, (y + 1) * 5 + x + 2, (y + 1) * 5 + x + 1)
             for y in range(4) for x in range(4)]
    plane = Mesh(obj_parse("\n".join(grid + quads)))
    plane.set_scale(0.25)
    plane.set_position((-0.5, -0.5, 1.0))
    cube = Mesh.primitive_cube()
    cube.set_scale(0.2)
    cube.set_rotation(0.4, 0.7, 0.1)
    tri = Mesh.primitive_triangle()
    tri.set_scale(0.3)
    tri.set_position((0.3, 0.3, -0.5))
    tris, models = scene_arrays([plane, cube, tri])
    sc = rtq_ctx.upload(tris, models)
    try:
        sc.build(16, 4, 0)
        cam = scenes.pinhole(eye=(0.0, 0.0, -3.0), target=(0.0, 0.0, 0.0), vfov_deg=30.0, width=64, height=64)
        rays = oracle.raygen_pinhole(*scenes.pinhole_arrays(cam), 64, 64)
        v9 = oracle.world_vertices(tris, models)
        got = sc.trace_host(rays)
        ref = oracle.trace_brute(v9, rays)
        assert np.array_equal(got.view(np.uint8), ref.view(np.uint8))
        ids = got["id"][got["id"] >= 0]
        # every mesh is seen
        assert (ids < 32).any() and ((ids >= 32) & (ids < 44)).any() and (ids == 44).any()
    finally:
        sc.release()
