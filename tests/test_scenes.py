"""Synthetic workload generators: sizes, ranges, winding and determinism."""
import numpy as np

from realtimeraytracing_amd import scenes


def test_soup_ranges_and_determinism():
    t = scenes.soup(4096, seed=0)
    assert len(t) == 4096
    v = np.stack([t["p0"][:, :3], t["p1"][:, :3], t["p2"][:, :3]], axis=1).astype(np.float64)
    c = v.mean(axis=1)
    assert np.all(np.abs(v - c[:, None, :]) <= 0.02 + 1e-6)
    assert c.min() > -0.01 and c.max() < 1.01
    assert np.all(t["p0"][:, 3] == 1.0)
    assert np.array_equal(t.view(np.uint8), scenes.soup(4096, seed=0).view(np.uint8))
    assert not np.array_equal(t.view(np.uint8), scenes.soup(4096, seed=1).view(np.uint8))


def test_terrain_size_and_winding():
    for res in (2, 3, 17):
        t = scenes.terrain(res, seed=0)
        assert len(t) == 2 * (res - 1) ** 2
    t = scenes.terrain(64, seed=0)
    p0, p1, p2 = (t[k][:, :3].astype(np.float64) for k in ("p0", "p1", "p2"))
    n = np.cross(p1 - p0, p2 - p0)
    assert np.all(n[:, 1] > 0)                          # cross(P1-P0, P2-P0) points +y
    xs = np.unique(np.concatenate([p0[:, 0], p1[:, 0], p2[:, 0]]))
    assert len(xs) == 64 and xs[0] == 0.0 and xs[-1] == 1.0
    assert np.abs(p0[:, 1]).max() < 0.2                 # 0.1 * sum 0.5^k < 0.2
    assert len(scenes.terrain(1024)) == 2_093_058


def test_sphere_closed_and_jittered():
    t = scenes.sphere(16, seed=0, jitter=1e-3)
    v = np.concatenate([t["p0"][:, :3], t["p1"][:, :3], t["p2"][:, :3]]).astype(np.float64)
    r = np.linalg.norm(v, axis=1)
    assert np.all(np.abs(r - 1.0) < 2e-3)
    area = np.linalg.norm(np.cross(t["p1"][:, :3] - t["p0"][:, :3], t["p2"][:, :3] - t["p0"][:, :3]), axis=1)
    assert np.all(area > 0)


def test_pinhole_basis():
    cam = scenes.pinhole(**scenes.TERRAIN_CAMERA)
    eye, right, up, fwd = scenes.pinhole_arrays(cam)
    assert abs(np.linalg.norm(fwd) - 1) < 1e-6
    assert abs(right @ fwd) < 1e-6 and abs(up @ fwd) < 1e-6 and abs(right @ up) < 1e-6
    assert abs(np.linalg.norm(right) / np.linalg.norm(up) - 1920 / 1080) < 1e-5
