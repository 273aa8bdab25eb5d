"""Deterministic synthetic scenes, cameras and lights for the benchmark workloads.

Every random choice comes from a counter-based hash: SplitMix64 applied as
h = sm(sm(sm(seed) ^ stream) ^ index), u = (h >> 40) / 2^24 in [0, 1).
Geometry is generated in float64 and rounded once to float32.

Workloads (sizes as used by the benchmark and the parity tests):
  soup(65536)          random triangle soup: centres U[0,1]^3, three vertex
                       offsets U[-0.01, 0.01]^3 each, seed 0
  terrain(1024)        1024 x 1024 height-field vertices over x, z in [0,1],
                       2 * 1023^2 = 2,093,058 triangles, 6-octave value noise
                       (amplitude 0.1, persistence 0.5, lacunarity 2, base
                       frequency 4)
  sphere(k)            tessellated unit sphere with seeded vertex jitter
"""
from __future__ import annotations

import math

import numpy as np

from .rtq import MODEL_DTYPE, TRIANGLE_DTYPE, Pinhole

_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def _sm(z: np.ndarray) -> np.ndarray:
    with np.errstate(over="ignore"):
        z = (z + np.uint64(0x9E3779B97F4A7C15)) & _M64
        z = ((z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)) & _M64
        z = ((z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)) & _M64
        return z ^ (z >> np.uint64(31))


def rng_hash(seed: int, stream: int, index) -> np.ndarray:
    idx = np.asarray(index, dtype=np.uint64)
    s = _sm(np.asarray(seed, dtype=np.uint64))
    s = _sm(s ^ np.uint64(stream))
    return _sm(s ^ idx)


def u01(seed: int, stream: int, index) -> np.ndarray:
    """Uniform in [0,1) with 24 bits, as float64 (exactly representable in float32)."""
    return (rng_hash(seed, stream, index) >> np.uint64(40)).astype(np.float64) * (1.0 / 16777216.0)


def identity_models(n: int = 1) -> np.ndarray:
    m = np.zeros(n, dtype=MODEL_DTYPE)
    m["m"][:] = np.eye(4, dtype=np.float32).reshape(16)
    return m


def triangles_from_vertices(v0: np.ndarray, v1: np.ndarray, v2: np.ndarray, model_id: int = 0) -> np.ndarray:
    n = len(v0)
    t = np.zeros(n, dtype=TRIANGLE_DTYPE)
    for name, v in (("p0", v0), ("p1", v1), ("p2", v2)):
        t[name][:, :3] = np.asarray(v, dtype=np.float64).astype(np.float32)
        t[name][:, 3] = 1.0
    t["model_id"] = model_id
    return t


def soup(n: int, seed: int = 0, extent: float = 1.0, offset: float = 0.01, origin=(0.0, 0.0, 0.0)) -> np.ndarray:
    i = np.arange(n, dtype=np.uint64)
    centres = np.stack([u01(seed, 0, 3 * i + k) for k in range(3)], axis=1) * extent + np.asarray(origin)
    verts = []
    for v in range(3):
        off = np.stack([u01(seed, 1, 9 * i + 3 * v + k) for k in range(3)], axis=1)
        verts.append(centres + (off * 2.0 - 1.0) * offset)
    return triangles_from_vertices(*verts)


def value_noise(x: np.ndarray, z: np.ndarray, seed: int, octave: int) -> np.ndarray:
    """2-D value noise: lattice values 2u-1 hashed per integer point, smoothstep blend."""
    xi = np.floor(x)
    zi = np.floor(z)
    fx = x - xi
    fz = z - zi
    xi = xi.astype(np.int64)
    zi = zi.astype(np.int64)

    def lattice(a, b):
        key = ((a.astype(np.uint64) & np.uint64(0xFFFFFFFF)) << np.uint64(32)) | (b.astype(np.uint64) & np.uint64(0xFFFFFFFF))
        return u01(seed, 1000 + octave, key) * 2.0 - 1.0

    sx = fx * fx * (3.0 - 2.0 * fx)
    sz = fz * fz * (3.0 - 2.0 * fz)
    v00, v10 = lattice(xi, zi), lattice(xi + 1, zi)
    v01, v11 = lattice(xi, zi + 1), lattice(xi + 1, zi + 1)
    a = v00 + (v10 - v00) * sx
    b = v01 + (v11 - v01) * sx
    return a + (b - a) * sz


def terrain_heights(res: int, seed: int = 0, octaves: int = 6, amplitude: float = 0.1, persistence: float = 0.5,
                    lacunarity: float = 2.0, base_freq: float = 4.0, x0: float = 0.0, z0: float = 0.0,
                    size_x: float = 1.0, size_z: float = 1.0, res_z: int | None = None):
    res_z = res if res_z is None else res_z
    xs = x0 + np.arange(res, dtype=np.float64) * (size_x / (res - 1))
    zs = z0 + np.arange(res_z, dtype=np.float64) * (size_z / (res_z - 1))
    X, Z = np.meshgrid(xs, zs, indexing="ij")  # X[i, j] = xs[i]
    h = np.zeros_like(X)
    amp, freq = amplitude, base_freq
    for o in range(octaves):
        h += amp * value_noise(X * freq, Z * freq, seed, o)
        amp *= persistence
        freq *= lacunarity
    return X, h, Z


def terrain(res: int = 1024, seed: int = 0, res_z: int | None = None, **kw) -> np.ndarray:
    """Height field with 2 triangles per cell; both wound so cross(P1-P0, P2-P0) points +y."""
    X, Y, Z = terrain_heights(res, seed, res_z=res_z, **kw)
    rz = Y.shape[1]
    P = np.stack([X, Y, Z], axis=-1)
    a = P[:-1, :-1].reshape(-1, 3)   # (i, j)
    b = P[1:, :-1].reshape(-1, 3)    # (i+1, j)
    c = P[:-1, 1:].reshape(-1, 3)    # (i, j+1)
    d = P[1:, 1:].reshape(-1, 3)     # (i+1, j+1)
    n_cells = (res - 1) * (rz - 1)
    v0 = np.empty((2 * n_cells, 3))
    v1 = np.empty_like(v0)
    v2 = np.empty_like(v0)
    v0[0::2], v1[0::2], v2[0::2] = a, c, d
    v0[1::2], v1[1::2], v2[1::2] = a, d, b
    return triangles_from_vertices(v0, v1, v2)


def sphere(subdiv: int, seed: int = 0, jitter: float = 1e-3, centre=(0.0, 0.0, 0.0), radius: float = 1.0) -> np.ndarray:
    """UV-tessellated unit sphere (subdiv rings x 2*subdiv segments), vertices jittered U[-j, j]^3."""
    rings, segs = subdiv, 2 * subdiv
    th = np.linspace(0.0, math.pi, rings + 1)
    ph = np.linspace(0.0, 2.0 * math.pi, segs + 1)
    T, P = np.meshgrid(th, ph, indexing="ij")
    V = np.stack([np.sin(T) * np.cos(P), np.cos(T), np.sin(T) * np.sin(P)], axis=-1)
    flat = V.reshape(-1, 3)
    k = np.arange(len(flat), dtype=np.uint64)
    jit = np.stack([u01(seed, 2, 3 * k + c) for c in range(3)], axis=1) * 2.0 - 1.0
    flat = flat + jit * jitter
    # seam / pole duplicates get the same jitter as their twin so the mesh stays closed
    V = flat.reshape(rings + 1, segs + 1, 3)
    V[:, -1] = V[:, 0]
    V[0, :] = V[0, 0]
    V[-1, :] = V[-1, 0]
    V = V * radius + np.asarray(centre)
    a = V[:-1, :-1].reshape(-1, 3)
    b = V[1:, :-1].reshape(-1, 3)
    c = V[:-1, 1:].reshape(-1, 3)
    d = V[1:, 1:].reshape(-1, 3)
    tris0 = (a, b, d)
    tris1 = (a, d, c)
    v0 = np.concatenate([tris0[0], tris1[0]])
    v1 = np.concatenate([tris0[1], tris1[1]])
    v2 = np.concatenate([tris0[2], tris1[2]])
    # drop the degenerate triangles at the poles
    area = np.linalg.norm(np.cross(v1 - v0, v2 - v0), axis=1)
    keep = area > 1e-12
    return triangles_from_vertices(v0[keep], v1[keep], v2[keep])


def pinhole(eye, target, up=(0.0, 1.0, 0.0), vfov_deg: float = 60.0, width: int = 256, height: int = 256) -> Pinhole:
    """Basis for rtq_raygen_pinhole, computed in float64 and rounded once."""
    eye = np.asarray(eye, np.float64)
    f = np.asarray(target, np.float64) - eye
    f /= np.linalg.norm(f)
    r = np.cross(f, np.asarray(up, np.float64))
    r /= np.linalg.norm(r)
    u = np.cross(r, f)
    h = 2.0 * math.tan(math.radians(vfov_deg) / 2.0)
    aspect = width / height
    cam = Pinhole()
    for k in range(3):
        cam.eye[k] = eye[k]
        cam.right[k] = r[k] * h * aspect
        cam.up[k] = u[k] * h
        cam.forward[k] = f[k]
    cam.width, cam.height = width, height
    return cam


def pinhole_arrays(cam: Pinhole):
    return (np.array(cam.eye[:], np.float32), np.array(cam.right[:], np.float32),
            np.array(cam.up[:], np.float32), np.array(cam.forward[:], np.float32))


# ---- the benchmark workloads --------------------------------------------------

SOUP_CAMERA = dict(eye=(0.5, 0.5, -1.0), target=(0.5, 0.5, 0.0), up=(0.0, 1.0, 0.0), vfov_deg=60.0,
                   width=256, height=256)
# Looking down onto the terrain so every primary ray of the 16:9 image lands on it.
TERRAIN_CAMERA = dict(eye=(0.5, 0.95, 0.3), target=(0.5, 0.0, 0.5), up=(0.0, 1.0, 0.0), vfov_deg=26.0,
                      width=1920, height=1080)
LIGHTS = np.array([[0.25, 2.0, 0.25], [0.75, 2.0, 0.25], [0.25, 2.0, 0.75], [0.75, 2.0, 0.75]], np.float32)
SECONDARY_EPS = 1e-4
DIFFUSE_SEED = 7


def sphere_with(n_tris: int, seed: int, jitter: float = 1e-3) -> np.ndarray:
    """A jittered unit sphere with about n_tris triangles (4k^2 - 4k for k rings)."""
    k = max(2, int(round(math.sqrt(max(n_tris, 8) / 4.0))))
    return sphere(k, seed=seed, jitter=jitter)


def random_rays(n: int, seed: int, lo, hi) -> np.ndarray:
    """Origins uniform in the box [lo, hi], directions uniform on the sphere."""
    from .rtq import RAY_DTYPE
    i = np.arange(n, dtype=np.uint64)
    lo = np.asarray(lo, np.float64)
    hi = np.asarray(hi, np.float64)
    o = np.stack([u01(seed, 10 + k, i) for k in range(3)], axis=1) * (hi - lo) + lo
    # uniform direction: z = 2u - 1, phi = 2 pi v
    z = 2.0 * u01(seed, 13, i) - 1.0
    phi = 2.0 * math.pi * u01(seed, 14, i)
    r = np.sqrt(np.maximum(0.0, 1.0 - z * z))
    d = np.stack([r * np.cos(phi), r * np.sin(phi), z], axis=1)
    rays = np.zeros(n, RAY_DTYPE)
    rays["o"] = o.astype(np.float32)
    rays["d"] = d.astype(np.float32)
    rays["tmin"] = 0.0
    rays["tmax"] = np.inf
    return rays


def batched_scene(s: int, seed: int = 0, tri_range=(1_000, 4_000_000), ray_range=(65_536, 2_097_152)):
    """Scene s of the uneven batch: a triangle count log-uniform in tri_range,
    random soup for even s and a jittered sphere for odd s, with a log-uniform
    number of random-direction rays from inside its bounds.  Returns
    (triangles, rays)."""
    lt = math.log(tri_range[0]) + float(u01(seed, 20, s)) * (math.log(tri_range[1]) - math.log(tri_range[0]))
    lr = math.log(ray_range[0]) + float(u01(seed, 21, s)) * (math.log(ray_range[1]) - math.log(ray_range[0]))
    n_tris, n_rays = int(math.exp(lt)), int(math.exp(lr))
    if s % 2 == 0:
        tris = soup(n_tris, seed=seed * 1000 + s)
        lo, hi = (0.0, 0.0, 0.0), (1.0, 1.0, 1.0)
    else:
        tris = sphere_with(n_tris, seed=seed * 1000 + s)
        lo, hi = (-0.9, -0.9, -0.9), (0.9, 0.9, 0.9)
    return tris, random_rays(n_rays, seed * 1000 + s, lo, hi)


def batched_workload(n_scenes: int = 64, seed: int = 0, tri_range=(1_000, 4_000_000), ray_range=(65_536, 2_097_152)):
    """The uneven batch (config 2 at the default ranges) as a list of (triangles, rays)."""
    return [batched_scene(s, seed, tri_range, ray_range) for s in range(n_scenes)]


def tiled_terrains(res: int = 1024, tiles: int = 2, seed: int = 0) -> np.ndarray:
    """tiles x tiles terrains of res x res vertices side by side over
    [0, tiles]^2 in x and z, tile (a, b) with noise seed seed + a * tiles + b,
    so neighbouring tiles differ (jittered heights) and share no vertices."""
    out = []
    for a in range(tiles):
        for b in range(tiles):
            t = terrain(res, seed + a * tiles + b)
            for v in ("p0", "p1", "p2"):
                t[v][:, 0] += a
                t[v][:, 2] += b
            out.append(t)
    return np.concatenate(out)


# Camera over the 2 x 2 tiled terrains: every primary ray hits
TILED_CAMERA = dict(eye=(1.0, 1.4, 0.2), target=(1.0, 0.0, 1.0), up=(0.0, 1.0, 0.0), vfov_deg=40.0,
                    width=2048, height=2048)


def soup_scene(n: int = 65536, seed: int = 0):
    return soup(n, seed), identity_models()


def terrain_scene(res: int = 1024, seed: int = 0):
    return terrain(res, seed), identity_models()


def occluded_scene(terrain_res: int = 1024, soup_n: int = 2_097_152, seed: int = 0):
    """Terrain plus a random soup over the unit cube (soup seed = seed + 1): the
    learned-ordering workload, where shadow rays are often occluded."""
    return np.concatenate([terrain(terrain_res, seed), soup(soup_n, seed + 1)]), identity_models()
