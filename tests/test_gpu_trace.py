"""GPU traversal parity against the CPU oracle.

Closest hit: ids, t, u, v bit-identical to the oracle's traversal (which is
itself checked equal to brute force); ties go to the lowest triangle id.
RefMT: bit-identical to the restated reference traversal (raytracer.glsl).
Any hit: occlusion flags identical to "closest hit exists in [tmin, tmax]".
Ray generation: bit-identical rays.
"""
import numpy as np
import pytest

import oracle
from realtimeraytracing_amd import (HIT_DTYPE, RAY_DTYPE, RTQ_TRACE_ANY, RTQ_TRACE_LEAF_REFMT, devmem,
                                    scenes)
from conftest import first_diff, hits_equal

pytestmark = pytest.mark.gpu


class Built:
    def __init__(self, ctx, tris, models, leaf=4):
        self.tris, self.models = tris, models
        self.sc = ctx.upload(tris, models)
        self.info = self.sc.build(16, leaf, 0)
        self.nodes = self.sc.export_ref()
        self.v9 = oracle.world_vertices(tris, models)


@pytest.fixture(scope="module")
def soup(rtq_ctx):
    b = Built(rtq_ctx, *scenes.soup_scene())
    yield b
    b.sc.release()


def soup_rays():
    cam = scenes.pinhole(**scenes.SOUP_CAMERA)
    return oracle.raygen_pinhole(*scenes.pinhole_arrays(cam), cam.width, cam.height)


def random_rays(n, seed, tmax=np.inf, lo=-0.2, hi=1.2):
    rng = np.random.default_rng(seed)
    r = np.zeros(n, RAY_DTYPE)
    r["o"] = rng.uniform(lo, hi, (n, 3)).astype(np.float32)
    d = rng.normal(size=(n, 3))
    r["d"] = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    r["tmin"] = 0.0
    r["tmax"] = tmax
    return r


def test_pinhole_raygen_bit_exact(rtq_ctx):
    for cfg in (scenes.SOUP_CAMERA, scenes.TERRAIN_CAMERA):
        cam = scenes.pinhole(**cfg)
        n = cam.width * cam.height
        d = devmem.empty(n, RAY_DTYPE)
        rtq_ctx.raygen_pinhole(cam, d.data_ptr())
        rtq_ctx.synchronize()
        got = devmem.to_host(d, RAY_DTYPE)
        want = oracle.raygen_pinhole(*scenes.pinhole_arrays(cam), cam.width, cam.height)
        assert hits_equal(got, want), first_diff(got, want)


def test_closest_hit_soup(soup):
    rays = soup_rays()
    got = soup.sc.trace_host(rays)
    want, _, _ = oracle.trace_bvh(soup.nodes, soup.v9, rays)
    assert hits_equal(got, want), first_diff(got, want)
    sub = np.ascontiguousarray(rays[::37])
    brute = oracle.trace_brute(soup.v9, sub)
    assert hits_equal(got[::37], brute), first_diff(got[::37], brute)
    assert 0.3 < (got["id"] >= 0).mean() < 0.6


def test_refmt_matches_reference_traversal(soup):
    rays = soup_rays()
    got = soup.sc.trace_host(rays, RTQ_TRACE_LEAF_REFMT)
    want = oracle.trace_ref(soup.nodes, soup.v9, rays, leaf=oracle.LEAF_REFMT)
    assert hits_equal(got, want), first_diff(got, want)
    sub = np.ascontiguousarray(rays[::53])
    brute = oracle.trace_brute(soup.v9, sub, leaf=oracle.LEAF_REFMT)
    assert hits_equal(got[::53], brute)


@pytest.mark.parametrize("leaf", [1, 2, 8])
def test_leaf_sizes_same_hits(rtq_ctx, leaf):
    b = Built(rtq_ctx, *scenes.soup_scene(20000, seed=5), leaf=leaf)
    try:
        rays = random_rays(30000, seed=leaf)
        got = b.sc.trace_host(rays)
        want, _, _ = oracle.trace_bvh(b.nodes, b.v9, rays)
        assert hits_equal(got, want), first_diff(got, want)
    finally:
        b.sc.release()


def test_random_rays_closest_and_any(soup):
    rays = random_rays(100_000, seed=11)
    got = soup.sc.trace_host(rays)
    want, _, _ = oracle.trace_bvh(soup.nodes, soup.v9, rays)
    assert hits_equal(got, want), first_diff(got, want)
    seg = random_rays(100_000, seed=12, tmax=0.05)
    occ = soup.sc.trace_host(seg, RTQ_TRACE_ANY)
    ref, _, _ = oracle.trace_bvh(soup.nodes, soup.v9, seg)
    assert np.array_equal(occ["id"] >= 0, ref["id"] >= 0)
    hit = occ["id"] >= 0
    assert hit.any() and (~hit).any()
    # an any-hit result is a genuine hit of that triangle within the interval
    for i in np.nonzero(hit)[0][:200]:
        h, t, u, v = oracle.leaf_test(seg[i], soup.v9[occ["id"][i]])
        assert h and t == occ["t"][i] and u == occ["u"][i] and v == occ["v"][i]


def test_ties_lowest_id(rtq_ctx):
    # the same triangle three times (ids 0..2) and a copy far away; every hit is a tie
    base = scenes.soup(1, seed=1)
    tris = np.concatenate([base, base, base, base])
    tris["p0"][3, :3] += 5.0
    tris["p1"][3, :3] += 5.0
    tris["p2"][3, :3] += 5.0
    models = scenes.identity_models()
    b = Built(rtq_ctx, tris, models, leaf=1)
    try:
        c = (tris["p0"][0, :3] + tris["p1"][0, :3] + tris["p2"][0, :3]) / 3.0
        rays = np.zeros(3, RAY_DTYPE)
        rays["o"] = c + np.array([0.0, 0.0, -1.0], np.float32)
        rays["d"] = (0.0, 0.0, 1.0)
        rays["tmax"] = np.inf
        rays["d"][1] = (0.0, 0.0, -1.0)
        rays["o"][1] = c + np.array([0.0, 0.0, 1.0], np.float32)
        rays["o"][2] = c + np.array([0.0, 1.0, 0.0], np.float32)
        rays["d"][2] = (0.0, -1.0, 0.0)
        got = b.sc.trace_host(rays)
        want = oracle.trace_brute(b.v9, rays)
        assert hits_equal(got, want), first_diff(got, want)
        assert (got["id"] == 0).all()
    finally:
        b.sc.release()


def test_edge_cases(rtq_ctx):
    # single triangle, degenerate rays, empty intervals, axis-parallel directions
    tri = scenes.triangles_from_vertices(np.array([[0.0, 0.0, 0.0]]), np.array([[1.0, 0.0, 0.0]]),
                                         np.array([[0.0, 1.0, 0.0]]))
    b = Built(rtq_ctx, tri, scenes.identity_models(), leaf=1)
    try:
        rays = np.zeros(8, RAY_DTYPE)
        rays["tmax"] = np.inf
        rays["o"][0], rays["d"][0] = (0.25, 0.25, -1.0), (0.0, 0.0, 1.0)       # straight hit
        rays["o"][1], rays["d"][1] = (0.25, 0.25, 1.0), (0.0, 0.0, -1.0)       # from behind (two-sided)
        rays["o"][2], rays["d"][2] = (0.25, 0.25, -1.0), (0.0, 0.0, 0.0)       # zero direction
        rays["o"][3], rays["d"][3] = (0.25, 0.25, -1.0), (0.0, 0.0, 1.0)
        rays["tmin"][3], rays["tmax"][3] = 2.0, 1.0                             # empty interval
        rays["o"][4], rays["d"][4] = (0.5, 0.5, -1.0), (0.0, 0.0, 1.0)         # exactly on the hypotenuse
        rays["o"][5], rays["d"][5] = (0.0, 0.0, -1.0), (0.0, 0.0, 1.0)         # exactly on a vertex
        rays["o"][6], rays["d"][6] = (-1.0, 0.25, 0.0), (1.0, 0.0, 0.0)        # in the triangle's plane
        rays["o"][7], rays["d"][7] = (0.25, 0.25, -1.0), (0.0, 0.0, 1.0)
        rays["tmax"][7] = 0.5                                                   # interval ends before the hit
        got = b.sc.trace_host(rays)
        want = oracle.trace_brute(b.v9, rays)
        assert hits_equal(got, want), first_diff(got, want)
        assert got["id"].tolist()[:4] == [0, 0, -1, -1]
        assert got["id"][7] == -1
        assert got["t"][0] == 1.0 and got["u"][0] == 0.25 and got["v"][0] == 0.25
    finally:
        b.sc.release()


def test_watertight_shared_edges(rtq_ctx):
    # rays aimed exactly at grid vertices and edge midpoints of a flat grid never leak
    tris = scenes.terrain(33, seed=0, amplitude=0.0)
    b = Built(rtq_ctx, tris, scenes.identity_models())
    try:
        g = np.arange(1, 32) / 32.0
        xs, zs = np.meshgrid(g, np.concatenate([g, g + 1.0 / 64.0]), indexing="ij")
        n = xs.size
        rays = np.zeros(n, RAY_DTYPE)
        rays["o"][:, 0] = xs.reshape(-1)
        rays["o"][:, 1] = 1.0
        rays["o"][:, 2] = zs.reshape(-1)
        rays["d"] = (0.0, -1.0, 0.0)
        rays["tmax"] = np.inf
        keep = rays["o"][:, 2] < 1.0
        rays = np.ascontiguousarray(rays[keep])
        got = b.sc.trace_host(rays)
        assert (got["id"] >= 0).all()
        want = oracle.trace_brute(b.v9, rays)
        assert hits_equal(got, want), first_diff(got, want)
    finally:
        b.sc.release()


@pytest.fixture(scope="module")
def terrain(rtq_ctx):
    b = Built(rtq_ctx, *scenes.terrain_scene())
    yield b
    b.sc.release()


def test_terrain_workload_parity(rtq_ctx, terrain, torch_mod):
    """Primary rays on the GPU, secondary rays generated on the GPU; every
    ray buffer and a sample of each hit buffer checked against the oracle."""
    cam = scenes.pinhole(**scenes.TERRAIN_CAMERA)
    n = cam.width * cam.height
    d_prim = devmem.empty(n, RAY_DTYPE)
    d_hit = devmem.empty(n, HIT_DTYPE)
    rtq_ctx.raygen_pinhole(cam, d_prim.data_ptr())
    terrain.sc.trace(d_prim.data_ptr(), n, d_hit.data_ptr())
    d_diff = devmem.empty(n, RAY_DTYPE)
    d_shad = devmem.empty(4 * n, RAY_DTYPE)
    terrain.sc.raygen_diffuse(d_prim.data_ptr(), d_hit.data_ptr(), n, scenes.DIFFUSE_SEED, scenes.SECONDARY_EPS,
                              d_diff.data_ptr())
    terrain.sc.raygen_shadow(d_prim.data_ptr(), d_hit.data_ptr(), n, scenes.LIGHTS, scenes.SECONDARY_EPS,
                             d_shad.data_ptr())
    d_hd = devmem.empty(n, HIT_DTYPE)
    d_hs = devmem.empty(4 * n, HIT_DTYPE)
    terrain.sc.trace(d_diff.data_ptr(), n, d_hd.data_ptr())
    terrain.sc.trace(d_shad.data_ptr(), 4 * n, d_hs.data_ptr(), RTQ_TRACE_ANY)
    rtq_ctx.synchronize()
    prim = devmem.to_host(d_prim, RAY_DTYPE)
    hit = devmem.to_host(d_hit, HIT_DTYPE)
    want_prim = oracle.raygen_pinhole(*scenes.pinhole_arrays(cam), cam.width, cam.height)
    assert hits_equal(prim, want_prim)
    want_hit, _, _ = oracle.trace_bvh(terrain.nodes, terrain.v9, prim)
    assert hits_equal(hit, want_hit), first_diff(hit, want_hit)
    assert (hit["id"] >= 0).all(), "every primary ray lands on the terrain"
    diff = devmem.to_host(d_diff, RAY_DTYPE)
    shad = devmem.to_host(d_shad, RAY_DTYPE)
    assert hits_equal(diff, oracle.raygen_diffuse(terrain.v9, prim, hit, scenes.DIFFUSE_SEED, scenes.SECONDARY_EPS))
    assert hits_equal(shad, oracle.raygen_shadow(terrain.v9, prim, hit, scenes.LIGHTS, scenes.SECONDARY_EPS))
    hd = devmem.to_host(d_hd, HIT_DTYPE)
    hs = devmem.to_host(d_hs, HIT_DTYPE)
    want_d, _, _ = oracle.trace_bvh(terrain.nodes, terrain.v9, diff)
    assert hits_equal(hd, want_d), first_diff(hd, want_d)
    want_s, _, _ = oracle.trace_bvh(terrain.nodes, terrain.v9, shad)
    assert np.array_equal(hs["id"] >= 0, want_s["id"] >= 0)


def test_stats_counters(rtq_ctx, terrain):
    cam = scenes.pinhole(**scenes.TERRAIN_CAMERA)
    n = cam.width * cam.height
    d_prim = devmem.empty(n, RAY_DTYPE)
    rtq_ctx.raygen_pinhole(cam, d_prim.data_ptr())
    d_h1, d_h2 = devmem.empty(n, HIT_DTYPE), devmem.empty(n, HIT_DTYPE)
    d_nv, d_nt = devmem.empty(n, np.uint32), devmem.empty(n, np.uint32)
    terrain.sc.trace(d_prim.data_ptr(), n, d_h1.data_ptr())
    terrain.sc.trace_stats(d_prim.data_ptr(), n, d_h2.data_ptr(), d_nv.data_ptr(), d_nt.data_ptr())
    rtq_ctx.synchronize()
    assert hits_equal(devmem.to_host(d_h1, HIT_DTYPE), devmem.to_host(d_h2, HIT_DTYPE))
    nv = devmem.to_host(d_nv, np.uint32)
    nt = devmem.to_host(d_nt, np.uint32)
    assert (nv >= 1).all() and (nt >= 1).all()
    assert nv.max() < terrain.info.n_trav_nodes


def test_vmem_probe_rates():
    """The dependent-load probe bench.py prices the trace kernels against:
    per-lane 16-B loads that hit L1 go faster than loads that miss it, and
    both rates are in the range of one CU clock per load or two."""
    import bench
    r = bench.probe_vmem()
    assert r is not None
    assert 1e11 < r["l1_miss"] < r["l1_hit"] < 5e12, r
    # a wave with only half its lanes loading takes more than half a full
    # wave's time per load: its per-lane rate is lower, though above half
    for k in ("l1_miss", "l1_hit"):
        assert 0.5 * r[k] < r[k + "_32"] < 0.95 * r[k], (k, r)
