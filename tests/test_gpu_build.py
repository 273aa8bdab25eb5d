"""GPU build parity: Morton codes, radix sort, scan, PLOC clusters and the
pre-order node array against the CPU oracle, bit for bit."""
import numpy as np
import pytest

import oracle
from realtimeraytracing_amd import RTQ_BUILD_PLOC_SPLIT, RTQ_BUILD_REF_PADDED_BOUNDS, devmem, scenes
from conftest import first_diff, hits_equal

pytestmark = pytest.mark.gpu

SIZES = [1, 2, 3, 5, 8, 64, 1000, 4096, 4097, 9000, 65536, 200_000]


def _scene(n, seed):
    return scenes.soup(n, seed=seed), scenes.identity_models()


@pytest.mark.parametrize("n", SIZES)
def test_morton_codes(rtq_ctx, n):
    tris, models = _scene(n, seed=n)
    sc = rtq_ctx.upload(tris, models)
    try:
        for flags in (0, RTQ_BUILD_REF_PADDED_BOUNDS):
            got = sc.morton_codes(flags)
            want = oracle.morton_codes(tris, models, pad_origin=bool(flags))
            assert np.array_equal(got, want), f"flags={flags}"
    finally:
        sc.release()


def _flat_scene(n, seed, fix):
    """Triangles whose vertices share the coordinates in `fix` (axis -> value):
    the scene box is flat along those axes."""
    rng = np.random.default_rng(seed)
    v = [rng.uniform(0.0, 1.0, (n, 3)) for _ in range(3)]
    for p in v:
        for axis, value in fix.items():
            p[:, axis] = value
    return scenes.triangles_from_vertices(*v), scenes.identity_models()


@pytest.mark.parametrize("case", ["y_branch", "z_branch", "single_point"])
def test_cube_axis_branches_on_gpu(rtq_ctx, case):
    """getCircumscribedCube takes X whenever distX > 0 (bvh.cpp:263-272); a
    scene flat in x takes the Y branch, one flat in x and y the Z branch, and a
    scene of one repeated point the Z branch with a zero-size cube.  Codes and
    the built tree equal the oracle's in each case."""
    fix = {"y_branch": {0: 0.3}, "z_branch": {0: 0.3, 1: 0.7}, "single_point": {0: 0.3, 1: 0.7, 2: 0.1}}[case]
    tris, models = _flat_scene(700, seed=len(case), fix=fix)
    sc = rtq_ctx.upload(tris, models)
    try:
        for flags in (0, RTQ_BUILD_REF_PADDED_BOUNDS):
            got = sc.morton_codes(flags)
            want = oracle.morton_codes(tris, models, pad_origin=bool(flags))
            assert np.array_equal(got, want), (case, flags)
        if case != "single_point":
            assert len(np.unique(want)) > 100   # the branch spreads the codes over the cube
    finally:
        sc.release()
    if case != "single_point":
        sc, _, _ = _check_build(rtq_ctx, tris, models)
        sc.release()


def _check_build(rtq_ctx, tris, models, flags=0, radius=16, leaf=4):
    n = len(tris)
    sc = rtq_ctx.upload(tris, models)
    try:
        info = sc.build(radius, leaf, flags)
        got = sc.export_clusters()
        want = oracle.ploc_build(tris, models, pad_origin=bool(flags & RTQ_BUILD_REF_PADDED_BOUNDS), radius=radius)
        assert want["passes"] >= 0
        for k in ("sorted_codes", "sorted_indices", "triangle_id", "left", "right", "parent"):
            assert np.array_equal(got[k], want[k]), k
        assert hits_equal(got["boxes"], want["boxes"]), first_diff(got["boxes"], want["boxes"])
        nodes = sc.export_ref()
        want_nodes = oracle.flatten_preorder(n, want)
        assert hits_equal(nodes, want_nodes), first_diff(nodes, want_nodes)
        assert info.n_ref_nodes == 2 * n - 1
        assert info.ploc_iterations == want["passes"]
        assert info.depth == oracle.tree_depth(want_nodes)
        return sc, info, nodes
    except Exception:
        sc.release()
        raise


@pytest.mark.parametrize("n", SIZES)
def test_ploc_and_flatten(rtq_ctx, n):
    tris, models = _scene(n, seed=100 + n)
    sc, info, _ = _check_build(rtq_ctx, tris, models)
    sc.release()


@pytest.mark.parametrize("radius,leaf", [(1, 1), (4, 2), (32, 8)])
def test_ploc_radius_and_leaf_size(rtq_ctx, radius, leaf):
    tris, models = _scene(3000, seed=radius)
    sc, info, _ = _check_build(rtq_ctx, tris, models, radius=radius, leaf=leaf)
    assert info.max_leaf_size == leaf
    sc.release()


@pytest.mark.parametrize("n,radius", [(4097, 16), (65536, 1)])
def test_split_ploc_schedule(rtq_ctx, n, radius):
    """The multi-launch PLOC schedule (RTQ_BUILD_PLOC_SPLIT) against the oracle."""
    tris, models = _scene(n, seed=7 * n)
    sc, _, _ = _check_build(rtq_ctx, tris, models, flags=RTQ_BUILD_PLOC_SPLIT, radius=radius)
    sc.release()


@pytest.mark.parametrize("n,radius", [(4481, 16), (30000, 1), (30000, 32), (500_000, 16)])
def test_fused_and_split_ploc_agree(rtq_ctx, n, radius):
    """Fused passes (one launch each, decoupled look-back between tiles, then one
    workgroup for the last 4096 clusters) and the multi-launch schedule give the
    same clusters, node array, pass count and traversal tree."""
    tris, models = _scene(n, seed=n + radius)
    sc = rtq_ctx.upload(tris, models)
    try:
        out = []
        for flags in (0, RTQ_BUILD_PLOC_SPLIT):
            info = sc.build(radius, 4, flags)
            cl = sc.export_clusters()
            trav, rng = sc.export_trav()
            out.append((info, cl, sc.export_ref(), trav.view(np.uint32).copy(), rng.copy()))
        (i0, c0, r0, t0, g0), (i1, c1, r1, t1, g1) = out
        for k in ("triangle_id", "left", "right", "parent", "sorted_indices"):
            assert np.array_equal(c0[k], c1[k]), k
        assert hits_equal(c0["boxes"], c1["boxes"])
        assert hits_equal(r0, r1)
        assert np.array_equal(t0, t1) and np.array_equal(g0, g1)
        assert (i0.ploc_iterations, i0.depth) == (i1.ploc_iterations, i1.depth)
    finally:
        sc.release()


@pytest.mark.parametrize("n", [9000, 200_000])
def test_split_schedule_8wide_nodes(rtq_ctx, n):
    """The 8-wide collapse reads the PLOC passes the build recorded: both
    schedules give the same 8-wide tree, and its hits equal the binary ones."""
    from realtimeraytracing_amd import RAY_DTYPE, RTQ_TRACE_ANY, RTQ_TRACE_BINARY
    tris, models = _scene(n, seed=n + 3)
    rng = np.random.default_rng(n)
    rays = np.zeros(30_000, RAY_DTYPE)
    rays["o"] = rng.uniform(-0.2, 1.2, (len(rays), 3)).astype(np.float32)
    d = rng.normal(size=(len(rays), 3))
    rays["d"] = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    rays["tmax"] = np.inf
    sc = rtq_ctx.upload(tris, models)
    try:
        out = []
        for flags in (0, RTQ_BUILD_PLOC_SPLIT):
            info = sc.build(16, 4, flags, node_width=8)
            assert info.node_width == 8
            wide, binary = sc.trace_host(rays), sc.trace_host(rays, RTQ_TRACE_BINARY)
            assert hits_equal(wide, binary), first_diff(wide, binary)
            occ_w, occ_b = sc.trace_host(rays, RTQ_TRACE_ANY), sc.trace_host(rays, RTQ_TRACE_ANY | RTQ_TRACE_BINARY)
            assert np.array_equal(occ_w["id"] >= 0, occ_b["id"] >= 0)
            out.append((info.n_wide_nodes, info.wide_depth, wide))
        assert out[0][:2] == out[1][:2]
        assert hits_equal(out[0][2], out[1][2])
    finally:
        sc.release()


@pytest.mark.parametrize("kind", ["copies", "identical"])
def test_fused_ploc_with_ties(rtq_ctx, kind):
    """Equal Morton codes and equal merged areas across tile boundaries of the
    fused pass: 2,500 distinct triangles 4 times over, or one triangle 5,000
    times (every neighbour distance ties, so the lowest index wins and a pass
    merges about one pair: 4,999 passes, most of them in the tail)."""
    base = scenes.soup(2500, seed=11)
    tris = np.concatenate([base] * 4) if kind == "copies" else np.concatenate([base[:1]] * 5000)
    models = scenes.identity_models()
    sc, _, _ = _check_build(rtq_ctx, tris, models)
    sc.release()
    sc, _, _ = _check_build(rtq_ctx, tris, models, flags=RTQ_BUILD_PLOC_SPLIT)
    sc.release()


def test_padded_bounds_flag(rtq_ctx):
    # a scene away from the origin: the reference's zero padding moves the cube
    tris = scenes.soup(500, seed=3, origin=(2.0, 3.0, 4.0))
    models = scenes.identity_models()
    sc, _, _ = _check_build(rtq_ctx, tris, models, flags=RTQ_BUILD_REF_PADDED_BOUNDS)
    sc.release()


def test_model_matrix_and_duplicates(rtq_ctx):
    # two meshes, one scaled and translated; identical triangles force equal codes
    base = scenes.soup(200, seed=9)
    tris = np.concatenate([base, base[:50]])
    tris["model_id"][200:] = 1
    models = scenes.identity_models(2)
    m = np.eye(4, dtype=np.float32)
    m[0, 0] = m[1, 1] = m[2, 2] = 0.5
    m[3, :3] = (1.0, -2.0, 0.25)  # column 3 (translation) in column-major storage
    models["m"][1] = m.reshape(16)
    sc, _, _ = _check_build(rtq_ctx, tris, models)
    sc.release()
    dup = np.concatenate([base[:64]] * 4)
    sc, _, _ = _check_build(rtq_ctx, dup, scenes.identity_models())
    sc.release()


def test_terrain_full_size(rtq_ctx):
    tris, models = scenes.terrain_scene()
    sc, info, _ = _check_build(rtq_ctx, tris, models)
    print(f"terrain build: {info.build_ms:.1f} ms, passes {info.ploc_iterations}, depth {info.depth}, "
          f"traversal nodes {info.n_trav_nodes}", flush=True)
    sc.release()


@pytest.mark.parametrize("n,bits", [(1, 32), (2047, 30), (2048, 30), (2049, 32), (100_000, 30), (1_000_003, 32),
                                    (300_000, 8), (50_000, 12), (4_194_305, 30), (65_536, 1)])
def test_radix_sort_pairs(rtq_ctx, torch_mod, n, bits):
    rng = np.random.default_rng(n + bits)
    hi = 1 << bits
    keys = rng.integers(0, min(hi, 1 << 32), n, dtype=np.uint64).astype(np.uint32)
    keys[: n // 3] = keys[n // 3: 2 * (n // 3)]  # plenty of duplicates
    vals = np.arange(n, dtype=np.uint32)
    dk, dv = devmem.to_device(keys), devmem.to_device(vals)
    rtq_ctx.sort_pairs(dk.data_ptr(), dv.data_ptr(), n, bits)
    rtq_ctx.synchronize()
    got_k = devmem.to_host(dk, np.uint32)
    got_v = devmem.to_host(dv, np.uint32)
    order = np.argsort(keys, kind="stable")
    assert np.array_equal(got_k, keys[order])
    assert np.array_equal(got_v, vals[order])
    ck, cv = oracle.sort_pairs(keys, vals)
    assert np.array_equal(got_k, ck) and np.array_equal(got_v, cv)


@pytest.mark.parametrize("n", [1, 7, 2048, 2049, 65536, 4_194_305])
def test_exclusive_scan(rtq_ctx, n):
    rng = np.random.default_rng(n)
    a = rng.integers(0, 1 << 31, n, dtype=np.uint64).astype(np.uint32)
    din, dout = devmem.to_device(a), devmem.empty(n, np.uint32)
    rtq_ctx.scan_exclusive(din.data_ptr(), dout.data_ptr(), n)
    rtq_ctx.synchronize()
    got = devmem.to_host(dout, np.uint32)
    assert np.array_equal(got, oracle.scan_exclusive(a))


def test_scan_reference_known_values(rtq_ctx):
    # testHistogramPrefixSum.cpp:54-84: exclusive scan within each group of 4 digits
    groups = [[37, 41, 49, 53], [37, 48, 44, 51], [35, 51, 53, 41]]
    expect = [[0, 37, 78, 127], [0, 37, 85, 129], [0, 35, 86, 139]]
    for g, e in zip(groups, expect):
        a = np.array(g, np.uint32)
        din, dout = devmem.to_device(a), devmem.empty(4, np.uint32)
        rtq_ctx.scan_exclusive(din.data_ptr(), dout.data_ptr(), 4)
        rtq_ctx.synchronize()
        assert devmem.to_host(dout, np.uint32).tolist() == e
