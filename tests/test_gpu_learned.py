"""Learned child ordering on the GPU: the training kernels' gradient against
the oracle's FP64 gradient, the BF16 MFMA inference against the oracle's
BF16-rounded forward pass, the node features and order words against the
oracle, exact hit records with learned order (also with garbage weights), the
checkpoint file, and the node-visit gain of a trained model over the nearer-
child order on occluded shadow rays."""
import math

import numpy as np
import pytest

import oracle
from realtimeraytracing_amd import (HIT_DTYPE, ORDER_PARAMS, ORDER_SAMPLE, RAY_DTYPE, RTQ_TRACE_ANY,
                                    RTQ_TRACE_BINARY, RTQ_TRACE_LEARNED, OrderModelHandle, RtqError, TrainParams,
                                    devmem, scenes)
from realtimeraytracing_amd.order import OrderModel
from conftest import first_diff, hits_equal

pytestmark = pytest.mark.gpu

W1, B1, W2, B2, W3, B3 = 0, 2048, 2112, 6208, 6272, 6336


def rand_params(seed):
    rng = np.random.default_rng(seed)
    p = np.zeros(ORDER_PARAMS, np.float32)
    p[W1:B1] = rng.uniform(-1, 1, 2048) * math.sqrt(6 / 32)
    p[B1:W2] = rng.uniform(-0.1, 0.1, 64)
    p[W2:B2] = rng.uniform(-1, 1, 4096) * math.sqrt(6 / 64)
    p[B2:W3] = rng.uniform(-0.1, 0.1, 64)
    p[W3:B3] = rng.uniform(-1, 1, 64) * math.sqrt(6 / 64)
    p[B3] = 0.05
    return p


def rand_samples(n, seed):
    rng = np.random.default_rng(seed)
    s = np.zeros((n, ORDER_SAMPLE), np.float32)
    s[:, :32] = rng.uniform(-1, 1, (n, 32))
    s[:, 32] = rng.integers(0, 2, n)
    s[:, 33] = rng.uniform(0.5, 2.0, n)
    return s


def dev(a, torch):
    """Device copy; the caller keeps the tensor alive while the library uses it."""
    t = torch.from_numpy(np.ascontiguousarray(a)).cuda()
    torch.cuda.synchronize()
    return t


@pytest.mark.parametrize("n", [1, 64, 300, 4096])
def test_gradient_matches_oracle(rtq_ctx, torch_mod, n):
    torch = torch_mod
    P, S = rand_params(n), rand_samples(n, n + 1)
    g = torch.zeros(ORDER_PARAMS, dtype=torch.float32, device="cuda")
    dP, dS = dev(P, torch), dev(S, torch)
    loss = rtq_ctx.order_grad(dP.data_ptr(), dS.data_ptr(), n, g.data_ptr())
    want_loss, want = oracle.mlp_grad(P, S)
    got = g.cpu().numpy().astype(np.float64)
    scale = np.abs(want).max()
    assert loss == pytest.approx(want_loss, rel=1e-5, abs=1e-6)
    assert np.abs(got - want).max() <= 1e-5 * scale + 1e-7, np.abs(got - want).max() / scale
    # every parameter block carries gradient (a dropped term would zero one)
    for a, b in ((W1, B1), (B1, W2), (W2, B2), (B2, W3), (W3, B3), (B3, B3 + 1)):
        assert np.abs(got[a:b]).max() > 0


def test_mfma_inference_matches_bf16_oracle(rtq_ctx, torch_mod):
    torch = torch_mod
    P = rand_params(11)
    X = rand_samples(1000, 12)          # stride 34: features are the first 32 floats
    prob = torch.zeros(1000, dtype=torch.float32, device="cuda")
    dP, dX = dev(P, torch), dev(X, torch)
    rtq_ctx.order_predict(dP.data_ptr(), dX.data_ptr(), 1000, ORDER_SAMPLE, prob.data_ptr())
    p = prob.cpu().numpy().astype(np.float64)
    z = np.log(p / (1 - p))
    zb = oracle.mlp_logits(P, X, bf16=True)
    zf = oracle.mlp_logits(P, X, bf16=False)
    d = np.abs(z - zb)
    # FP32 accumulation vs FP64, plus rare BF16 rounding-boundary flips of an
    # activation: tight in the median, bounded at the worst
    assert np.median(d) < 2e-4, np.median(d)
    assert d.max() < 5e-2, d.max()
    # the BF16 model stays close to the FP32 one; a transposed operand would not
    assert np.abs(z - zf).max() < 0.1 * (1 + np.abs(zf).max())
    assert np.corrcoef(z, zf)[0, 1] > 0.999


@pytest.fixture(scope="module")
def occ(rtq_ctx, torch_mod):
    """Terrain 256 x 256 + 131,072 soup triangles; primary, diffuse and shadow rays."""
    torch = torch_mod
    tris, models = scenes.occluded_scene(256, 131_072, seed=0)
    sc = rtq_ctx.upload(tris, models)
    sc.build(16, 4, 0)
    cam = scenes.pinhole(**{**scenes.TERRAIN_CAMERA, "width": 256, "height": 256})
    n = cam.width * cam.height
    prim, phit = devmem.empty(n, RAY_DTYPE), devmem.empty(n, HIT_DTYPE)
    rtq_ctx.raygen_pinhole(cam, prim.data_ptr())
    sc.trace(prim.data_ptr(), n, phit.data_ptr())
    diff, shad = devmem.empty(n, RAY_DTYPE), devmem.empty(4 * n, RAY_DTYPE)
    sc.raygen_diffuse(prim.data_ptr(), phit.data_ptr(), n, 7, scenes.SECONDARY_EPS, diff.data_ptr())
    sc.raygen_shadow(prim.data_ptr(), phit.data_ptr(), n, scenes.LIGHTS, scenes.SECONDARY_EPS, shad.data_ptr())
    rtq_ctx.synchronize()
    torch.cuda.synchronize()
    yield dict(sc=sc, tris=tris, models=models, n=n, prim=prim, diff=diff, shad=shad)
    sc.release()


def trace(sc, rays, n, flags, stats=False):
    hits = devmem.empty(n, HIT_DTYPE)
    if stats:
        nv, nt = devmem.empty(n, np.uint32), devmem.empty(n, np.uint32)
        sc.trace_stats(rays.data_ptr(), n, hits.data_ptr(), nv.data_ptr(), nt.data_ptr(), flags)
        sc.ctx.synchronize()
        return (devmem.to_host(hits, HIT_DTYPE), devmem.to_host(nv, np.uint32).astype(np.int64),
                devmem.to_host(nt, np.uint32).astype(np.int64))
    sc.trace(rays.data_ptr(), n, hits.data_ptr(), flags)
    sc.ctx.synchronize()
    return devmem.to_host(hits, HIT_DTYPE)


def relative_boxes(nodes):
    """Features 5-16 of every node: child boxes relative to the node box, float32."""
    l0, h0 = nodes[:, [0, 2, 4]], nodes[:, [1, 3, 5]]
    l1, h1 = nodes[:, [6, 8, 10]], nodes[:, [7, 9, 11]]
    lo = np.minimum(l0, l1)
    ext = np.maximum(h0, h1) - lo
    e = np.maximum(np.float32(1e-30), ext.max(axis=1))
    inv = (np.float32(1) / e)[:, None]
    return np.concatenate([(l0 - lo) * inv, (h0 - lo) * inv, (l1 - lo) * inv, (h1 - lo) * inv], axis=1)


def test_log_features_match_oracle(rtq_ctx, occ, torch_mod):
    sc = occ["sc"]
    model = OrderModel(rtq_ctx, seed=3)
    nodes, ranges = sc.export_trav()
    for any_hit, rays, n in ((False, occ["diff"], occ["n"]), (True, occ["shad"], 4 * occ["n"])):
        samples, k = model.log(sc, rays, n, any_hit, stride=7, capacity=1 << 18)
        assert k > 1000
        S = samples.cpu().numpy()
        assert np.isfinite(S).all()
        assert set(np.unique(S[:, 32])) <= {0.0, 1.0}
        assert (S[:, 3] == (1.0 if any_hit else 0.0)).all()
        assert (S[:, 33] >= 1).all()          # log2(1 + cost difference), difference >= 1
        # each sample's features are the oracle's features of some node: find
        # the candidates by depth and relative child boxes (features 5-16, IEEE
        # float32 ops in numpy), then compare all 32 with the oracle's
        rel = relative_boxes(nodes)
        rng = np.random.default_rng(0)
        for i in rng.choice(k, size=200, replace=False):
            f = S[i, :32]
            octant = int(f[0] < 0) | (int(f[1] < 0) << 1) | (int(f[2] < 0) << 2)
            depth = int(round(f[4] * 32))
            cand = np.nonzero((ranges[:, 3] == depth) & (rel == f[5:17]).all(axis=1))[0]
            assert any(np.allclose(oracle.order_features(nodes[c], ranges[c], octant, any_hit), f, rtol=0, atol=1e-6)
                       for c in cand), f"sample {i} matches no node"


def test_training_reduces_loss_and_checkpoint(rtq_ctx, occ, torch_mod, tmp_path):
    torch = torch_mod
    sc = occ["sc"]
    model = OrderModel(rtq_ctx, seed=5)
    samples, k = model.log(sc, occ["shad"], 4 * occ["n"], True, stride=3)
    P0 = model.params.cpu().numpy()
    S = samples.cpu().numpy()
    sub = S[:20000]
    loss0 = oracle.mlp_loss(P0, sub)
    losses = model.train(samples, epochs=4, batch=1024, lr=2e-3)
    assert len(losses) == 4 and losses[-1] < losses[0]
    P1 = model.params.cpu().numpy()
    assert oracle.mlp_loss(P1, sub) < loss0
    assert model.step == 4 * ((k + 1023) // 1024)
    # checkpoint round trip (parameters, moments, step)
    path = tmp_path / "order.bin"
    model.save(path)
    other = OrderModel(rtq_ctx, seed=99)
    other.load(path)
    assert other.step == model.step
    for a, b in ((other.params, model.params), (other.m, model.m), (other.v, model.v)):
        assert torch.equal(a, b)
    data = bytearray(path.read_bytes())
    bad = tmp_path / "bad.bin"
    data[100] ^= 1
    bad.write_bytes(bytes(data))
    with pytest.raises(RtqError):
        other.load(bad)
    data = bytearray(path.read_bytes())
    data[4] = 2                                       # version
    bad.write_bytes(bytes(data))
    with pytest.raises(RtqError):
        other.load(bad)
    bad.write_bytes(path.read_bytes()[:-9])           # truncated
    with pytest.raises(RtqError):
        other.load(bad)
    # deterministic: the same samples, seed and steps give the same parameters
    again = OrderModel(rtq_ctx, seed=5)
    again.train(samples, epochs=4, batch=1024, lr=2e-3)
    assert torch.equal(again.params, model.params)


def test_order_words_match_oracle(rtq_ctx, occ, torch_mod):
    sc = occ["sc"]
    model = OrderModel(rtq_ctx, seed=8)
    model.apply(sc, confidence_any=0.6, confidence_closest=0.6)
    nodes, ranges = sc.export_trav()
    P = model.params.cpu().numpy()
    words = nodes[:, 14].view(np.uint32)
    rng = np.random.default_rng(1)
    idx = rng.choice(len(nodes), size=400, replace=False)
    zmin = math.log(0.6 / 0.4)
    checked = 0
    for c in idx:
        feats = np.stack([oracle.order_features(nodes[c], ranges[c], o, a) for a in (0, 1) for o in range(8)])
        z = oracle.mlp_logits(P, feats, bf16=True)
        for b in range(16):
            if abs(z[b]) < 0.02 or abs(abs(z[b]) - zmin) < 0.02:
                continue                               # too close to a threshold to call
            assert ((words[c] >> b) & 1) == (z[b] > 0), (c, b, z[b])
            assert ((words[c] >> (16 + b)) & 1) == (abs(z[b]) >= zmin), (c, b, z[b])
            checked += 1
    assert checked > 5000
    sc.order_apply(None)
    nodes2, _ = sc.export_trav()
    assert (nodes2[:, 14].view(np.uint32) == 0).all()
    nodes2[:, 14] = nodes[:, 14]
    assert np.array_equal(nodes2.view(np.uint32), nodes.view(np.uint32))   # only the order word changed


def test_learned_order_exact(rtq_ctx, occ, torch_mod):
    torch = torch_mod
    sc = occ["sc"]
    n = occ["n"]
    with pytest.raises(RtqError):
        trace(sc, occ["diff"], n, RTQ_TRACE_LEARNED)      # no order bits applied
    plain_d = trace(sc, occ["diff"], n, RTQ_TRACE_BINARY)
    plain_s = trace(sc, occ["shad"], 4 * n, RTQ_TRACE_ANY | RTQ_TRACE_BINARY)
    model = OrderModel(rtq_ctx, seed=9)
    samples, _ = model.log(sc, occ["shad"], 4 * n, True, stride=5)
    model.train(samples, epochs=2)
    garbage = torch.full((ORDER_PARAMS,), float("nan"), device="cuda")
    huge = dev(rand_params(4) * np.float32(1e30), torch)
    for params, conf in ((model.params, 0.5), (model.params, 0.8), (garbage, 0.5), (huge, 0.5)):
        torch.cuda.synchronize()
        sc.order_apply(params.data_ptr(), conf, conf)
        got_d = trace(sc, occ["diff"], n, RTQ_TRACE_LEARNED)
        assert hits_equal(got_d, plain_d), first_diff(got_d, plain_d)
        got_s = trace(sc, occ["shad"], 4 * n, RTQ_TRACE_ANY | RTQ_TRACE_LEARNED)
        assert np.array_equal(got_s["id"] >= 0, plain_s["id"] >= 0)
    # the oracle agrees on a sample of the closest hits
    nodes = sc.export_ref()
    v9 = oracle.world_vertices(occ["tris"], occ["models"])
    rays = devmem.to_host(occ["diff"], RAY_DTYPE)[:4000]
    want, _, _ = oracle.trace_bvh(nodes, v9, np.ascontiguousarray(rays))
    assert hits_equal(got_d[:4000], want)
    # learned items in a batch
    sc.order_apply(model.params.data_ptr(), 0.5)
    hd, hs = devmem.empty(n, HIT_DTYPE), devmem.empty(4 * n, HIT_DTYPE)
    items = rtq_ctx.items([(sc, occ["diff"].data_ptr(), n, hd.data_ptr(), RTQ_TRACE_LEARNED),
                           (sc, occ["shad"].data_ptr(), 4 * n, hs.data_ptr(), RTQ_TRACE_ANY | RTQ_TRACE_LEARNED)])
    rtq_ctx.trace_batch(items)
    rtq_ctx.synchronize()
    assert hits_equal(devmem.to_host(hd, HIT_DTYPE), plain_d)
    assert np.array_equal(devmem.to_host(hs, HIT_DTYPE)["id"] >= 0, plain_s["id"] >= 0)
    # learned and plain binary items share one launch (each keeps its own order)
    mixed = rtq_ctx.items([(sc, occ["diff"].data_ptr(), n, hd.data_ptr(), RTQ_TRACE_LEARNED),
                           (sc, occ["shad"].data_ptr(), 4 * n, hs.data_ptr(), RTQ_TRACE_ANY | RTQ_TRACE_BINARY)])
    rtq_ctx.trace_batch(mixed)
    rtq_ctx.synchronize()
    assert hits_equal(devmem.to_host(hd, HIT_DTYPE), plain_d)
    assert np.array_equal(devmem.to_host(hs, HIT_DTYPE)["id"] >= 0, plain_s["id"] >= 0)
    # with learned decisions in the batch, a scene built with 8-wide nodes is
    # traced on its binary nodes in the same launch
    wide = rtq_ctx.upload(*scenes.soup_scene(2000, 3))
    wide.build(16, 4, 0, node_width=8)
    try:
        assert sc.info().order_modes != 0 and wide.info().order_modes == 0
        hw = devmem.empty(1000, HIT_DTYPE)
        both = rtq_ctx.items([(sc, occ["diff"].data_ptr(), n, hd.data_ptr(), RTQ_TRACE_LEARNED),
                              (wide, occ["diff"].data_ptr(), 1000, hw.data_ptr(), 0)])
        rtq_ctx.trace_batch(both)
        rtq_ctx.synchronize()
        assert hits_equal(devmem.to_host(hd, HIT_DTYPE), plain_d)
        assert hits_equal(devmem.to_host(hw, HIT_DTYPE), trace(wide, occ["diff"], 1000, 0))
        # explicit binary items and 8-wide items cannot share a launch
        bad = rtq_ctx.items([(wide, occ["diff"].data_ptr(), 1000, hw.data_ptr(), RTQ_TRACE_BINARY),
                             (wide, occ["diff"].data_ptr(), 1000, hs.data_ptr(), 0)])
        with pytest.raises(RtqError):
            rtq_ctx.trace_batch(bad)
    finally:
        wide.release()
    # a mode without decisions traces like the plain kernel
    sc.order_apply(model.params.data_ptr(), 0.5, 1.0)
    assert hits_equal(trace(sc, occ["diff"], n, RTQ_TRACE_LEARNED), plain_d)
    # a rebuild clears the order bits
    sc.build(16, 4, 0)
    with pytest.raises(RtqError):
        trace(sc, occ["diff"], n, RTQ_TRACE_LEARNED)


LEAF = 0x80000000


def _leaf_range(r):
    r = r.astype(np.int64)
    first = (r & 0x7FFFFFFF) >> 3
    return first, first + (r & 7) + 1


def _node_ranges(refs, valid):
    """[lo, hi) leaf-order triangle range of every node of a tree whose children
    follow their parent: refs (n x w u32) with leaf refs or node indices."""
    n, w = refs.shape
    lo = np.full(n, 1 << 40, np.int64)
    hi = np.full(n, -1, np.int64)
    is_leaf = (refs & LEAF) != 0
    llo, lhi = _leaf_range(refs)
    for k in range(n - 1, -1, -1):
        for j in range(w):
            if not valid[k, j]:
                continue
            r = int(refs[k, j])
            a, b = (int(llo[k, j]), int(lhi[k, j])) if is_leaf[k, j] else (lo[r], hi[r])
            lo[k], hi[k] = min(lo[k], a), max(hi[k], b)
    return lo, hi


def wide_order_tables(trav, wide):
    """The learned first-child table of every 8-wide node (words 24, 25), restated
    from the binary nodes' order words: entry e = octant + 8 * any_hit follows the
    confident binary decisions (bit 16 + e set; bit e set = child 0) from the
    binary node covering the same triangles as the wide node until they reach
    one of its children (a leaf child by its ref, an internal child by its
    triangle range); 8 | slot there, 0 where a decision is missing."""
    tu = trav.view(np.uint32)
    brefs, order = tu[:, 12:14], tu[:, 14]
    n_trav = len(tu)
    blo, bhi = _node_ranges(brefs, np.ones((n_trav, 2), bool))
    by_range = {(int(a), int(b)): x for x, (a, b) in enumerate(zip(blo, bhi))}
    valid = ((wide[:, 3] >> 24)[:, None] >> np.arange(8)) & 1 != 0
    wrefs = wide[:, 16:24]
    wlo, whi = _node_ranges(wrefs, valid)
    n_wide = len(wide)
    # every child as the binary layout refers to it
    as_bin = np.full((n_wide, 8), 0xFFFFFFFF, np.int64)
    for k in range(n_wide):
        for j in range(8):
            if valid[k, j]:
                r = int(wrefs[k, j])
                as_bin[k, j] = r if r & LEAF else by_range.get((int(wlo[r]), int(whi[r])), -2)
    root = np.array([by_range.get((int(a), int(b)), -1) for a, b in zip(wlo, whi)], np.int64)
    tables = np.zeros((n_wide, 2), np.uint32)
    for e in range(16):
        x = root.copy()
        live = x >= 0
        entry = np.zeros(n_wide, np.uint32)
        for _ in range(8):
            xs = np.where(live, x, 0)
            ow = order[xs]
            live &= ((ow >> (16 + e)) & 1) != 0
            child = np.where((ow >> e) & 1 != 0, brefs[xs, 0], brefs[xs, 1]).astype(np.int64)
            match = as_bin == child[:, None]
            found = live & match.any(axis=1)
            slot = 7 - np.argmax(match[:, ::-1], axis=1)      # the last matching slot
            entry[found] = 8 | slot[found]
            live &= ~found & ((child & LEAF) == 0) & (child < n_trav)
            x = np.where(live, child, x)
        tables[:, e >> 3] |= entry << np.uint32(4 * (e & 7))
    return tables


def test_learned_wide_order(rtq_ctx, occ, torch_mod):
    """Learned order on 8-wide nodes: the first-child tables against their
    restatement from the binary order words, hit records against the plain
    8-wide and binary kernels (trained, NaN and saturated weights), one launch
    per set, batches (also with plain items, items of a binary-only scene, and
    graph replays), and what clears the tables."""
    torch = torch_mod
    n = occ["n"]
    wsc = rtq_ctx.upload(occ["tris"], occ["models"])
    wsc.build(16, 4, 0, node_width=8)
    try:
        assert wsc.info().n_wide_nodes > 0
        assert not wsc.export_wide()[:, 24:26].any()           # tables empty after a build
        with pytest.raises(RtqError):
            trace(wsc, occ["diff"], n, RTQ_TRACE_LEARNED)
        plain_d = trace(wsc, occ["diff"], n, RTQ_TRACE_BINARY)
        plain_s = trace(wsc, occ["shad"], 4 * n, RTQ_TRACE_ANY | RTQ_TRACE_BINARY)
        assert hits_equal(trace(wsc, occ["diff"], n, 0), plain_d)
        _, wide_v, wide_t = trace(wsc, occ["shad"], 4 * n, RTQ_TRACE_ANY, stats=True)
        _, bin_v, _ = trace(wsc, occ["shad"], 4 * n, RTQ_TRACE_ANY | RTQ_TRACE_BINARY, stats=True)
        model = OrderModel(rtq_ctx, seed=4)
        samples, _ = model.log(wsc, occ["shad"], 4 * n, True, stride=3)
        model.train(samples, epochs=3)
        garbage = torch.full((ORDER_PARAMS,), float("nan"), device="cuda")
        huge = dev(rand_params(5) * np.float32(1e30), torch)
        for params, conf, label in ((model.params, 0.5, "trained"), (model.params, 0.8, "trained 0.8"),
                                    (garbage, 0.5, "nan"), (huge, 0.5, "huge")):
            torch.cuda.synchronize()
            wsc.order_apply(params.data_ptr(), conf, conf)
            trav, _ = wsc.export_trav()
            wide = wsc.export_wide()
            want = wide_order_tables(trav, wide)
            assert np.array_equal(wide[:, 24:26], want), label
            entries = (want[:, :, None] >> (4 * np.arange(8, dtype=np.uint32))) & 15
            if label == "nan":
                assert not entries.any()
            else:
                # training is not deterministic (log order): a loose floor
                assert (entries & 8).mean() > (0.05 if conf == 0.5 else 0.01), label
            # a decided slot is a valid child
            valid = (wide[:, 3] >> 24)[:, None, None] >> (entries & 7)
            assert np.all((valid & 1)[(entries & 8) != 0] == 1)
            got_d = trace(wsc, occ["diff"], n, RTQ_TRACE_LEARNED)
            assert hits_equal(got_d, plain_d), f"{label}: {first_diff(got_d, plain_d)}"
            got_s, lv, lt = trace(wsc, occ["shad"], 4 * n, RTQ_TRACE_ANY | RTQ_TRACE_LEARNED, stats=True)
            assert np.array_equal(got_s["id"] >= 0, plain_s["id"] >= 0), label
            # traced on the 8-wide nodes: about as many node visits as the plain
            # 8-wide kernel, far fewer than on binary nodes
            assert lv.sum() < 0.7 * bin_v.sum(), label
            if label == "trained":
                print(f"8-wide shadow work: nearest first {(wide_v + wide_t).sum()}, learned {(lv + lt).sum()} "
                      f"({(lv + lt).sum() / (wide_v + wide_t).sum():.3f}x)")
            # learned order on binary nodes, same records
            assert hits_equal(trace(wsc, occ["diff"], n, RTQ_TRACE_LEARNED | RTQ_TRACE_BINARY), plain_d), label
        wsc.order_apply(model.params.data_ptr(), 0.5, 0.5)
        hd, hs = devmem.empty(n, HIT_DTYPE), devmem.empty(4 * n, HIT_DTYPE)
        for flags_s in (RTQ_TRACE_ANY | RTQ_TRACE_LEARNED, RTQ_TRACE_ANY):   # learned + learned, learned + plain
            items = rtq_ctx.items([(wsc, occ["diff"].data_ptr(), n, hd.data_ptr(), RTQ_TRACE_LEARNED),
                                   (wsc, occ["shad"].data_ptr(), 4 * n, hs.data_ptr(), flags_s)])
            rtq_ctx.trace_batch(items)
            rtq_ctx.synchronize()
            assert hits_equal(devmem.to_host(hd, HIT_DTYPE), plain_d), flags_s
            assert np.array_equal(devmem.to_host(hs, HIT_DTYPE)["id"] >= 0, plain_s["id"] >= 0), flags_s
        # graph replays, with eager work and a clear of the outputs between them
        items = rtq_ctx.items([(wsc, occ["diff"].data_ptr(), n, hd.data_ptr(), RTQ_TRACE_LEARNED),
                               (wsc, occ["shad"].data_ptr(), 4 * n, hs.data_ptr(), RTQ_TRACE_ANY | RTQ_TRACE_LEARNED)])
        g = rtq_ctx.batch(items)
        try:
            for rep in range(3):
                hd.fill_(255)
                hs.fill_(255)
                torch.cuda.synchronize()
                g.launch()
                rtq_ctx.synchronize()
                assert hits_equal(devmem.to_host(hd, HIT_DTYPE), plain_d), rep
                assert np.array_equal(devmem.to_host(hs, HIT_DTYPE)["id"] >= 0, plain_s["id"] >= 0), rep
                assert hits_equal(trace(wsc, occ["diff"], 1000, 0), plain_d[:1000])
        finally:
            g.destroy()
        # with an item of a binary-only scene the batch traces binary nodes
        # with the learned order (occ's scene holds no order bits here)
        sc = occ["sc"]
        sc.order_apply(None)
        h2 = devmem.empty(n, HIT_DTYPE)
        both = rtq_ctx.items([(wsc, occ["diff"].data_ptr(), n, hd.data_ptr(), RTQ_TRACE_LEARNED),
                              (sc, occ["diff"].data_ptr(), n, h2.data_ptr(), 0)])
        rtq_ctx.trace_batch(both)
        rtq_ctx.synchronize()
        assert hits_equal(devmem.to_host(hd, HIT_DTYPE), plain_d)
        assert hits_equal(devmem.to_host(h2, HIT_DTYPE), plain_d)
        # clearing the order empties the tables; a rebuild does as well
        wsc.order_apply(None)
        assert not wsc.export_wide()[:, 24:26].any()
        with pytest.raises(RtqError):
            trace(wsc, occ["diff"], n, RTQ_TRACE_LEARNED)
        wsc.order_apply(model.params.data_ptr(), 0.5, 0.5)
        wsc.build(16, 4, 0, node_width=8)
        assert not wsc.export_wide()[:, 24:26].any()
        with pytest.raises(RtqError):
            trace(wsc, occ["diff"], n, RTQ_TRACE_LEARNED)
        # a refit keeps the 8-wide nodes and empties their tables; the order
        # applied again writes them from the kept node mapping
        wsc.order_apply(model.params.data_ptr(), 0.5, 0.5)
        wsc.update(occ["tris"])
        wsc.refit()
        assert not wsc.export_wide()[:, 24:26].any()
        with pytest.raises(RtqError):
            trace(wsc, occ["diff"], n, RTQ_TRACE_LEARNED)
        wsc.order_apply(model.params.data_ptr(), 0.5, 0.5)
        trav, _ = wsc.export_trav()
        wide = wsc.export_wide()
        assert np.array_equal(wide[:, 24:26], wide_order_tables(trav, wide))
        got_s = trace(wsc, occ["shad"], 4 * n, RTQ_TRACE_ANY | RTQ_TRACE_LEARNED)
        assert np.array_equal(got_s["id"] >= 0, plain_s["id"] >= 0)
        assert hits_equal(trace(wsc, occ["diff"], n, RTQ_TRACE_LEARNED), plain_d)
    finally:
        wsc.release()


def test_learned_order_gain(rtq_ctx, occ, torch_mod):
    """A model trained on the shadow-ray logs of one light set needs fewer node
    visits + triangle tests than the nearer-child order on the shadow rays of
    another light set, with the same occlusion flags."""
    sc = occ["sc"]
    n = occ["n"]
    model = OrderModel(rtq_ctx, seed=2)
    samples, k = model.log(sc, occ["shad"], 4 * n, True, stride=2)
    model.train(samples, epochs=6, batch=1024, lr=2e-3)
    # held-out rays: other lights, same primary hits
    lights = np.array([[0.1, 2.0, 0.5], [0.9, 2.0, 0.5], [0.5, 2.0, 0.1], [0.5, 2.0, 0.9]], np.float32)
    phit = devmem.empty(n, HIT_DTYPE)
    sc.trace(occ["prim"].data_ptr(), n, phit.data_ptr())
    held = devmem.empty(4 * n, RAY_DTYPE)
    sc.raygen_shadow(occ["prim"].data_ptr(), phit.data_ptr(), n, lights, scenes.SECONDARY_EPS, held.data_ptr())
    rtq_ctx.synchronize()
    base_h, base_v, base_t = trace(sc, held, 4 * n, RTQ_TRACE_ANY | RTQ_TRACE_BINARY, stats=True)
    model.apply(sc, confidence_any=0.5)
    got_h, got_v, got_t = trace(sc, held, 4 * n, RTQ_TRACE_ANY | RTQ_TRACE_LEARNED, stats=True)
    assert np.array_equal(got_h["id"] >= 0, base_h["id"] >= 0)
    base = (base_v + base_t).sum()
    got = (got_v + got_t).sum()
    print(f"shadow work: nearer-child {base}, learned {got} ({got / base:.3f}x)")
    assert got < 0.97 * base


def test_order_model_handle(rtq_ctx, occ, tmp_path):
    """rtq_order_model_*: the model owning its device memory, end to end."""
    sc, n = occ["sc"], occ["n"]
    plain_s = trace(sc, occ["shad"], 4 * n, RTQ_TRACE_ANY | RTQ_TRACE_BINARY)
    plain_d = trace(sc, occ["diff"], n, RTQ_TRACE_BINARY)
    m = OrderModelHandle(rtq_ctx, seed=4, sample_capacity=50_000)
    try:
        with pytest.raises(RtqError):
            m.train(TrainParams(1, 1024, 2e-3, 0.9, 0.999, 0, 1))       # nothing logged yet
        k1 = m.log(sc, occ["shad"].data_ptr(), 4 * n, RTQ_TRACE_ANY, 3)
        k2 = m.log(sc, occ["shad"].data_ptr(), 4 * n, RTQ_TRACE_ANY, 3)
        assert 0 < k1 and k1 + k2 == m.info()[0] <= 50_000
        assert k1 + k2 == 50_000 or k2 == k1                          # a full store keeps what fits
        with pytest.raises(RtqError):
            m.train(TrainParams(1, 100, 2e-3, 0.9, 0.999, 0, 1))        # batch not a multiple of 64
        losses = m.train(TrainParams(3, 1024, 2e-3, 0.9, 0.999, 0, 1))
        assert losses[-1] < losses[0] and m.info()[1] > 0
        m.apply(sc, 1.0, 0.5)
        got_s = trace(sc, occ["shad"], 4 * n, RTQ_TRACE_ANY | RTQ_TRACE_LEARNED)
        assert np.array_equal(got_s["id"] >= 0, plain_s["id"] >= 0)
        assert hits_equal(trace(sc, occ["diff"], n, RTQ_TRACE_LEARNED), plain_d)
        path = tmp_path / "model.bin"
        m.save(path)
        other = OrderModelHandle(rtq_ctx, seed=9, sample_capacity=0)
        try:
            other.load(path)
            assert other.info() == (0, m.info()[1])
            with pytest.raises(RtqError):
                other.train(TrainParams(1, 1024, 2e-3, 0.9, 0.999, 0, 1))
            assert other.log(sc, occ["shad"].data_ptr(), 4 * n, RTQ_TRACE_ANY, 3) == 0   # no room
        finally:
            other.destroy()
        m.clear_samples()
        assert m.info()[0] == 0
    finally:
        m.destroy()
        sc.order_apply(None)


def test_learned_without_decisions_traces_like_plain_on_wide_nodes(rtq_ctx, torch_mod):
    """Order words applied with both modes turned off (confidence 1): learned
    rays trace as plain rays do, on the scene's 8-wide nodes, bit for bit; with
    any-hit decisions on, closest-hit rays still match and a batch of both
    modes runs on binary nodes with the same results."""
    torch = torch_mod
    tris, models = scenes.soup_scene(20000, 4)
    sc = rtq_ctx.upload(tris, models)
    try:
        sc.build(16, 4, 0, node_width=8)
        rays = devmem.to_device(scenes.random_rays(8192, 6, (-0.1, -0.1, -0.1), (1.1, 1.1, 1.1)))
        n = 8192
        plain_c, plain_a = trace(sc, rays, n, 0), trace(sc, rays, n, RTQ_TRACE_ANY)
        params = dev(rand_params(21), torch)
        torch.cuda.synchronize()
        sc.order_apply(params.data_ptr(), 1.0, 1.0)
        assert sc.info().order_modes == 0
        assert hits_equal(trace(sc, rays, n, RTQ_TRACE_LEARNED), plain_c)
        assert hits_equal(trace(sc, rays, n, RTQ_TRACE_ANY | RTQ_TRACE_LEARNED), plain_a)
        sc.order_apply(params.data_ptr(), 0.5, 1.0)
        assert sc.info().order_modes == 2
        assert hits_equal(trace(sc, rays, n, RTQ_TRACE_LEARNED), plain_c)
        got_a = trace(sc, rays, n, RTQ_TRACE_ANY | RTQ_TRACE_LEARNED)
        assert np.array_equal(got_a["id"] >= 0, plain_a["id"] >= 0)
        hc, ha = devmem.empty(n, HIT_DTYPE), devmem.empty(n, HIT_DTYPE)
        rtq_ctx.trace_batch(rtq_ctx.items([(sc, rays.data_ptr(), n, hc.data_ptr(), RTQ_TRACE_LEARNED),
                                           (sc, rays.data_ptr(), n, ha.data_ptr(), RTQ_TRACE_ANY | RTQ_TRACE_LEARNED)]))
        rtq_ctx.synchronize()
        assert hits_equal(devmem.to_host(hc, HIT_DTYPE), plain_c)
        assert np.array_equal(devmem.to_host(ha, HIT_DTYPE)["id"] >= 0, plain_a["id"] >= 0)
    finally:
        sc.release()
