// Drives rtq::BVH (include/rtq/bvh.hpp) the way the reference's Scene drives
// cr::BVH (scene.cpp:146-171): build from a triangle vector and a model vector,
// read _InternalStruct, flatten to pre-order, trace.  Then the frame-batch,
// autotuner and order-model wrappers on device buffers (HIP runtime API).
// Prints "wrapper_check OK" and exits 0 when every check holds.
#include <hip/hip_runtime_api.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "rtq/bvh.hpp"
#include "rtq/cr_bvh.hpp"

#define CHECK(c)                                                         \
    do {                                                                 \
        if (!(c)) {                                                      \
            std::fprintf(stderr, "check failed: %s (line %d)\n", #c, __LINE__); \
            return 1;                                                    \
        }                                                                \
    } while (0)

// Layout twins of the reference structs (glm-free), as a caller would pass them.
struct Vec4 { float x, y, z, w; };
struct TriangleGPU { Vec4 p0, p1, p2; alignas(16) uint32_t model_id; };
struct MeshModelGPU { float m[16]; uint32_t material_id; };
// what the reference's triangle.hpp / mesh.hpp give a caller
namespace cr {
using TriangleGPU = ::TriangleGPU;
using MeshModelGPU = ::MeshModelGPU;
}

// The reference Scene's use of cr::BVH (scene.cpp:148, 189-208), written
// against the drop-in header: build in the constructor, then a recursive
// pre-order walk over _InternalStruct from the root cluster 2N - 2.
struct ReferenceStyleScene {
    uint32_t _NbTriangles = 0;
    cr::BVH_Ptr _BVH;

    void walk(std::vector<cr::BVH_NodeGPU>& out, cr::BVH_Ptr bvh, uint32_t id) const {
        cr::BVH_NodeGPU node = bvh->_InternalStruct._Clusters[id].value();
        const uint32_t at = static_cast<uint32_t>(out.size());
        out.push_back(node);
        if (bvh->_InternalStruct._IsLeaf[id]) return;
        out[at]._LeftChild = static_cast<uint32_t>(out.size());
        walk(out, bvh, bvh->_InternalStruct._LeftChild[id].value());
        out[at]._RightChild = static_cast<uint32_t>(out.size());
        walk(out, bvh, bvh->_InternalStruct._RightChild[id].value());
    }
    std::vector<cr::BVH_NodeGPU> nodesForGpu(cr::BVH_Ptr bvh) const {
        std::vector<cr::BVH_NodeGPU> out;
        walk(out, bvh, 2 * _NbTriangles - 2);
        return out;
    }
    void bind(const std::vector<cr::TriangleGPU>& triangleGPU, const std::vector<cr::MeshModelGPU>& modelsGPU) {
        _BVH = cr::BVH_Ptr(new cr::BVH(_NbTriangles, triangleGPU, modelsGPU));
    }
};

int main() {
    // a 32 x 32 flat grid at y = 0 (2 x 31^2 triangles)
    const int res = 32;
    std::vector<TriangleGPU> tris;
    auto P = [&](int i, int j) { return Vec4{i / float(res - 1), 0.0f, j / float(res - 1), 1.0f}; };
    for (int i = 0; i + 1 < res; ++i)
        for (int j = 0; j + 1 < res; ++j) {
            tris.push_back({P(i, j), P(i, j + 1), P(i + 1, j + 1), 0});
            tris.push_back({P(i, j), P(i + 1, j + 1), P(i + 1, j), 0});
        }
    MeshModelGPU identity{};
    for (int k = 0; k < 4; ++k) identity.m[k * 5] = 1.0f;
    std::vector<MeshModelGPU> models{identity};
    const uint32_t n = static_cast<uint32_t>(tris.size());

    // ---- the reference's call pattern through cr::BVH on the default context
    {
        ReferenceStyleScene scene;
        scene._NbTriangles = n;
        scene.bind(tris, models);
        const std::vector<cr::BVH_NodeGPU> walked = scene.nodesForGpu(scene._BVH);
        const std::vector<cr::BVH_NodeGPU> exported = scene._BVH->preorder();
        CHECK(sizeof(cr::BVH_NodeGPU) == 48 && walked.size() == 2 * n - 1 && exported.size() == walked.size());
        for (size_t i = 0; i < walked.size(); ++i) {
            const cr::BVH_NodeGPU &a = walked[i], &b = exported[i];
            CHECK(a._TriangleId == b._TriangleId && a._LeftChild == b._LeftChild && a._RightChild == b._RightChild);
            for (int k = 0; k < 3; ++k)
                CHECK(a._BoundingBox._Min[k] == b._BoundingBox._Min[k] && a._BoundingBox._Max[k] == b._BoundingBox._Max[k]);
        }
        CHECK(walked[0]._LeftChild == 1 && scene._BVH->_InternalStruct._TriangleIndices.size() == n);
        CHECK(cr::AABB::getSurfaceArea(walked[0]._BoundingBox) == 2.0f);   // the unit square in y = 0
        CHECK(walked[0]._BoundingBox._Min.x == 0.0f && walked[0]._BoundingBox._Max[2] == 1.0f);
        for (const rtq_hit& h : scene._BVH->gpu().trace(std::vector<rtq_ray>{{{0.3f, 1.0f, 0.6f}, 0.0f,
                                                                               {0.0f, -1.0f, 0.0f}, INFINITY}}))
            CHECK(h.id >= 0 && h.t == 1.0f);
    }

    auto ctx = cr::defaultContext();
    rtq::BVH bvh(ctx, n, tris, models);
    const rtq::BVHParams& p = bvh._InternalStruct;
    CHECK(p._NbTriangles == n);
    CHECK(p._Clusters.size() == 2 * n - 1);
    // leaves first, each with its sorted triangle; the root is the last cluster
    for (uint32_t c = 0; c < n; ++c) {
        CHECK(p._IsLeaf[c].has_value());
        CHECK(p._Clusters[c]->triangle_id == p._TriangleIndices[c]);
    }
    CHECK(!p._Parent[2 * n - 2].has_value());
    for (uint32_t c = n; c < 2 * n - 1; ++c) {
        CHECK(!p._IsLeaf[c].has_value());
        CHECK(*p._Parent[*p._LeftChild[c]] == c && *p._Parent[*p._RightChild[c]] == c);
    }
    // pre-order array: root 0, left child adjacent
    std::vector<rtq_bvh_node> nodes = bvh.preorder();
    CHECK(nodes.size() == 2 * n - 1);
    uint32_t leaves = 0;
    for (size_t i = 0; i < nodes.size(); ++i) {
        if (nodes[i].left_child == 0 && nodes[i].right_child == 0) ++leaves;
        else CHECK(nodes[i].left_child == i + 1);
    }
    CHECK(leaves == n);
    // straight-down rays through cell centres all hit, at t = 1
    std::vector<rtq_ray> rays;
    for (int i = 0; i + 1 < res; ++i)
        for (int j = 0; j + 1 < res; ++j) {
            rtq_ray r{};
            r.o[0] = (i + 0.25f) / float(res - 1);
            r.o[1] = 1.0f;
            r.o[2] = (j + 0.5f) / float(res - 1);
            r.d[1] = -1.0f;
            r.tmax = INFINITY;
            rays.push_back(r);
        }
    std::vector<rtq_hit> hits = bvh.trace(rays);
    for (const rtq_hit& h : hits) CHECK(h.id >= 0 && h.t == 1.0f);
    std::vector<rtq_hit> occ = bvh.trace(rays, RTQ_TRACE_ANY);
    for (const rtq_hit& h : occ) CHECK(h.id >= 0);
    // errors surface as exceptions carrying the status
    bool threw = false;
    try {
        rtq_build_params bad{};
        rtq::BVH b2(ctx, n, tris, models, &bad);   // radius 0 is invalid
    } catch (const rtq::Error& e) {
        threw = (e.status == RTQ_EINVAL);
    }
    CHECK(threw);

    // ---- scene edit: the grid lifted to y = 0.5 by its model matrix, refit --
    {
        std::vector<MeshModelGPU> lifted = models;
        lifted[0].m[13] = 0.5f;                      // column 3, row 1: translation y
        bvh.update(std::vector<TriangleGPU>{}, lifted);
        CHECK(p._Clusters[2 * n - 2]->box.min[1] == 0.5f && p._Clusters[2 * n - 2]->box.max[1] == 0.5f);
        for (const rtq_hit& h : bvh.trace(rays)) CHECK(h.id >= 0 && h.t == 0.5f);
        // shading of those hits with one material (getColor: base + colour)
        rtq_hit* d_h = nullptr;
        rtq_material* d_m = nullptr;
        float* d_px = nullptr;
        std::vector<rtq_hit> lifted_hits = bvh.trace(rays);
        const rtq_material mat{{0.2f, 0.3f, 0.1f, 1.0f}};
        CHECK(hipMalloc(&d_h, lifted_hits.size() * sizeof(rtq_hit)) == hipSuccess);
        CHECK(hipMalloc(&d_m, sizeof mat) == hipSuccess);
        CHECK(hipMalloc(&d_px, lifted_hits.size() * 4 * sizeof(float)) == hipSuccess);
        CHECK(hipMemcpy(d_h, lifted_hits.data(), lifted_hits.size() * sizeof(rtq_hit), hipMemcpyHostToDevice) ==
              hipSuccess);
        CHECK(hipMemcpy(d_m, &mat, sizeof mat, hipMemcpyHostToDevice) == hipSuccess);
        bvh.shade(d_h, lifted_hits.size(), d_m, 1, d_px);
        CHECK(rtq_synchronize(ctx->get()) == RTQ_OK);
        std::vector<float> px(lifted_hits.size() * 4);
        CHECK(hipMemcpy(px.data(), d_px, px.size() * sizeof(float), hipMemcpyDeviceToHost) == hipSuccess);
        for (size_t i = 0; i < lifted_hits.size(); ++i)
            CHECK(px[4 * i] == 0.2f && px[4 * i + 1] == 0.3f && px[4 * i + 2] == 0.1f && px[4 * i + 3] == 2.0f);
        (void)hipFree(d_h);
        (void)hipFree(d_m);
        (void)hipFree(d_px);
        bvh.update(std::vector<TriangleGPU>{}, models);   // back to y = 0 for the checks below

        // the grid's vertices raised to y = 0.5 in device memory (as an
        // animation kernel would write them), then back from the host
        std::vector<TriangleGPU> raised = tris;
        for (TriangleGPU& t : raised) {
            t.p0.y += 0.5f;
            t.p1.y += 0.5f;
            t.p2.y += 0.5f;
        }
        TriangleGPU* d_tris = nullptr;
        CHECK(hipMalloc(&d_tris, raised.size() * sizeof(TriangleGPU)) == hipSuccess);
        CHECK(hipMemcpy(d_tris, raised.data(), raised.size() * sizeof(TriangleGPU), hipMemcpyHostToDevice) ==
              hipSuccess);
        bvh.update_device(reinterpret_cast<const rtq_triangle*>(d_tris), n);
        for (const rtq_hit& h : bvh.trace(rays)) CHECK(h.id >= 0 && h.t == 0.5f);
        threw = false;
        try {
            bvh.update_device(reinterpret_cast<const rtq_triangle*>(raised.data()), n);   // host memory
        } catch (const rtq::Error& e) {
            threw = (e.status == RTQ_EINVAL);
        }
        CHECK(threw);
        (void)hipFree(d_tris);
        bvh.update(tris, std::vector<MeshModelGPU>{});
        for (const rtq_hit& h : bvh.trace(rays)) CHECK(h.id >= 0 && h.t == 1.0f);
    }

    // ---- frame batch + autotuner + learned order on device buffers --------
    // a second scene: the grid plus a cloud of small occluders above it
    std::vector<TriangleGPU> occ_tris = tris;
    uint64_t st = 12345;
    auto u01 = [&]() { st = st * 6364136223846793005ull + 1442695040888963407ull; return (st >> 40) / 16777216.0f; };
    for (int k = 0; k < 4000; ++k) {
        const float cx = u01(), cy = 0.05f + 0.5f * u01(), cz = u01();
        occ_tris.push_back({Vec4{cx - 0.01f, cy, cz - 0.01f, 1.0f}, Vec4{cx + 0.01f, cy, cz, 1.0f},
                            Vec4{cx, cy + 0.01f, cz + 0.01f, 1.0f}, 0});
    }
    rtq::BVH occ_bvh(ctx, static_cast<uint32_t>(occ_tris.size()), occ_tris, models);
    // shadow rays from the grid towards a light above it, and closest-hit rays
    std::vector<rtq_ray> shadow, closest;
    for (int k = 0; k < 20000; ++k) {
        rtq_ray r{};
        r.o[0] = u01(); r.o[1] = 1e-4f; r.o[2] = u01();
        r.d[0] = 0.5f - r.o[0]; r.d[1] = 2.0f; r.d[2] = 0.5f - r.o[2];
        r.tmax = 1.0f;
        shadow.push_back(r);
        rtq_ray c = r;
        c.o[1] = 1.5f; c.d[0] = 0.0f; c.d[1] = -1.0f; c.d[2] = 0.0f; c.tmax = INFINITY;
        closest.push_back(c);
    }
    const size_t n_s = shadow.size(), n_c = closest.size();
    rtq_ray *d_shadow = nullptr, *d_closest = nullptr;
    rtq_hit *d_hs = nullptr, *d_hc = nullptr, *d_hs2 = nullptr, *d_hc2 = nullptr;
    CHECK(hipMalloc(&d_shadow, n_s * sizeof(rtq_ray)) == hipSuccess);
    CHECK(hipMalloc(&d_closest, n_c * sizeof(rtq_ray)) == hipSuccess);
    CHECK(hipMalloc(&d_hs, n_s * sizeof(rtq_hit)) == hipSuccess && hipMalloc(&d_hs2, n_s * sizeof(rtq_hit)) == hipSuccess);
    CHECK(hipMalloc(&d_hc, n_c * sizeof(rtq_hit)) == hipSuccess && hipMalloc(&d_hc2, n_c * sizeof(rtq_hit)) == hipSuccess);
    CHECK(hipMemcpy(d_shadow, shadow.data(), n_s * sizeof(rtq_ray), hipMemcpyHostToDevice) == hipSuccess);
    CHECK(hipMemcpy(d_closest, closest.data(), n_c * sizeof(rtq_ray), hipMemcpyHostToDevice) == hipSuccess);
    auto hits_of = [&](const rtq_hit* d, size_t n) {
        std::vector<rtq_hit> h(n);
        (void)hipMemcpy(h.data(), d, n * sizeof(rtq_hit), hipMemcpyDeviceToHost);
        return h;
    };
    // reference results: one trace per set
    ctx->check(rtq_trace(ctx->get(), occ_bvh.scene_id(), d_shadow, n_s, RTQ_TRACE_ANY, d_hs, nullptr));
    ctx->check(rtq_trace(ctx->get(), occ_bvh.scene_id(), d_closest, n_c, RTQ_TRACE_CLOSEST, d_hc, nullptr));
    ctx->check(rtq_synchronize(ctx->get()));
    const std::vector<rtq_hit> ref_s = hits_of(d_hs, n_s), ref_c = hits_of(d_hc, n_c);
    size_t occluded = 0;
    for (const rtq_hit& h : ref_s) occluded += h.id >= 0;
    CHECK(occluded > n_s / 10 && occluded < n_s);
    // both sets in one graph-replayed launch with autotuned settings
    std::vector<rtq_batch_item> items = {
        {occ_bvh.scene_id(), RTQ_TRACE_ANY, d_shadow, d_hs2, n_s},
        {occ_bvh.scene_id(), RTQ_TRACE_CLOSEST, d_closest, d_hc2, n_c},
    };
    const rtq_tune_result tuned = rtq::autotune(*ctx, items, 1);
    CHECK(tuned.n_tried > 1 && tuned.best_ms <= tuned.default_ms);
    {
        rtq::FrameBatch frame(ctx, items);
        for (int rep = 0; rep < 3; ++rep) {
            CHECK(hipMemset(d_hc2, 0, n_c * sizeof(rtq_hit)) == hipSuccess);
            frame.launch();
            ctx->check(rtq_synchronize(ctx->get()));
            const std::vector<rtq_hit> got_c = hits_of(d_hc2, n_c), got_s = hits_of(d_hs2, n_s);
            CHECK(std::memcmp(got_c.data(), ref_c.data(), n_c * sizeof(rtq_hit)) == 0);
            for (size_t i = 0; i < n_s; ++i) CHECK((got_s[i].id >= 0) == (ref_s[i].id >= 0));
        }
    }
    // learned order: log, train, apply, trace with RTQ_TRACE_LEARNED, checkpoint
    rtq::OrderModel model(ctx, 7, 1u << 20);
    const uint64_t logged = model.log(occ_bvh, d_shadow, n_s, RTQ_TRACE_ANY, 1);
    CHECK(logged > 0 && model.samples() == logged);
    const rtq_train_params tp{3, 1024, 2e-3f, 0.9f, 0.999f, 0, 7};
    const std::vector<float> loss = model.train(tp);
    CHECK(loss.size() == 3 && loss[2] < loss[0] && model.steps() > 0);
    model.apply(occ_bvh, 1.0f, 0.5f);
    ctx->check(rtq_trace(ctx->get(), occ_bvh.scene_id(), d_shadow, n_s, RTQ_TRACE_ANY | RTQ_TRACE_LEARNED, d_hs2,
                         nullptr));
    ctx->check(rtq_trace(ctx->get(), occ_bvh.scene_id(), d_closest, n_c, RTQ_TRACE_LEARNED, d_hc2, nullptr));
    ctx->check(rtq_synchronize(ctx->get()));
    {
        const std::vector<rtq_hit> got_c = hits_of(d_hc2, n_c), got_s = hits_of(d_hs2, n_s);
        CHECK(std::memcmp(got_c.data(), ref_c.data(), n_c * sizeof(rtq_hit)) == 0);
        for (size_t i = 0; i < n_s; ++i) CHECK((got_s[i].id >= 0) == (ref_s[i].id >= 0));
    }
    const char* ckpt = "wrapper_check_order.bin";
    model.save(ckpt);
    rtq::OrderModel other(ctx, 99, 16);
    other.load(ckpt);
    CHECK(other.steps() == model.steps());
    std::remove(ckpt);
    (void)hipFree(d_shadow);
    (void)hipFree(d_closest);
    (void)hipFree(d_hs);
    (void)hipFree(d_hs2);
    (void)hipFree(d_hc);
    (void)hipFree(d_hc2);

    const rtq_bvh_info info = bvh.info();
    std::printf("wrapper_check OK: %u triangles, %u traversal nodes, depth %u, build %.3f ms\n", info.n_tris,
                info.n_trav_nodes, info.depth, info.build_ms);
    return 0;
}
