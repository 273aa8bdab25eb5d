// Drives rtq::BVH (include/rtq/bvh.hpp) the way the reference's Scene drives
// cr::BVH (scene.cpp:146-171): build from a triangle vector and a model vector,
// read _InternalStruct, flatten to pre-order, trace.  Prints "wrapper_check OK"
// and exits 0 when every check holds.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "rtq/bvh.hpp"

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

    auto ctx = std::make_shared<rtq::Context>(0);
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
    const rtq_bvh_info info = bvh.info();
    std::printf("wrapper_check OK: %u triangles, %u traversal nodes, depth %u, build %.3f ms\n", info.n_tris,
                info.n_trav_nodes, info.depth, info.build_ms);
    return 0;
}
