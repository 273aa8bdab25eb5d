// Drop-in for the reference's srcCommon/scene/geometry/bvh.hpp: the cr::BVH
// API with the GPU build behind it.
//
// A caller written against the reference (srcOpenGL/scene/scene.cpp:148,
// 189-208) compiles with only its include changed to this header:
//   - cr::BVH(uint32_t nbTriangles, const std::vector<TriangleGPU>&,
//     const std::vector<MeshModelGPU>&) builds in its constructor
//     (bvh.hpp:89-91, bvh.cpp:11-24), here on the GPU of the default context;
//   - cr::BVH_Ptr (bvh.hpp:16);
//   - _InternalStruct is a BVH_Params (bvh.hpp:44-63) with optional-valued
//     _Clusters / _IsLeaf / _Parent / _LeftChild / _RightChild and
//     _TriangleIndices, sized 2N-1 and N instead of the 65,536 cap;
//   - cr::BVH_NodeGPU (bvh.hpp:36-42) with _BoundingBox, _TriangleId,
//     _LeftChild, _RightChild and cr::AABB_GPU (bvh.hpp:22-26) with _Min and
//     _Max, byte for byte the reference's 48-B and 32-B layouts.
// The triangle and model vectors may hold the reference's own glm structs:
// any element type with the 64-B TriangleGPU / 68-B MeshModelGPU layout is
// accepted.  This header defines neither of those types, so it can sit next
// to the reference's triangle.hpp and mesh.hpp.
//
// Differences from the reference, each deliberate:
//   - _Min / _Max are a glm-free three-float type (x, y, z, operator[]) with
//     glm::vec3's layout;
//   - BVH_Params does not keep copies of the input vectors
//     (_UnsortedTriangles, _MeshesInTheScene), which the reference only reads
//     while building;
//   - failures throw rtq::Error where the reference calls exit()
//     (errorHandler.cpp:13-45).
#pragma once

#include <cmath>
#include <cstddef>
#include <cstdint>
#include <memory>
#include <mutex>
#include <optional>
#include <vector>

#include "bvh.hpp"

namespace cr {

class BVH;
using BVH_Ptr = std::shared_ptr<BVH>;

enum Axis { X, Y, Z };

// Three floats laid out as glm::vec3.
struct Vec3 {
    float x, y, z;
    float& operator[](int i) { return i == 0 ? x : (i == 1 ? y : z); }
    float operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
};

struct AABB_GPU {
    Vec3 _Min = {INFINITY, INFINITY, INFINITY};
    alignas(16) Vec3 _Max = {-INFINITY, -INFINITY, -INFINITY};
};

struct BVH_NodeGPU {
    AABB_GPU _BoundingBox;
    uint32_t _TriangleId;
    uint32_t _LeftChild;
    uint32_t _RightChild;
    // both children 0: a leaf
};

static_assert(sizeof(AABB_GPU) == sizeof(rtq_aabb) && offsetof(AABB_GPU, _Max) == offsetof(rtq_aabb, max),
              "cr::AABB_GPU is 32 B with _Max at 16");
static_assert(sizeof(BVH_NodeGPU) == sizeof(rtq_bvh_node) &&
                  offsetof(BVH_NodeGPU, _TriangleId) == offsetof(rtq_bvh_node, triangle_id) &&
                  offsetof(BVH_NodeGPU, _LeftChild) == offsetof(rtq_bvh_node, left_child) &&
                  offsetof(BVH_NodeGPU, _RightChild) == offsetof(rtq_bvh_node, right_child),
              "cr::BVH_NodeGPU is 48 B, fields at 32/36/40");

// The box helpers of the reference's AABB class (bvh.cpp:374-413).
class AABB {
  public:
    static float getDiagonal(const AABB_GPU& b) {
        const float dx = b._Max.x - b._Min.x, dy = b._Max.y - b._Min.y, dz = b._Max.z - b._Min.z;
        return std::sqrt(dx * dx + dy * dy + dz * dz);
    }
    // 2 (dx dy + dy dz + dz dx), summed in that order
    static float getSurfaceArea(const AABB_GPU& b) {
        const float dx = b._Max.x - b._Min.x, dy = b._Max.y - b._Min.y, dz = b._Max.z - b._Min.z;
        return 2.0f * ((dx * dy + dy * dz) + dz * dx);
    }
    static AABB_GPU merge(const AABB_GPU& a, const AABB_GPU& b) {
        AABB_GPU m;
        for (int k = 0; k < 3; ++k) {
            m._Min[k] = b._Min[k] < a._Min[k] ? b._Min[k] : a._Min[k];
            m._Max[k] = a._Max[k] < b._Max[k] ? b._Max[k] : a._Max[k];
        }
        return m;
    }
};

// The device context the reference-signature constructor builds on: set one
// with setDefaultContext, or the first BVH creates one on HIP device 0.
inline std::shared_ptr<rtq::Context>& defaultContextSlot() {
    static std::shared_ptr<rtq::Context> ctx;
    return ctx;
}
inline std::mutex& defaultContextMutex() {
    static std::mutex m;
    return m;
}
inline void setDefaultContext(std::shared_ptr<rtq::Context> ctx) {
    std::lock_guard<std::mutex> lock(defaultContextMutex());
    defaultContextSlot() = std::move(ctx);
}
inline std::shared_ptr<rtq::Context> defaultContext() {
    std::lock_guard<std::mutex> lock(defaultContextMutex());
    std::shared_ptr<rtq::Context>& c = defaultContextSlot();
    if (!c) c = std::make_shared<rtq::Context>(0);
    return c;
}

struct BVH_Params {
    size_t _NbTriangles = 0;
    std::vector<std::optional<BVH_NodeGPU>> _Clusters;
    std::vector<std::optional<bool>> _IsLeaf;
    std::vector<std::optional<uint32_t>> _Parent;
    std::vector<std::optional<uint32_t>> _LeftChild;
    std::vector<std::optional<uint32_t>> _RightChild;
    std::vector<uint32_t> _TriangleIndices;
};

inline BVH_NodeGPU toNodeGPU(const rtq_bvh_node& n) {
    BVH_NodeGPU o;
    o._BoundingBox._Min = {n.box.min[0], n.box.min[1], n.box.min[2]};
    o._BoundingBox._Max = {n.box.max[0], n.box.max[1], n.box.max[2]};
    o._TriangleId = n.triangle_id;
    o._LeftChild = n.left_child;
    o._RightChild = n.right_child;
    return o;
}

class BVH {
  public:
    BVH_Params _InternalStruct = {};

    // bvh.hpp:89-91: the reference's constructor, on the default context
    template <class Tri, class Model>
    BVH(uint32_t nbTriangles, const std::vector<Tri>& unsortedTriangles, const std::vector<Model>& meshesInTheScene)
        : impl_(defaultContext(), nbTriangles, unsortedTriangles, meshesInTheScene) {
        copy_params();
    }
    // the same on a given context, with build parameters (radius, leaf size, node width)
    template <class Tri, class Model>
    BVH(std::shared_ptr<rtq::Context> ctx, uint32_t nbTriangles, const std::vector<Tri>& unsortedTriangles,
        const std::vector<Model>& meshesInTheScene, const rtq_build_params* params = nullptr)
        : impl_(std::move(ctx), nbTriangles, unsortedTriangles, meshesInTheScene, params) {
        copy_params();
    }
    BVH(const BVH&) = delete;
    BVH& operator=(const BVH&) = delete;

    // What Scene::getBVH_NodesToGPUData returns (scene.cpp:203-208), from the device.
    std::vector<BVH_NodeGPU> preorder() const {
        const std::vector<rtq_bvh_node> n = impl_.preorder();
        std::vector<BVH_NodeGPU> out;
        out.reserve(n.size());
        for (const rtq_bvh_node& x : n) out.push_back(toNodeGPU(x));
        return out;
    }
    // The GPU-side object: traces, edits, shading, scene id (rtq::BVH).
    rtq::BVH& gpu() { return impl_; }
    const rtq::BVH& gpu() const { return impl_; }

  private:
    void copy_params() {
        const rtq::BVHParams& p = impl_._InternalStruct;
        BVH_Params& q = _InternalStruct;
        q._NbTriangles = p._NbTriangles;
        q._Clusters.assign(p._Clusters.size(), std::nullopt);
        for (size_t c = 0; c < p._Clusters.size(); ++c) {
            if (p._Clusters[c]) q._Clusters[c] = toNodeGPU(*p._Clusters[c]);
        }
        q._IsLeaf = p._IsLeaf;
        q._Parent = p._Parent;
        q._LeftChild = p._LeftChild;
        q._RightChild = p._RightChild;
        q._TriangleIndices = p._TriangleIndices;
    }

    rtq::BVH impl_;
};

}  // namespace cr
