// rtq::BVH — C++ mirror of the reference's cr::BVH (bvh.hpp:84-124) over the C ABI.
//
// Same constructor arguments (triangle count, triangle vector, mesh-model vector,
// bvh.hpp:89-91), the build runs on the GPU in the constructor, and the result
// is exposed in the form of BVH_Params (bvh.hpp:44-63): optional-valued cluster,
// leaf, parent and child arrays plus the sorted triangle indices.  preorder()
// returns what Scene::getBVH_NodesToGPUData builds (scene.cpp:189-208).
//
// FrameBatch, autotune and OrderModel wrap the batched-launch, autotuner and
// learned-ordering entry points the same way.
//
// Triangle and model element types only need the reference byte layouts
// (64-B TriangleGPU, 68-B MeshModelGPU), so vectors of the reference's own glm
// structs can be passed unchanged.  Failures throw rtq::Error (the reference
// exits the process instead, errorHandler.cpp:13-45).
#pragma once

#include <cstdint>
#include <memory>
#include <optional>
#include <stdexcept>
#include <string>
#include <vector>

#include "rtq.h"

namespace rtq {

class Error : public std::runtime_error {
  public:
    Error(rtq_status s, const std::string& msg) : std::runtime_error(msg), status(s) {}
    rtq_status status;
};

// One device context; share it between BVHs on the same device.
class Context {
  public:
    explicit Context(int device = 0) {
        rtq_ctx* c = nullptr;
        rtq_status s = rtq_create(&c, device);
        if (s != RTQ_OK) throw Error(s, std::string("rtq_create: ") + rtq_status_string(s));
        ctx_.reset(c);
    }
    rtq_ctx* get() const { return ctx_.get(); }
    void check(rtq_status s) const {
        if (s != RTQ_OK) throw Error(s, rtq_last_error(ctx_.get()));
    }

  private:
    struct Del {
        void operator()(rtq_ctx* c) const { rtq_destroy(c); }
    };
    std::unique_ptr<rtq_ctx, Del> ctx_;
};

// The reference's BVH_Params, sized 2N-1 (N) instead of the 65,536 cap.
struct BVHParams {
    size_t _NbTriangles = 0;
    std::vector<std::optional<rtq_bvh_node>> _Clusters;
    std::vector<std::optional<bool>> _IsLeaf;
    std::vector<std::optional<uint32_t>> _Parent;
    std::vector<std::optional<uint32_t>> _LeftChild;
    std::vector<std::optional<uint32_t>> _RightChild;
    std::vector<uint32_t> _TriangleIndices;
};

class BVH {
  public:
    BVHParams _InternalStruct;

    template <class Tri, class Model>
    BVH(std::shared_ptr<Context> ctx, uint32_t nbTriangles, const std::vector<Tri>& unsortedTriangles,
        const std::vector<Model>& meshesInTheScene, const rtq_build_params* params = nullptr)
        : ctx_(std::move(ctx)) {
        static_assert(sizeof(Tri) == sizeof(rtq_triangle), "triangle records must be 64-B TriangleGPU");
        static_assert(sizeof(Model) == sizeof(rtq_mesh_model), "model records must be 68-B MeshModelGPU");
        if (nbTriangles > unsortedTriangles.size()) throw Error(RTQ_EINVAL, "nbTriangles exceeds the triangle vector");
        ctx_->check(rtq_scene_upload(ctx_->get(), reinterpret_cast<const rtq_triangle*>(unsortedTriangles.data()),
                                     nbTriangles,
                                     reinterpret_cast<const rtq_mesh_model*>(meshesInTheScene.data()),
                                     static_cast<uint32_t>(meshesInTheScene.size()), &scene_));
        try {
            ctx_->check(rtq_bvh_build(ctx_->get(), scene_, params));
            fill(nbTriangles);
        } catch (...) {
            rtq_scene_release(ctx_->get(), scene_);
            throw;
        }
    }
    BVH(const BVH&) = delete;
    BVH& operator=(const BVH&) = delete;
    ~BVH() { rtq_scene_release(ctx_->get(), scene_); }

    // Scene::getBVH_NodesToGPUData (scene.cpp:203-208)
    std::vector<rtq_bvh_node> preorder() const {
        uint64_t n = 0;
        ctx_->check(rtq_bvh_export_ref(ctx_->get(), scene_, nullptr, 0, &n));
        std::vector<rtq_bvh_node> out(n);
        ctx_->check(rtq_bvh_export_ref(ctx_->get(), scene_, out.data(), n, &n));
        return out;
    }

    rtq_bvh_info info() const {
        rtq_bvh_info i{};
        ctx_->check(rtq_bvh_info_get(ctx_->get(), scene_, &i));
        return i;
    }

    // Synchronous trace of host rays (closest hit unless flags say otherwise).
    std::vector<rtq_hit> trace(const std::vector<rtq_ray>& rays, uint32_t flags = RTQ_TRACE_CLOSEST) const {
        std::vector<rtq_hit> hits(rays.size());
        ctx_->check(rtq_trace_host(ctx_->get(), scene_, rays.data(), rays.size(), flags, hits.data()));
        return hits;
    }

    // Per-frame edit: new vertices and/or model matrices (same counts; an empty
    // vector keeps that array), then new boxes for the kept tree.  The
    // _InternalStruct boxes are refreshed.
    template <class Tri, class Model>
    void update(const std::vector<Tri>& triangles, const std::vector<Model>& meshes) {
        static_assert(sizeof(Tri) == sizeof(rtq_triangle), "triangle records must be 64-B TriangleGPU");
        static_assert(sizeof(Model) == sizeof(rtq_mesh_model), "model records must be 68-B MeshModelGPU");
        ctx_->check(rtq_scene_update(
            ctx_->get(), scene_, triangles.empty() ? nullptr : reinterpret_cast<const rtq_triangle*>(triangles.data()),
            static_cast<uint32_t>(triangles.size()),
            meshes.empty() ? nullptr : reinterpret_cast<const rtq_mesh_model*>(meshes.data()),
            static_cast<uint32_t>(meshes.size())));
        ctx_->check(rtq_bvh_refit(ctx_->get(), scene_));
        fill(_InternalStruct._NbTriangles);
    }

    // The same edit with the new triangles in device memory (n_triangles 64-B
    // records, e.g. written by an animation kernel): no host copy.
    void update_device(const rtq_triangle* triangles_dev, uint32_t n_triangles) {
        ctx_->check(rtq_scene_update_device(ctx_->get(), scene_, triangles_dev, n_triangles));
        ctx_->check(rtq_bvh_refit(ctx_->get(), scene_));
        fill(_InternalStruct._NbTriangles);
    }

    // getColor for device hit records into device RGBA32F pixels (rtq_shade).
    void shade(const rtq_hit* hits_dev, uint64_t n, const rtq_material* materials_dev, uint32_t n_materials,
               float* rgba_dev, uint32_t flags = 0, const float* overlay_dev = nullptr, void* stream = nullptr) const {
        ctx_->check(rtq_shade(ctx_->get(), scene_, hits_dev, n, materials_dev, n_materials, overlay_dev, flags,
                              rgba_dev, stream));
    }

    uint32_t scene_id() const { return scene_; }
    const std::shared_ptr<Context>& context() const { return ctx_; }

  private:
    void fill(uint32_t n) {
        const uint64_t nc = 2ull * n - 1;
        std::vector<rtq_aabb> boxes(nc);
        std::vector<uint32_t> tri(nc), left(nc), right(nc), parent(nc);
        _InternalStruct._TriangleIndices.assign(n, 0);
        ctx_->check(rtq_bvh_export_clusters(ctx_->get(), scene_, nc, boxes.data(), tri.data(), left.data(),
                                            right.data(), parent.data(), _InternalStruct._TriangleIndices.data(),
                                            nullptr));
        BVHParams& p = _InternalStruct;
        p._NbTriangles = n;
        p._Clusters.assign(nc, std::nullopt);
        p._IsLeaf.assign(nc, std::nullopt);
        p._Parent.assign(nc, std::nullopt);
        p._LeftChild.assign(nc, std::nullopt);
        p._RightChild.assign(nc, std::nullopt);
        for (uint64_t c = 0; c < nc; ++c) {
            rtq_bvh_node node{};
            node.box = boxes[c];
            node.triangle_id = tri[c];   // internal clusters carry 0 and no children, as mergeBVH_Nodes
            p._Clusters[c] = node;
            if (left[c] == 0xFFFFFFFFu) p._IsLeaf[c] = true;
            else { p._LeftChild[c] = left[c]; p._RightChild[c] = right[c]; }
            if (parent[c] != 0xFFFFFFFFu) p._Parent[c] = parent[c];
        }
    }

    std::shared_ptr<Context> ctx_;
    uint32_t scene_ = 0;
};

// A frame's ray sets (any scenes of one context) as one persistent launch
// captured in a hipGraph; launch() replays it.  Item buffers must outlive it.
class FrameBatch {
  public:
    FrameBatch(std::shared_ptr<Context> ctx, const std::vector<rtq_batch_item>& items) : ctx_(std::move(ctx)) {
        rtq_batch* b = nullptr;
        ctx_->check(rtq_batch_create(ctx_->get(), items.data(), static_cast<uint32_t>(items.size()), &b));
        batch_.reset(b);
    }
    void launch(void* stream = nullptr) const { ctx_->check(rtq_batch_launch(batch_.get(), stream)); }

  private:
    struct Del {
        void operator()(rtq_batch* b) const { rtq_batch_destroy(b); }
    };
    std::shared_ptr<Context> ctx_;
    std::unique_ptr<rtq_batch, Del> batch_;
};

// Launch settings for these items from the autotuner (cached per workload).
inline rtq_tune_result autotune(const Context& ctx, const std::vector<rtq_batch_item>& items, uint32_t reps = 2) {
    rtq_tune_result r{};
    ctx.check(rtq_autotune(ctx.get(), items.data(), static_cast<uint32_t>(items.size()), reps, &r));
    return r;
}

// Learned child ordering: logs from a scene's rays, GPU training, order
// words applied to a scene (rtq_order_model_*).  Traces then opt in with
// RTQ_TRACE_LEARNED; results are those of the plain traversal.
class OrderModel {
  public:
    explicit OrderModel(std::shared_ptr<Context> ctx, uint64_t seed = 1, uint64_t sample_capacity = 1u << 22)
        : ctx_(std::move(ctx)) {
        rtq_order_model* m = nullptr;
        ctx_->check(rtq_order_model_create(ctx_->get(), seed, sample_capacity, &m));
        model_.reset(m);
    }
    // samples from every stride-th device ray (flags: RTQ_TRACE_ANY or closest hit)
    uint64_t log(const BVH& bvh, const rtq_ray* rays_dev, uint64_t n_rays, uint32_t flags, uint32_t stride = 1) {
        uint64_t added = 0;
        ctx_->check(rtq_order_model_log(model_.get(), bvh.scene_id(), rays_dev, n_rays, flags, stride, &added));
        return added;
    }
    // mean loss of each epoch
    std::vector<float> train(const rtq_train_params& tp) {
        std::vector<float> loss(tp.epochs);
        ctx_->check(rtq_order_model_train(model_.get(), &tp, loss.data()));
        return loss;
    }
    void apply(const BVH& bvh, float confidence_closest = 1.0f, float confidence_any = 0.5f) {
        ctx_->check(rtq_order_model_apply(model_.get(), bvh.scene_id(), confidence_closest, confidence_any));
    }
    void save(const std::string& path) { ctx_->check(rtq_order_model_save(model_.get(), path.c_str())); }
    void load(const std::string& path) { ctx_->check(rtq_order_model_load(model_.get(), path.c_str())); }
    void clear_samples() { ctx_->check(rtq_order_model_clear_samples(model_.get())); }
    uint64_t samples() const {
        uint64_t n = 0;
        ctx_->check(rtq_order_model_info(model_.get(), &n, nullptr));
        return n;
    }
    uint32_t steps() const {
        uint32_t s = 0;
        ctx_->check(rtq_order_model_info(model_.get(), nullptr, &s));
        return s;
    }

  private:
    struct Del {
        void operator()(rtq_order_model* m) const { rtq_order_model_destroy(m); }
    };
    std::shared_ptr<Context> ctx_;   // declared first: destroyed after the model
    std::unique_ptr<rtq_order_model, Del> model_;
};

}  // namespace rtq
