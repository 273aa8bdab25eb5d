// Internal declarations shared by the rtq translation units (gfx950 only).
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <utility>
#include <vector>

#include "../../include/rtq/rtq.h"
#include "trace_layout.hpp"

namespace rtq {

// Model matrix as uploaded: 16 column-major floats + material id (68 B).
struct DevModel {
    float m[16];
    uint32_t material_id;
};
static_assert(sizeof(DevModel) == 68, "MeshModelGPU is 68 B");

// A device buffer that outlives the build it was allocated for: the next
// build of the scene takes it back when it is large enough (take()), so a
// rebuild does not free and allocate its node arrays again.
struct Kept {
    void* p = nullptr;
    size_t bytes = 0;
};

struct Scene {
    uint32_t n_tris = 0;
    uint32_t n_models = 0;
    rtq_triangle* tris_ref = nullptr;   // N x 64 B, as uploaded
    DevModel* models = nullptr;         // n_models x 68 B
    float4* tri_world = nullptr;        // N x 3 float4: world vertices, original order
    rtq_triangle* tris_stage = nullptr; // N x 64 B: an update's triangles until they pass the checks
    uint32_t* check_words = nullptr;    // 2 words: first triangle failing each device check

    // build products
    bool built = false;
    rtq_build_params params{};
    uint32_t n_clusters = 0;            // 2N - 1
    rtq_aabb* cbox = nullptr;           // cluster boxes
    uint32_t* ctri = nullptr;           // cluster triangle id (leaves)
    uint32_t* cleft = nullptr;
    uint32_t* cright = nullptr;
    uint32_t* cparent = nullptr;
    uint32_t* ccount = nullptr;         // leaves below each cluster
    uint32_t* sorted_idx = nullptr;     // N
    uint32_t* sorted_codes = nullptr;   // N
    rtq_bvh_node* ref_nodes = nullptr;  // 2N - 1 pre-order nodes (reference layout), written on demand
    bool ref_ready = false;             // ref_nodes matches the current tree and boxes
    uint32_t* preorder = nullptr;       // pre-order position of every cluster (2N - 1)
    uint32_t* first_leaf = nullptr;     // rank of the first leaf below every cluster (2N - 1)
    uint32_t* node_depth = nullptr;     // depth of every cluster (2N - 1)
    uint32_t* trav_index = nullptr;     // binary traversal node of every cluster (2N - 1), kNone if none
    float4* nodes = nullptr;            // traversal nodes, 4 float4 (64 B) each
    uint4* node_aux = nullptr;          // per traversal node: child leaf ranges [A,B) [B,C), depth
    bool ordered = false;               // learned child-order words stored in the nodes
    bool order_closest = false;         // ... with decisions for closest-hit rays
    bool order_any = false;             // ... with decisions for any-hit rays
    float4* top = nullptr;              // top-of-tree copy, kTopNodes x 64 B (refs rewritten)
    uint32_t n_top = 0;                 // nodes in it (fewer for small trees)
    float4* tris = nullptr;             // traversal triangles, leaf order, 3 float4 (48 B) each
    uint32_t n_trav = 0;
    float4* wide = nullptr;             // compressed 8-wide nodes, 8 float4 (128 B) each
    uint32_t* wide_trav = nullptr;      // binary traversal node of every 8-wide node's root cluster
    uint32_t* wide_kids = nullptr;      // 8 child clusters of every 8-wide node (kNone past the last)
    bool wide_ordered = false;          // learned first-child tables written in the 8-wide nodes
    uint32_t n_wide = 0;
    uint32_t wide_depth = 0;            // levels of the wide tree
    uint32_t depth = 0;
    uint32_t ploc_passes = 0;
    // (first cluster id, count) of the merges of each PLOC pass, in build order
    std::vector<std::pair<uint32_t, uint32_t>> passes;
    // the same on the device (and its host staging copy); the clusters of the
    // passes from lds_from on fit in one workgroup's LDS together
    // (kLdsClusters): placement, collapse costs and refit merges run those
    // passes in one launch
    uint2* d_passes = nullptr;
    std::vector<uint2> h_passes;
    uint32_t lds_from = 0;
    // the memory behind nodes, node_aux, wide, wide_trav, wide_kids and
    // d_passes, kept from build to build
    Kept kept_nodes, kept_aux, kept_wide, kept_wide_trav, kept_wide_kids, kept_passes;
    // vertices changed since the last build or refit (rtq_scene_update)
    bool stale = false;
    float box_pad = 0.0f;
    float build_ms = 0.0f;
    uint64_t scratch_spill = 0;         // bytes of the last build's temporaries beyond the reserved scratch
    // set by each build from the context's counter, 0 while unbuilt: a batch
    // graph records it and refuses to replay over buffers of another build
    uint64_t generation = 0;
};

}  // namespace rtq

namespace rtq {
// Device scratch of persistent trace launches: counter = the per-part work
// counters and the finished-block counter (all zero between launches: the
// last block of a launch resets them), spill = the deep part of the traversal
// stacks.
struct TraceScratch {
    uint32_t* counter = nullptr;
    uint32_t* spill = nullptr;
    size_t spill_words = 0;
    // fixed: the spill buffer cannot grow (rtq_trace_noalloc); a launch takes
    // fewer workgroups instead, so every spill column fits in spill_words
    bool fixed = false;
};
}  // namespace rtq

namespace rtq {
// Device scratch kept by a context for the temporaries of builds, refits,
// scans and sorts (see Temps).
struct Arena {
    char* base = nullptr;
    size_t cap = 0;
    size_t top = 0;    // bytes handed out
    size_t extra = 0;  // bytes of live allocations made because the arena was full
    size_t peak = 0;   // largest top + extra seen
};
}  // namespace rtq

struct rtq_ctx {
    int device = 0;
    int n_cu = 0;
    hipStream_t stream = nullptr;
    // a second stream (default priority) for build work that runs beside the
    // main stream's, forked and joined with the two events
    hipStream_t side = nullptr;
    hipEvent_t fork = nullptr, join = nullptr;
    // pinned host memory (rtq::kPinnedBytes) for the build's small device-to-
    // host reads (rtq::HostReads)
    void* pinned = nullptr;
    std::string err;
    std::map<uint32_t, rtq::Scene*> scenes;
    uint32_t next_scene = 1;
    uint64_t builds = 0;  // build generations handed out
    rtq::Arena arena;
    rtq::TraceScratch scratch;
    // traversal launch settings: refill a wave once `refill` lanes are idle,
    // taking `chunk` consecutive rays per atomic on a work counter; trace
    // 8-wide nodes when a scene has them and prefer_wide is set.  Defaults
    // from a sweep of single-scene launches on the 2M-triangle terrain
    // (diffuse + shadow): against refill 16 / slack 4, 8-wide -5%, binary -1%
    uint32_t refill = 32;
    uint32_t chunk = 128;
    // leaf tests start once at most spec_slack lanes are still descending
    uint32_t spec_slack = 16;
    // node visits resume once at most leaf_slack lanes still test triangles
    uint32_t leaf_slack = 8;
    bool prefer_wide = true;
    // LDS stack entries per lane of batched launches (0: the layout's default)
    uint32_t stack_lds = 0;
    // autotuner results: key -> launch settings
    std::map<uint64_t, rtq_launch_params> tuned;
};

namespace rtq {
constexpr size_t kPinnedBytes = 128 << 10;

// Device-to-host reads staged through the context's pinned buffer (a copy to
// pageable memory is staged and waited for one by one): add() queues a copy
// on st into the next free bytes, wait() synchronises st once and hands every
// read to its destination.  One HostReads at a time per context.
class HostReads {
public:
    HostReads(rtq_ctx* ctx, hipStream_t st) : ctx_(ctx), st_(st) {}
    hipError_t add(void* dst, const void* src, size_t bytes) {
        const size_t off = (used_ + 15) & ~size_t(15);
        if (n_ == kMaxReads || off + bytes > kPinnedBytes) return hipErrorInvalidValue;
        const hipError_t e =
            hipMemcpyAsync(static_cast<char*>(ctx_->pinned) + off, src, bytes, hipMemcpyDeviceToHost, st_);
        if (e != hipSuccess) return e;
        reads_[n_++] = Read{dst, off, bytes};
        used_ = off + bytes;
        return hipSuccess;
    }
    hipError_t wait() {
        const hipError_t e = hipStreamSynchronize(st_);
        if (e != hipSuccess) return e;
        for (int i = 0; i < n_; ++i)
            std::memcpy(reads_[i].dst, static_cast<const char*>(ctx_->pinned) + reads_[i].off, reads_[i].bytes);
        n_ = 0;
        used_ = 0;
        return hipSuccess;
    }

private:
    struct Read {
        void* dst;
        size_t off, bytes;
    };
    static constexpr int kMaxReads = 8;
    rtq_ctx* ctx_;
    hipStream_t st_;
    Read reads_[kMaxReads] = {};
    int n_ = 0;
    size_t used_ = 0;
};
}  // namespace rtq

struct rtq_batch {
    rtq_ctx* ctx = nullptr;
    std::vector<rtq_batch_item> items;
    std::vector<uint64_t> generations;  // of each item's scene at capture
    rtq::TraceScratch scratch;
    hipStream_t capture = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

namespace rtq {

rtq_status fail(rtq_ctx* ctx, rtq_status s, const char* fmt, ...);
rtq_status hip_fail(rtq_ctx* ctx, hipError_t e, const char* what);
Scene* find_scene(rtq_ctx* ctx, uint32_t id);
hipStream_t pick_stream(rtq_ctx* ctx, void* stream);
void free_build(Scene* s);
void reset_build(Scene* s);

template <typename T>
hipError_t dev_alloc(T** p, size_t count) {
    return hipMalloc(reinterpret_cast<void**>(p), sizeof(T) * (count ? count : 1));
}

// *p = count elements of k's memory, reallocated only when k holds fewer bytes.
template <typename T>
hipError_t take(Kept& k, T** p, size_t count) {
    const size_t bytes = sizeof(T) * (count ? count : 1);
    if (k.bytes < bytes) {
        if (k.p) {
            const hipError_t e = hipFree(k.p);
            k.p = nullptr;
            k.bytes = 0;
            if (e != hipSuccess) return e;
        }
        const hipError_t e = hipMalloc(&k.p, bytes);
        if (e != hipSuccess) {
            k.p = nullptr;
            return e;
        }
        k.bytes = bytes;
    }
    *p = static_cast<T*>(k.p);
    return hipSuccess;
}

// Temporaries of one scope, taken from the top of the context's arena and
// given back when the Temps goes out of scope (scopes nest).  Every user
// queues its work on one stream or waits for it before returning, so memory
// given back is taken again only by kernels queued after the ones that used
// it.  A request past the arena's capacity gets its own allocation (freed,
// with an implicit device synchronisation, by the Temps) and raises the
// arena's peak; arena_reserve grows the arena to that peak between
// operations, so repeating an operation allocates nothing.
class Temps {
  public:
    explicit Temps(rtq_ctx* ctx) : a_(&ctx->arena), mark_(ctx->arena.top) {}
    Temps(const Temps&) = delete;
    Temps& operator=(const Temps&) = delete;
    ~Temps() {
        for (const auto& o : own_) {
            (void)hipFree(o.first);
            a_->extra -= o.second;
        }
        a_->top = mark_;
    }
    template <typename T>
    hipError_t get(T** p, size_t count) {
        const size_t bytes = (sizeof(T) * (count ? count : 1) + 255) & ~size_t(255);
        if (a_->top + bytes <= a_->cap) {
            *p = reinterpret_cast<T*>(a_->base + a_->top);
            a_->top += bytes;
        } else {
            void* q = nullptr;
            const hipError_t e = hipMalloc(&q, bytes);
            if (e != hipSuccess) {
                *p = nullptr;
                return e;
            }
            own_.emplace_back(q, bytes);
            a_->extra += bytes;
            *p = static_cast<T*>(q);
        }
        a_->peak = std::max(a_->peak, a_->top + a_->extra);
        return hipSuccess;
    }

  private:
    Arena* a_;
    size_t mark_;
    std::vector<std::pair<void*, size_t>> own_;
};

// At the start of an operation (no Temps alive): grow the arena to the
// largest total an earlier operation needed.
rtq_status arena_reserve(rtq_ctx* ctx);

// Exclusive scan of u32 on device (wrapping); `total` (device, may be null)
// receives the sum.  Temporaries come from the context's arena; the scan is
// queued on st and not waited for.
rtq_status scan_u32(rtq_ctx* ctx, const uint32_t* in, uint32_t* out, uint64_t n,
                    uint32_t* total, hipStream_t st);
// Exclusive scan of pairs of u32 (component-wise) with a caller workspace of
// scan_u32x2_workspace(n) elements: no allocation, no synchronisation (n > 0).
uint64_t scan_u32x2_workspace(uint64_t n);
rtq_status scan_u32x2_ws(rtq_ctx* ctx, const uint2* in, uint2* out, uint64_t n, uint2* total, uint2* ws,
                         hipStream_t st);
// Trace entry points (trace.hip).  `scratch` null = the context's own.
rtq_status trace(rtq_ctx* ctx, const Scene* s, const rtq_ray* rays, uint64_t n, uint32_t flags, rtq_hit* hits,
                 uint32_t* stat_nodes, uint32_t* stat_tris, bool stats, hipStream_t st);
rtq_status trace_batch(rtq_ctx* ctx, const rtq_batch_item* items, uint32_t n_items, TraceScratch* scratch,
                       bool size_only, hipStream_t st);
// Node layout a batch traces under the context's settings (checks the items):
// 0 binary, 1 8-wide, 2 binary with learned order, 3 8-wide with learned
// first children.
rtq_status batch_layout(rtq_ctx* ctx, const rtq_batch_item* items, uint32_t n_items, int* layout);
rtq_status scratch_alloc(rtq_ctx* ctx, TraceScratch* sc);
// rtq_trace_noalloc (trace.hip): per-device address of the fixed scratch
// array, looked up once per device at context creation (device current).
rtq_status noalloc_setup(rtq_ctx* ctx);
rtq_status trace_noalloc(rtq_ctx* ctx, const Scene* s, const rtq_ray* rays, uint64_t n, uint32_t flags,
                         rtq_hit* hits, hipStream_t st);
void scratch_free(TraceScratch* sc);
// Compressed 8-wide nodes from the cluster arrays (fl: first leaf rank per cluster).
// pos, bidx: pre-order position of every cluster and binary traversal index
// of every position (null for a one-triangle scene, which has no binary node).
// choice: the collapse's cost-table decisions (wide_cost_table), null for a
// one-triangle scene.
rtq_status build_wide(rtq_ctx* ctx, Scene* s, const uint32_t* choice, const uint32_t* fl, const uint32_t* pos,
                      const uint32_t* bidx, hipStream_t st);
// Clusters created by the passes from Scene::lds_from on: at most this many
// (their placement state, 12 B each, and cost rows, 32 B each, in LDS).
constexpr uint32_t kLdsClusters = 4096;
constexpr int kLdsThreads = 1024;   // 4 clusters per thread

// The SAH cost table of the 8-wide collapse (cost: 8 floats, choice: one word
// per internal cluster), bottom-up over the PLOC passes in wide_cost_steps(s)
// launches; wide_cost_table queues launches [first, last) of them.  It reads
// the cluster arrays only, so it can run beside the binary layout's placement.
uint32_t wide_cost_steps(const Scene* s);
rtq_status wide_cost_table(rtq_ctx* ctx, const Scene* s, float* cost, uint32_t* choice, uint32_t first,
                           uint32_t last, hipStream_t st);
// New boxes for the kept 8-wide nodes after a refit (the cluster boxes and
// box pad already refitted): one launch over the nodes, child refs kept.
rtq_status refit_wide(rtq_ctx* ctx, Scene* s, hipStream_t st);
// Learned first-child tables of the 8-wide nodes from the binary nodes' order
// words (learned.hip); clear = zero them.
rtq_status write_wide_order(rtq_ctx* ctx, Scene* s, bool clear, hipStream_t st);
// New boxes for the kept topology after rtq_scene_update (build.hip).
rtq_status refit_scene(rtq_ctx* ctx, Scene* s);
// (Re)writes s->top from s->nodes: after the traversal nodes or their order
// words change (build.hip).
rtq_status write_top(rtq_ctx* ctx, Scene* s, hipStream_t st);
// Writes s->ref_nodes (the reference's pre-order BVH_NodeGPU array) from the
// cluster arrays when the build or refit since the last call changed them;
// queued on st (build.hip).
rtq_status ensure_ref_nodes(rtq_ctx* ctx, Scene* s, hipStream_t st);
// Stable LSD radix sort of (key, value) pairs; `tmp_k`/`tmp_v` are n-entry buffers.
// Queued on st and not waited for; digit counts and look-back words come from
// the context's arena.  *err (device, zeroed by the caller) becomes non-zero
// when a pass's look-back gave up: the order is then wrong and the caller fails.
rtq_status radix_sort_pairs(rtq_ctx* ctx, uint32_t* keys, uint32_t* vals, uint32_t* tmp_k,
                            uint32_t* tmp_v, uint64_t n, uint32_t key_bits, uint32_t* err, hipStream_t st);

}  // namespace rtq

#define RTQ_HIP(ctx, call)                                   \
    do {                                                     \
        hipError_t e_ = (call);                              \
        if (e_ != hipSuccess) return rtq::hip_fail((ctx), e_, #call); \
    } while (0)
