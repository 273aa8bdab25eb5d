// GPU BVH build: scene bounds -> circumscribed cube -> Morton codes -> stable
// radix sort -> PLOC passes -> pre-order flatten -> traversal layout.
//
// Float arithmetic follows the reference's evaluation order exactly (built with
// -ffp-contract=off, std::min/std::max written as the same compares), so codes,
// cluster boxes and the flattened node array are bit-identical to a serial run
// of bvh.cpp / scene.cpp on the same input.
#include <cmath>
#include <cstring>

#include "rtq_internal.hpp"

namespace rtq {
namespace {

constexpr int kThreads = 256;

__device__ __forceinline__ float smin(float a, float b) { return (b < a) ? b : a; }
__device__ __forceinline__ float smax(float a, float b) { return (a < b) ? b : a; }
__host__ __device__ __forceinline__ float canon0(float x) { return x == 0.0f ? 0.0f : x; }

// glm mat4 * vec4 as evaluated at bvh.cpp:238-240 / triangle.cpp:33:
// (m0*x + m1*y) + (m2*z + m3*w), column-major m.
__device__ __forceinline__ void mat_vec(const float* m, float x, float y, float z, float w,
                                        float* out) {
#pragma unroll
    for (int r = 0; r < 3; ++r) {
        float a0 = m[0 + r] * x + m[4 + r] * y;
        float a1 = m[8 + r] * z + m[12 + r] * w;
        out[r] = a0 + a1;
    }
}

__device__ __forceinline__ void world_vertex(const DevModel& md, const rtq_vec4& p, float* out) {
    float m[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) m[i] = md.m[i];
    mat_vec(m, p.x, p.y, p.z, p.w, out);
}

// Triangle box, bvh.cpp:383-399.
__device__ __forceinline__ void tri_box(const rtq_triangle& t, const DevModel& md, float* lo, float* hi) {
    float p0[3], p1[3], p2[3];
    world_vertex(md, t.p0, p0);
    world_vertex(md, t.p1, p1);
    world_vertex(md, t.p2, p2);
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        lo[k] = smin(p0[k], smin(p1[k], p2[k]));
        hi[k] = smax(p0[k], smax(p1[k], p2[k]));
    }
}

// ------------------------------------------------------------- refit

// Leaf cluster c < N keeps its triangle; its box is the triangle's box for the
// current vertices (bvh.cpp:383-399), as k_leaf_init computes it.
__global__ __launch_bounds__(kThreads) void k_refit_leaves(const rtq_triangle* __restrict__ tris,
                                                           const DevModel* __restrict__ models,
                                                           const uint32_t* __restrict__ ctri, uint32_t n,
                                                           rtq_aabb* __restrict__ cbox) {
    const uint32_t i = blockIdx.x * kThreads + threadIdx.x;
    if (i >= n) return;
    const rtq_triangle t = tris[ctri[i]];
    float lo[3], hi[3];
    tri_box(t, models[t.model_id], lo, hi);
    rtq_aabb a{};
    for (int k = 0; k < 3; ++k) { a.min[k] = lo[k]; a.max[k] = hi[k]; }
    cbox[i] = a;
}

// Merged clusters [first, first + count) of one PLOC pass: the box is the
// merge of the children's (left operand first, as k_ploc_merge).
__global__ __launch_bounds__(kThreads) void k_refit_merge(uint32_t first, uint32_t count,
                                                          const uint32_t* __restrict__ cleft,
                                                          const uint32_t* __restrict__ cright,
                                                          rtq_aabb* __restrict__ cbox) {
    const uint32_t i = blockIdx.x * kThreads + threadIdx.x;
    if (i >= count) return;
    const uint32_t id = first + i;
    const rtq_aabb l = cbox[cleft[id]], r = cbox[cright[id]];
    rtq_aabb a{};
    for (int k = 0; k < 3; ++k) {
        a.min[k] = smin(l.min[k], r.min[k]);
        a.max[k] = smax(l.max[k], r.max[k]);
    }
    cbox[id] = a;
}

// ------------------------------------------------------------- top of the tree

// One thread: BFS from the root over internal-node refs, at most kTopNodes
// nodes.  A child's slot is its position in the BFS queue, known when it is
// appended, so each node is read once and written with its refs into the set
// already rewritten to kTopBit | slot (leaf refs and refs leaving the set
// unchanged).
__global__ void k_top_nodes(const float4* __restrict__ nodes, float4* __restrict__ top,
                            uint32_t* __restrict__ n_top) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    uint32_t q[kTopNodes > 0 ? kTopNodes : 1];
    uint32_t n = 0;
    q[n++] = 0;
    for (uint32_t i = 0; i < n; ++i) {
        const float4* src = nodes + 4ull * q[i];
        const float4 a = src[0], b = src[1], c = src[2];
        float4 w = src[3];
        uint32_t r[2] = {__float_as_uint(w.x), __float_as_uint(w.y)};
        for (int k = 0; k < 2; ++k)
            if (!(r[k] & kLeafBit) && n < kTopNodes) {
                q[n] = r[k];
                r[k] = kTopBit | n;
                ++n;
            }
        w.x = __uint_as_float(r[0]);
        w.y = __uint_as_float(r[1]);
        float4* dst = top + 4ull * i;
        dst[0] = a;
        dst[1] = b;
        dst[2] = c;
        dst[3] = w;
    }
    *n_top = n;
}

// ------------------------------------------------------------- scene bounds

__global__ __launch_bounds__(kThreads) void k_bounds_partial(const rtq_triangle* __restrict__ tris,
                                                             const DevModel* __restrict__ models,
                                                             uint32_t n, float* __restrict__ partial) {
    __shared__ float s[6][kThreads];
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (uint32_t i = blockIdx.x * kThreads + threadIdx.x; i < n; i += gridDim.x * kThreads) {
        float a[3], b[3];
        tri_box(tris[i], models[tris[i].model_id], a, b);
#pragma unroll
        for (int k = 0; k < 3; ++k) { lo[k] = smin(lo[k], a[k]); hi[k] = smax(hi[k], b[k]); }
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) { s[k][threadIdx.x] = lo[k]; s[3 + k][threadIdx.x] = hi[k]; }
    __syncthreads();
    for (int w = kThreads / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                s[k][threadIdx.x] = smin(s[k][threadIdx.x], s[k][threadIdx.x + w]);
                s[3 + k][threadIdx.x] = smax(s[3 + k][threadIdx.x], s[3 + k][threadIdx.x + w]);
            }
        }
        __syncthreads();
    }
    if (threadIdx.x < 6) partial[blockIdx.x * 6 + threadIdx.x] = s[threadIdx.x][0];
}

// Reduces the partial boxes, then (one thread) the circumscribed cube of
// bvh.cpp:253-303.  out[0] = scene box, out[1] = cube.
__global__ __launch_bounds__(kThreads) void k_bounds_final(const float* __restrict__ partial, uint32_t nb,
                                                           uint32_t n, uint32_t pad_origin,
                                                           rtq_aabb* __restrict__ out) {
    __shared__ float s[6][kThreads];
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (uint32_t b = threadIdx.x; b < nb; b += kThreads) {
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            lo[k] = smin(lo[k], partial[b * 6 + k]);
            hi[k] = smax(hi[k], partial[b * 6 + 3 + k]);
        }
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) { s[k][threadIdx.x] = lo[k]; s[3 + k][threadIdx.x] = hi[k]; }
    __syncthreads();
    for (int w = kThreads / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                s[k][threadIdx.x] = smin(s[k][threadIdx.x], s[k][threadIdx.x + w]);
                s[3 + k][threadIdx.x] = smax(s[3 + k][threadIdx.x], s[3 + k][threadIdx.x + w]);
            }
        }
        __syncthreads();
    }
    if (threadIdx.x != 0) return;
    rtq_aabb b{};
    for (int k = 0; k < 3; ++k) { b.min[k] = s[k][0]; b.max[k] = s[3 + k][0]; }
    if (pad_origin && n < 65536u) {
        for (int k = 0; k < 3; ++k) { b.max[k] = smax(b.max[k], 0.0f); b.min[k] = smin(b.min[k], 0.0f); }
    }
    for (int k = 0; k < 3; ++k) { b.min[k] = canon0(b.min[k]); b.max[k] = canon0(b.max[k]); }
    out[0] = b;
    rtq_aabb c = b;
    float dist_x = b.max[0] - b.min[0];
    float dist_y = b.max[1] - b.min[1];
    float dist_z = b.max[2] - b.min[2];
    float max_dist = 0.0f;
    int axis;
    if (dist_x > max_dist) { max_dist = dist_x; axis = 0; }
    else if (dist_y > max_dist) { max_dist = dist_y; axis = 1; }
    else { max_dist = dist_z; axis = 2; }
    float delta;
    if (axis == 0) {
        delta = (max_dist - dist_y) / 2.0f; c.max[1] += delta; c.min[1] -= delta;
        delta = (max_dist - dist_z) / 2.0f; c.max[2] += delta; c.min[2] -= delta;
    } else if (axis == 1) {
        delta = (max_dist - dist_x) / 2.0f; c.max[0] += delta; c.min[0] -= delta;
        delta = (max_dist - dist_z) / 2.0f; c.max[2] += delta; c.min[2] -= delta;
    } else {
        delta = (max_dist - dist_x) / 2.0f; c.max[0] += delta; c.min[0] -= delta;
        delta = (max_dist - dist_y) / 2.0f; c.max[1] += delta; c.min[1] -= delta;
    }
    out[1] = c;
}

// ------------------------------------------------------------- Morton codes

__device__ __forceinline__ uint32_t expand_bits(uint32_t v) {
    v = (v * 0x00010001u) & 0xFF0000FFu;
    v = (v * 0x00000101u) & 0x0F00F00Fu;
    v = (v * 0x00000011u) & 0xC30C30C3u;
    v = (v * 0x00000005u) & 0x49249249u;
    return v;
}

__device__ __forceinline__ uint32_t quantize10(float x) {
    x = smin(smax(x * 1024.0f, 0.0f), 1023.0f);
    if (x != x) return 0u;   // zero-extent axis: the reference cast is undefined
    return (uint32_t)x;
}

// bvh.cpp:305-372 for one triangle.
__global__ __launch_bounds__(kThreads) void k_morton(const rtq_triangle* __restrict__ tris,
                                                     const DevModel* __restrict__ models, uint32_t n,
                                                     const rtq_aabb* __restrict__ bounds,
                                                     uint32_t* __restrict__ codes,
                                                     uint32_t* __restrict__ idx) {
    const uint32_t i = blockIdx.x * kThreads + threadIdx.x;
    if (i >= n) return;
    const rtq_aabb cube = bounds[1];
    const rtq_triangle t = tris[i];
    const DevModel& md = models[t.model_id];
    const float third = 1.0f / 3.0f;
    float ms[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) ms[k] = third * md.m[k];
    float sx = (t.p0.x + t.p1.x) + t.p2.x;
    float sy = (t.p0.y + t.p1.y) + t.p2.y;
    float sz = (t.p0.z + t.p1.z) + t.p2.z;
    float sw = (t.p0.w + t.p1.w) + t.p2.w;
    float c[3];
    mat_vec(ms, sx, sy, sz, sw, c);
    uint32_t q[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        float len = cube.max[k] - cube.min[k];
        q[k] = quantize10((c[k] - cube.min[k]) / len);
    }
    codes[i] = (expand_bits(q[0]) << 2) | (expand_bits(q[1]) << 1) | expand_bits(q[2]);
    if (idx) idx[i] = i;
}

// --------------------------------------------------------------------- PLOC

struct Box6 {
    float lo[3], hi[3];
};

__device__ __forceinline__ float merged_sa(const Box6& a, const Box6& b) {
    // AABB::merge then AABB::getSurfaceArea (bvh.cpp:378-381, 401-413)
    float dx = smax(a.hi[0], b.hi[0]) - smin(a.lo[0], b.lo[0]);
    float dy = smax(a.hi[1], b.hi[1]) - smin(a.lo[1], b.lo[1]);
    float dz = smax(a.hi[2], b.hi[2]) - smin(a.lo[2], b.lo[2]);
    return 2.0f * ((dx * dy + dy * dz) + dz * dx);
}

__global__ __launch_bounds__(kThreads) void k_cluster_init(uint32_t n, uint32_t* __restrict__ cparent,
                                                           uint32_t* __restrict__ cleft,
                                                           uint32_t* __restrict__ cright,
                                                           uint32_t* __restrict__ ctri) {
    const uint32_t c = blockIdx.x * kThreads + threadIdx.x;
    if (c >= 2 * n - 1) return;
    cparent[c] = kNone;
    cleft[c] = kNone;
    cright[c] = kNone;
    ctri[c] = 0;
}

// plocPreprocessing, bvh.cpp:26-46.
__global__ __launch_bounds__(kThreads) void k_leaf_init(const rtq_triangle* __restrict__ tris,
                                                        const DevModel* __restrict__ models,
                                                        const uint32_t* __restrict__ sorted_idx,
                                                        uint32_t n, rtq_aabb* __restrict__ cbox,
                                                        uint32_t* __restrict__ ctri,
                                                        uint32_t* __restrict__ ccount,
                                                        uint32_t* __restrict__ c_in,
                                                        Box6* __restrict__ b_in) {
    const uint32_t i = blockIdx.x * kThreads + threadIdx.x;
    if (i >= n) return;
    const uint32_t ti = sorted_idx[i];
    const rtq_triangle t = tris[ti];
    Box6 b;
    tri_box(t, models[t.model_id], b.lo, b.hi);
    rtq_aabb a{};
    for (int k = 0; k < 3; ++k) { a.min[k] = b.lo[k]; a.max[k] = b.hi[k]; }
    cbox[i] = a;
    ctri[i] = ti;
    ccount[i] = 1;
    c_in[i] = i;
    b_in[i] = b;
}

// plocNearestNeighborSearch, bvh.cpp:193-210: window [i-r, i+r] of the active
// cluster list staged through LDS; strict '<' keeps the lowest j on ties.
// PLOC state on the device, so passes can be queued without reading the
// cluster count back: state[0] = clusters this pass, state[1] = clusters
// made so far (the next merge id).  Kernels are launched for an upper bound.
__global__ __launch_bounds__(kThreads) void k_ploc_nn(const Box6* __restrict__ b_in,
                                                      const uint32_t* __restrict__ state, uint32_t radius,
                                                      uint32_t* __restrict__ nn) {
    __shared__ Box6 s_box[kThreads + 64];
    const uint32_t iter = state[0];
    const uint32_t base = blockIdx.x * kThreads;
    const uint32_t lo = base > radius ? base - radius : 0u;
    const uint32_t hi = min(iter, base + kThreads + radius);
    for (uint32_t j = lo + threadIdx.x; j < hi; j += kThreads) s_box[j - lo] = b_in[j];
    __syncthreads();
    const uint32_t i = base + threadIdx.x;
    if (i >= iter) return;
    const Box6 bi = s_box[i - lo];
    const uint32_t start = i > radius ? i - radius : 0u;
    const uint32_t end = min(i + radius + 1, iter);
    float best = INFINITY;
    uint32_t arg = kNone;
    for (uint32_t j = start; j < end; ++j) {
        if (j == i) continue;
        float d = merged_sa(bi, s_box[j - lo]);
        if (d < best) { best = d; arg = j; }
    }
    if (arg != kNone) nn[i] = arg;
}

// Per position: x = 1 if this position starts a merge (mutual pair, lower
// index), y = 1 if a cluster stays at this position after merging.
__global__ __launch_bounds__(kThreads) void k_ploc_flags(const uint32_t* __restrict__ nn,
                                                         const uint32_t* __restrict__ state, uint32_t bound,
                                                         uint2* __restrict__ flags) {
    const uint32_t i = blockIdx.x * kThreads + threadIdx.x;
    const uint32_t iter = state[0];
    if (i >= bound) return;
    if (i >= iter) {
        flags[i] = make_uint2(0u, 0u);  // past this pass's clusters: counted as nothing
        return;
    }
    const uint32_t j = nn[i];
    const bool mutual = (j < iter) && (j != i) && nn[j] == i;
    flags[i] = make_uint2((mutual && i < j) ? 1u : 0u, (mutual && j < i) ? 0u : 1u);
}

// plocMerging + plocCompaction (bvh.cpp:150-191): merged cluster ids are
// n_total + (number of merges at lower positions), deterministic.
__global__ __launch_bounds__(kThreads) void k_ploc_merge(
    const uint32_t* __restrict__ state, const uint32_t* __restrict__ nn, const uint2* __restrict__ flags,
    const uint2* __restrict__ offs, const uint32_t* __restrict__ c_in, const Box6* __restrict__ b_in,
    uint32_t* __restrict__ c_out, Box6* __restrict__ b_out, rtq_aabb* __restrict__ cbox,
    uint32_t* __restrict__ cleft, uint32_t* __restrict__ cright, uint32_t* __restrict__ cparent,
    uint32_t* __restrict__ ccount, uint32_t* __restrict__ ctri) {
    const uint32_t i = blockIdx.x * kThreads + threadIdx.x;
    const uint32_t iter = state[0], n_total = state[1];
    if (i >= iter) return;
    const uint2 f = flags[i];
    if (!f.y) return;
    const uint2 o = offs[i];
    uint32_t cid = c_in[i];
    Box6 b = b_in[i];
    if (f.x) {
        const uint32_t j = nn[i];
        const uint32_t cj = c_in[j];
        const Box6 bj = b_in[j];
        const uint32_t id = n_total + o.x;
        Box6 m;
        rtq_aabb a{};
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            m.lo[k] = smin(b.lo[k], bj.lo[k]);
            m.hi[k] = smax(b.hi[k], bj.hi[k]);
            a.min[k] = m.lo[k];
            a.max[k] = m.hi[k];
        }
        cbox[id] = a;
        cleft[id] = cid;
        cright[id] = cj;
        cparent[cid] = id;
        cparent[cj] = id;
        ccount[id] = ccount[cid] + ccount[cj];
        ctri[id] = 0;
        cid = id;
        b = m;
    }
    c_out[o.y] = cid;
    b_out[o.y] = b;
}

// --------------------------------------------------------------- flattening

// End of a pass: log (first merge id, merges, clusters) and move the state on.
__global__ void k_ploc_advance(const uint2* __restrict__ tot, uint32_t* __restrict__ state,
                               uint4* __restrict__ pass_log, uint32_t slot) {
    const uint2 t = *tot;
    pass_log[slot] = make_uint4(state[1], t.x, state[0], 0u);
    state[1] += t.x;
    state[0] = t.y;
}

__global__ void k_root_init(uint32_t root, uint32_t* pos, uint32_t* fl, uint32_t* dep) {
    pos[root] = 0;
    fl[root] = 0;
    dep[root] = 0;
}

// Top-down pre-order placement for the clusters created in one PLOC pass
// (their parents were created in later passes, so are already placed):
// left child right after the parent, right child after the left subtree
// (scene.cpp:189-201); fl = rank of the subtree's first leaf.
__global__ __launch_bounds__(kThreads) void k_place_pass(uint32_t first, uint32_t count,
                                                         const uint32_t* __restrict__ cleft,
                                                         const uint32_t* __restrict__ cright,
                                                         const uint32_t* __restrict__ ccount,
                                                         uint32_t* __restrict__ pos,
                                                         uint32_t* __restrict__ fl,
                                                         uint32_t* __restrict__ dep) {
    const uint32_t k = blockIdx.x * kThreads + threadIdx.x;
    if (k >= count) return;
    const uint32_t p = first + k;
    const uint32_t l = cleft[p], r = cright[p];
    const uint32_t pp = pos[p], pf = fl[p], pd = dep[p];
    const uint32_t lc = ccount[l];
    pos[l] = pp + 1;
    pos[r] = pp + 2 * lc;
    fl[l] = pf;
    fl[r] = pf + lc;
    dep[l] = pd + 1;
    dep[r] = pd + 1;
}

// Reference node array (scene.cpp:189-208) + leaf-ordered traversal triangles.
__global__ __launch_bounds__(kThreads) void k_write_ref(
    uint32_t n_clusters, const rtq_aabb* __restrict__ cbox, const uint32_t* __restrict__ ctri,
    const uint32_t* __restrict__ cleft, const uint32_t* __restrict__ cright,
    const uint32_t* __restrict__ pos, const uint32_t* __restrict__ fl, const uint32_t* __restrict__ dep,
    const float4* __restrict__ tri_world, rtq_bvh_node* __restrict__ ref, float4* __restrict__ tris,
    uint32_t* __restrict__ max_depth) {
    const uint32_t c = blockIdx.x * kThreads + threadIdx.x;
    if (c >= n_clusters) return;
    rtq_bvh_node nd{};
    nd.box = cbox[c];
    nd.box.pad0_ = 0.0f;
    nd.box.pad1_ = 0.0f;
    const bool leaf = cleft[c] == kNone;
    nd.triangle_id = leaf ? ctri[c] : 0u;
    nd.left_child = leaf ? 0u : pos[cleft[c]];
    nd.right_child = leaf ? 0u : pos[cright[c]];
    nd.pad_ = 0u;
    ref[pos[c]] = nd;
    if (leaf) {
        const uint32_t t = ctri[c];
        const uint64_t dst = 3ull * fl[c];
        tris[dst + 0] = tri_world[3ull * t + 0];
        tris[dst + 1] = tri_world[3ull * t + 1];
        tris[dst + 2] = tri_world[3ull * t + 2];
        atomicMax(max_depth, dep[c]);
    }
}

__device__ __forceinline__ bool is_trav_node(uint32_t c, uint32_t root, uint32_t max_leaf,
                                             const uint32_t* cleft, const uint32_t* ccount) {
    return cleft[c] != kNone && (ccount[c] > max_leaf || c == root);
}

__global__ __launch_bounds__(kThreads) void k_mark_trav(uint32_t n_clusters, uint32_t root, uint32_t max_leaf,
                                                        const uint32_t* __restrict__ cleft,
                                                        const uint32_t* __restrict__ ccount,
                                                        const uint32_t* __restrict__ pos,
                                                        uint32_t* __restrict__ flag) {
    const uint32_t c = blockIdx.x * kThreads + threadIdx.x;
    if (c >= n_clusters) return;
    flag[pos[c]] = is_trav_node(c, root, max_leaf, cleft, ccount) ? 1u : 0u;
}

// Traversal node (64 B): child boxes widened outward by `pad`, interleaved as
//   [c0.lo.x c0.hi.x c0.lo.y c0.hi.y] [c0.lo.z c0.hi.z c1.lo.x c1.hi.x]
//   [c1.lo.y c1.hi.y c1.lo.z c1.hi.z] [ref0 ref1 order 0]
// where `order` holds the learned child-order bits (zero until a model is applied).
__device__ __forceinline__ void child_desc(uint32_t ch, uint32_t root, uint32_t max_leaf,
                                           const rtq_aabb* cbox, const uint32_t* cleft,
                                           const uint32_t* ccount, const uint32_t* pos,
                                           const uint32_t* fl, const uint32_t* bidx, float pad,
                                           float* lo, float* hi, uint32_t* ref) {
    const rtq_aabb a = cbox[ch];
    for (int k = 0; k < 3; ++k) { lo[k] = a.min[k] - pad; hi[k] = a.max[k] + pad; }
    if (is_trav_node(ch, root, max_leaf, cleft, ccount)) *ref = bidx[pos[ch]];
    else *ref = leaf_ref(fl[ch], ccount[ch]);
}

__global__ __launch_bounds__(kThreads) void k_write_trav(
    uint32_t n_clusters, uint32_t root, uint32_t max_leaf, float pad, const rtq_aabb* __restrict__ cbox,
    const uint32_t* __restrict__ cleft, const uint32_t* __restrict__ cright,
    const uint32_t* __restrict__ ccount, const uint32_t* __restrict__ pos, const uint32_t* __restrict__ fl,
    const uint32_t* __restrict__ bidx, const uint32_t* __restrict__ dep, float4* __restrict__ nodes,
    uint4* __restrict__ aux) {
    const uint32_t c = blockIdx.x * kThreads + threadIdx.x;
    if (c >= n_clusters) return;
    if (!is_trav_node(c, root, max_leaf, cleft, ccount)) return;
    float l0[3], h0[3], l1[3], h1[3];
    uint32_t r0, r1;
    child_desc(cleft[c], root, max_leaf, cbox, cleft, ccount, pos, fl, bidx, pad, l0, h0, &r0);
    child_desc(cright[c], root, max_leaf, cbox, cleft, ccount, pos, fl, bidx, pad, l1, h1, &r1);
    const uint32_t idx = bidx[pos[c]];
    float4* o = nodes + 4ull * idx;
    o[0] = make_float4(l0[0], h0[0], l0[1], h0[1]);
    o[1] = make_float4(l0[2], h0[2], l1[0], h1[0]);
    o[2] = make_float4(l1[1], h1[1], l1[2], h1[2]);
    // .z holds the learned child-order bits (none until an order model is applied)
    o[3] = make_float4(__uint_as_float(r0), __uint_as_float(r1), 0.0f, 0.0f);
    // leaf-order ranges of the children, [A, B) and [B, C), and the binary depth
    const uint32_t a0 = fl[cleft[c]], b0 = fl[cright[c]];
    aux[idx] = make_uint4(a0, b0, b0 + ccount[cright[c]], dep[c]);
}

// Single-triangle scene: one node whose two children are the same leaf.
__global__ void k_write_single(const rtq_aabb* __restrict__ cbox, float pad, float4* __restrict__ nodes,
                               uint4* __restrict__ aux) {
    const rtq_aabb a = cbox[0];
    float lo[3], hi[3];
    for (int k = 0; k < 3; ++k) { lo[k] = a.min[k] - pad; hi[k] = a.max[k] + pad; }
    const uint32_t r = leaf_ref(0, 1);
    nodes[0] = make_float4(lo[0], hi[0], lo[1], hi[1]);
    nodes[1] = make_float4(lo[2], hi[2], lo[0], hi[0]);
    nodes[2] = make_float4(lo[1], hi[1], lo[2], hi[2]);
    nodes[3] = make_float4(__uint_as_float(r), __uint_as_float(r), 0.0f, 0.0f);
    aux[0] = make_uint4(0u, 1u, 1u, 0u);
}

inline uint32_t blocks_for(uint64_t n) { return (uint32_t)((n + kThreads - 1) / kThreads); }

// Shared front half of the build and of rtq_morton_codes: scene box + cube.
// d_bounds holds 2 boxes followed by 6 floats per partial block (at most
// kBoundsPartials): kBoundsBoxes boxes of 32 B.
constexpr uint32_t kBoundsPartials = 1024;
constexpr uint32_t kBoundsBoxes = 2 + (6 * kBoundsPartials + 7) / 8;
rtq_status compute_bounds(rtq_ctx* ctx, Scene* s, uint32_t pad_origin, rtq_aabb* d_bounds,
                          hipStream_t st) {
    const uint32_t nb = std::min<uint32_t>(blocks_for(s->n_tris), kBoundsPartials);
    float* partial = reinterpret_cast<float*>(d_bounds + 2);
    hipLaunchKernelGGL(k_bounds_partial, dim3(nb), dim3(kThreads), 0, st, s->tris_ref, s->models,
                       s->n_tris, partial);
    RTQ_HIP(ctx, hipGetLastError());
    hipLaunchKernelGGL(k_bounds_final, dim3(1), dim3(kThreads), 0, st, partial, nb, s->n_tris,
                       pad_origin, d_bounds);
    RTQ_HIP(ctx, hipGetLastError());
    return RTQ_OK;
}

struct BuildTemps {
    std::vector<void*> ptrs;
    ~BuildTemps() {
        for (void* p : ptrs) (void)hipFree(p);
    }
    template <typename T>
    hipError_t get(T** p, size_t count) {
        hipError_t e = dev_alloc(p, count);
        if (e == hipSuccess) ptrs.push_back(*p);
        return e;
    }
};

// Steps after PLOC, shared by the build and the refit: pre-order placement,
// the reference node array and leaf-ordered triangles, the box pad from the
// scene extent, the binary traversal nodes and (node_width 8) the compressed
// wide nodes.  Buffers that exist already are written in place (a refit keeps
// their addresses, so captured batch graphs stay valid).
rtq_status write_layouts(rtq_ctx* ctx, Scene* s, const rtq_aabb* d_bounds, hipStream_t st) {
    const uint32_t n = s->n_tris;
    const uint32_t nc = s->n_clusters;
    BuildTemps tmp;
    uint32_t *pos = nullptr, *fl = nullptr, *dep = nullptr, *maxd = nullptr, *flag = nullptr,
             *bidx = nullptr, *ntrav = nullptr;
    RTQ_HIP(ctx, tmp.get(&pos, nc));
    RTQ_HIP(ctx, tmp.get(&fl, nc));
    RTQ_HIP(ctx, tmp.get(&dep, nc));
    RTQ_HIP(ctx, tmp.get(&maxd, 2));
    RTQ_HIP(ctx, tmp.get(&flag, nc));
    RTQ_HIP(ctx, tmp.get(&bidx, nc));
    RTQ_HIP(ctx, tmp.get(&ntrav, 1));
    RTQ_HIP(ctx, hipMemsetAsync(maxd, 0, sizeof(uint32_t) * 2, st));
    const uint32_t root = nc - 1;
    hipLaunchKernelGGL(k_root_init, dim3(1), dim3(1), 0, st, root, pos, fl, dep);
    RTQ_HIP(ctx, hipGetLastError());
    for (auto it = s->passes.rbegin(); it != s->passes.rend(); ++it) {
        hipLaunchKernelGGL(k_place_pass, dim3(blocks_for(it->second)), dim3(kThreads), 0, st, it->first,
                           it->second, s->cleft, s->cright, s->ccount, pos, fl, dep);
        RTQ_HIP(ctx, hipGetLastError());
    }
    hipLaunchKernelGGL(k_write_ref, dim3(blocks_for(nc)), dim3(kThreads), 0, st, nc, s->cbox, s->ctri,
                       s->cleft, s->cright, pos, fl, dep, s->tri_world, s->ref_nodes, s->tris, maxd);
    RTQ_HIP(ctx, hipGetLastError());

    rtq_aabb h_bounds[2];
    uint32_t h_depth = 0;
    RTQ_HIP(ctx, hipMemcpyAsync(h_bounds, d_bounds, sizeof h_bounds, hipMemcpyDeviceToHost, st));
    RTQ_HIP(ctx, hipMemcpyAsync(&h_depth, maxd, sizeof h_depth, hipMemcpyDeviceToHost, st));
    RTQ_HIP(ctx, hipStreamSynchronize(st));
    float extent = 0.0f;
    for (int k = 0; k < 3; ++k)
        extent = std::max(extent, std::max(std::fabs(h_bounds[0].min[k]), std::fabs(h_bounds[0].max[k])));
    s->box_pad = std::max(std::ldexp(extent, -18), 1e-30f);
    s->depth = h_depth;

    const uint32_t max_leaf = s->params.max_leaf_size;
    if (n == 1) {
        s->n_trav = 1;
        if (!s->nodes) RTQ_HIP(ctx, dev_alloc(&s->nodes, 4));
        if (!s->node_aux) RTQ_HIP(ctx, dev_alloc(&s->node_aux, 1));
        hipLaunchKernelGGL(k_write_single, dim3(1), dim3(1), 0, st, s->cbox, s->box_pad, s->nodes, s->node_aux);
        RTQ_HIP(ctx, hipGetLastError());
    } else {
        hipLaunchKernelGGL(k_mark_trav, dim3(blocks_for(nc)), dim3(kThreads), 0, st, nc, root, max_leaf,
                           s->cleft, s->ccount, pos, flag);
        RTQ_HIP(ctx, hipGetLastError());
        rtq_status rs = scan_u32(ctx, flag, bidx, nc, ntrav, st);
        if (rs != RTQ_OK) return rs;
        uint32_t h_ntrav = 0;
        RTQ_HIP(ctx, hipMemcpyAsync(&h_ntrav, ntrav, sizeof h_ntrav, hipMemcpyDeviceToHost, st));
        RTQ_HIP(ctx, hipStreamSynchronize(st));
        if (s->nodes && h_ntrav != s->n_trav)
            return fail(ctx, RTQ_EINTERNAL, "refit: %u traversal nodes, the tree has %u", h_ntrav, s->n_trav);
        s->n_trav = h_ntrav;
        if (!s->nodes) RTQ_HIP(ctx, dev_alloc(&s->nodes, 4ull * h_ntrav));
        if (!s->node_aux) RTQ_HIP(ctx, dev_alloc(&s->node_aux, h_ntrav));
        hipLaunchKernelGGL(k_write_trav, dim3(blocks_for(nc)), dim3(kThreads), 0, st, nc, root, max_leaf,
                           s->box_pad, s->cbox, s->cleft, s->cright, s->ccount, pos, fl, bidx, dep, s->nodes,
                           s->node_aux);
        RTQ_HIP(ctx, hipGetLastError());
    }
    if (s->params.node_width == 8) {
        rtq_status rs = build_wide(ctx, s, fl, st);
        if (rs != RTQ_OK) return rs;
    }
    rtq_status rs = write_top(ctx, s, st);
    if (rs != RTQ_OK) return rs;
    RTQ_HIP(ctx, hipStreamSynchronize(st));
    return RTQ_OK;
}

}  // namespace

rtq_status build_scene(rtq_ctx* ctx, Scene* s, const rtq_build_params& prm) {
    hipStream_t st = ctx->stream;
    const uint32_t n = s->n_tris;
    const uint32_t nc = 2 * n - 1;
    free_build(s);
    s->params = prm;
    s->n_clusters = nc;

    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    RTQ_HIP(ctx, hipEventCreate(&ev0));
    RTQ_HIP(ctx, hipEventCreate(&ev1));
    struct EvGuard {
        hipEvent_t a, b;
        ~EvGuard() { (void)hipEventDestroy(a); (void)hipEventDestroy(b); }
    } evg{ev0, ev1};
    RTQ_HIP(ctx, hipEventRecord(ev0, st));

    // persistent products
    RTQ_HIP(ctx, dev_alloc(&s->cbox, nc));
    RTQ_HIP(ctx, dev_alloc(&s->ctri, nc));
    RTQ_HIP(ctx, dev_alloc(&s->cleft, nc));
    RTQ_HIP(ctx, dev_alloc(&s->cright, nc));
    RTQ_HIP(ctx, dev_alloc(&s->cparent, nc));
    RTQ_HIP(ctx, dev_alloc(&s->ccount, nc));
    RTQ_HIP(ctx, dev_alloc(&s->sorted_idx, n));
    RTQ_HIP(ctx, dev_alloc(&s->sorted_codes, n));
    RTQ_HIP(ctx, dev_alloc(&s->ref_nodes, nc));
    RTQ_HIP(ctx, dev_alloc(&s->tris, 3ull * n));

    BuildTemps tmp;
    rtq_aabb* d_bounds = nullptr;
    uint32_t *tk = nullptr, *tv = nullptr, *c_in = nullptr, *c_out = nullptr, *nn = nullptr;
    Box6 *b_in = nullptr, *b_out = nullptr;
    uint2 *flags = nullptr, *offs = nullptr, *tot = nullptr;
    RTQ_HIP(ctx, tmp.get(&d_bounds, kBoundsBoxes));
    RTQ_HIP(ctx, tmp.get(&tk, n));
    RTQ_HIP(ctx, tmp.get(&tv, n));
    RTQ_HIP(ctx, tmp.get(&c_in, n));
    RTQ_HIP(ctx, tmp.get(&c_out, n));
    RTQ_HIP(ctx, tmp.get(&nn, n));
    RTQ_HIP(ctx, tmp.get(&b_in, n));
    RTQ_HIP(ctx, tmp.get(&b_out, n));
    RTQ_HIP(ctx, tmp.get(&flags, n));
    RTQ_HIP(ctx, tmp.get(&offs, n));
    RTQ_HIP(ctx, tmp.get(&tot, 1));
    RTQ_HIP(ctx, hipMemsetAsync(nn, 0, sizeof(uint32_t) * n, st));

    // 1. bounds, cube, Morton codes, sort
    rtq_status rs = compute_bounds(ctx, s, prm.flags & RTQ_BUILD_REF_PADDED_BOUNDS, d_bounds, st);
    if (rs != RTQ_OK) return rs;
    hipLaunchKernelGGL(k_morton, dim3(blocks_for(n)), dim3(kThreads), 0, st, s->tris_ref, s->models, n,
                       d_bounds, s->sorted_codes, s->sorted_idx);
    RTQ_HIP(ctx, hipGetLastError());
    rs = radix_sort_pairs(ctx, s->sorted_codes, s->sorted_idx, tk, tv, n, 30, st);
    if (rs != RTQ_OK) return rs;

    // 2. PLOC
    hipLaunchKernelGGL(k_cluster_init, dim3(blocks_for(nc)), dim3(kThreads), 0, st, n, s->cparent,
                       s->cleft, s->cright, s->ctri);
    RTQ_HIP(ctx, hipGetLastError());
    hipLaunchKernelGGL(k_leaf_init, dim3(blocks_for(n)), dim3(kThreads), 0, st, s->tris_ref, s->models,
                       s->sorted_idx, n, s->cbox, s->ctri, s->ccount, c_in, b_in);
    RTQ_HIP(ctx, hipGetLastError());
    // passes are queued kPlocBatch at a time against a device-side cluster
    // count; the host reads the count and the pass log once per batch (passes
    // queued after the last merge find one cluster and change nothing)
    constexpr uint32_t kPlocBatch = 8;
    std::vector<std::pair<uint32_t, uint32_t>> pass_ranges;
    uint32_t *state = nullptr;
    uint4* pass_log = nullptr;
    uint2* scan_ws = nullptr;
    RTQ_HIP(ctx, tmp.get(&state, 2));
    RTQ_HIP(ctx, tmp.get(&pass_log, kPlocBatch));
    RTQ_HIP(ctx, tmp.get(&scan_ws, std::max<uint64_t>(scan_u32x2_workspace(n), 1)));
    const uint32_t h_state0[2] = {n, n};
    RTQ_HIP(ctx, hipMemcpyAsync(state, h_state0, sizeof h_state0, hipMemcpyHostToDevice, st));
    uint32_t iter = n, n_total = n;
    while (iter > 1) {
        const uint32_t bound = iter;
        for (uint32_t k = 0; k < kPlocBatch; ++k) {
            hipLaunchKernelGGL(k_ploc_nn, dim3(blocks_for(bound)), dim3(kThreads), 0, st, b_in, state,
                               prm.search_radius, nn);
            RTQ_HIP(ctx, hipGetLastError());
            hipLaunchKernelGGL(k_ploc_flags, dim3(blocks_for(bound)), dim3(kThreads), 0, st, nn, state, bound, flags);
            RTQ_HIP(ctx, hipGetLastError());
            rs = scan_u32x2_ws(ctx, flags, offs, bound, tot, scan_ws, st);
            if (rs != RTQ_OK) return rs;
            hipLaunchKernelGGL(k_ploc_merge, dim3(blocks_for(bound)), dim3(kThreads), 0, st, state, nn, flags, offs,
                               c_in, b_in, c_out, b_out, s->cbox, s->cleft, s->cright, s->cparent, s->ccount,
                               s->ctri);
            RTQ_HIP(ctx, hipGetLastError());
            hipLaunchKernelGGL(k_ploc_advance, dim3(1), dim3(1), 0, st, tot, state, pass_log, k);
            RTQ_HIP(ctx, hipGetLastError());
            std::swap(c_in, c_out);
            std::swap(b_in, b_out);
        }
        uint4 h_log[kPlocBatch];
        uint32_t h_state[2];
        RTQ_HIP(ctx, hipMemcpyAsync(h_log, pass_log, sizeof h_log, hipMemcpyDeviceToHost, st));
        RTQ_HIP(ctx, hipMemcpyAsync(h_state, state, sizeof h_state, hipMemcpyDeviceToHost, st));
        RTQ_HIP(ctx, hipStreamSynchronize(st));
        for (uint32_t k = 0; k < kPlocBatch; ++k) {
            const uint4 e = h_log[k];  // (first merge id, merges, clusters at the pass)
            if (e.z <= 1) break;
            if (e.y == 0) return fail(ctx, RTQ_EINTERNAL, "PLOC pass with %u clusters made no merge", e.z);
            pass_ranges.emplace_back(e.x, e.y);
        }
        iter = h_state[0];
        n_total = h_state[1];
    }
    if (n_total != nc) return fail(ctx, RTQ_EINTERNAL, "PLOC produced %u clusters, expected %u", n_total, nc);
    s->ploc_passes = (uint32_t)pass_ranges.size();
    s->passes = pass_ranges;

    // 3-4. node arrays of every layout
    rs = write_layouts(ctx, s, d_bounds, st);
    if (rs != RTQ_OK) return rs;
    RTQ_HIP(ctx, hipEventRecord(ev1, st));
    RTQ_HIP(ctx, hipStreamSynchronize(st));
    float ms = 0.0f;
    RTQ_HIP(ctx, hipEventElapsedTime(&ms, ev0, ev1));
    s->build_ms = ms;
    s->built = true;
    return RTQ_OK;
}

rtq_status write_top(rtq_ctx* ctx, Scene* s, hipStream_t st) {
    if (kTopNodes == 0) return RTQ_OK;
    // the set depends on the topology only, so a refit or an order update
    // rewrites the same buffer with the same node count
    if (!s->top) RTQ_HIP(ctx, dev_alloc(&s->top, 4ull * kTopNodes));
    uint32_t* d_n = nullptr;
    RTQ_HIP(ctx, dev_alloc(&d_n, 1));
    hipLaunchKernelGGL(k_top_nodes, dim3(1), dim3(1), 0, st, s->nodes, s->top, d_n);
    hipError_t e = hipGetLastError();
    uint32_t n = 0;
    if (e == hipSuccess) e = hipMemcpyAsync(&n, d_n, sizeof n, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    (void)hipFree(d_n);
    if (e != hipSuccess) return hip_fail(ctx, e, "write_top");
    s->n_top = n;
    return RTQ_OK;
}

// New boxes for the kept cluster tree (rtq_bvh_refit): leaf boxes from the
// current vertices, then the merges of each PLOC pass in build order (a
// merge's children come from earlier passes), then every node array again,
// written into the buffers it already has.
rtq_status refit_scene(rtq_ctx* ctx, Scene* s) {
    hipStream_t st = ctx->stream;
    const uint32_t n = s->n_tris;
    BuildTemps tmp;
    rtq_aabb* d_bounds = nullptr;
    RTQ_HIP(ctx, tmp.get(&d_bounds, kBoundsBoxes));
    hipLaunchKernelGGL(k_refit_leaves, dim3(blocks_for(n)), dim3(kThreads), 0, st, s->tris_ref, s->models, s->ctri,
                       n, s->cbox);
    RTQ_HIP(ctx, hipGetLastError());
    for (const auto& pr : s->passes) {
        hipLaunchKernelGGL(k_refit_merge, dim3(blocks_for(pr.second)), dim3(kThreads), 0, st, pr.first, pr.second,
                           s->cleft, s->cright, s->cbox);
        RTQ_HIP(ctx, hipGetLastError());
    }
    rtq_status rs = compute_bounds(ctx, s, s->params.flags & RTQ_BUILD_REF_PADDED_BOUNDS, d_bounds, st);
    if (rs != RTQ_OK) return rs;
    return write_layouts(ctx, s, d_bounds, st);
}

rtq_status morton_codes(rtq_ctx* ctx, Scene* s, uint32_t flags, uint32_t* host_codes) {
    hipStream_t st = ctx->stream;
    BuildTemps tmp;
    rtq_aabb* d_bounds = nullptr;
    uint32_t* codes = nullptr;
    RTQ_HIP(ctx, tmp.get(&d_bounds, kBoundsBoxes));
    RTQ_HIP(ctx, tmp.get(&codes, s->n_tris));
    rtq_status rs = compute_bounds(ctx, s, flags & RTQ_BUILD_REF_PADDED_BOUNDS, d_bounds, st);
    if (rs != RTQ_OK) return rs;
    hipLaunchKernelGGL(k_morton, dim3(blocks_for(s->n_tris)), dim3(kThreads), 0, st, s->tris_ref, s->models,
                       s->n_tris, d_bounds, codes, (uint32_t*)nullptr);
    RTQ_HIP(ctx, hipGetLastError());
    RTQ_HIP(ctx, hipMemcpyAsync(host_codes, codes, sizeof(uint32_t) * s->n_tris, hipMemcpyDeviceToHost, st));
    RTQ_HIP(ctx, hipStreamSynchronize(st));
    return RTQ_OK;
}

}  // namespace rtq
