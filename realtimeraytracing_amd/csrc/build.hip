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
__device__ __forceinline__ void refit_merge_one(uint32_t id, const uint32_t* __restrict__ cleft,
                                                const uint32_t* __restrict__ cright, rtq_aabb* __restrict__ cbox) {
    const rtq_aabb l = cbox[cleft[id]], r = cbox[cright[id]];
    rtq_aabb a{};
    for (int k = 0; k < 3; ++k) {
        a.min[k] = smin(l.min[k], r.min[k]);
        a.max[k] = smax(l.max[k], r.max[k]);
    }
    cbox[id] = a;
}

__global__ __launch_bounds__(kThreads) void k_refit_merge(uint32_t first, uint32_t count,
                                                          const uint32_t* __restrict__ cleft,
                                                          const uint32_t* __restrict__ cright,
                                                          rtq_aabb* __restrict__ cbox) {
    const uint32_t i = blockIdx.x * kThreads + threadIdx.x;
    if (i < count) refit_merge_one(first + i, cleft, cright, cbox);
}

// The passes [k0, k1) at the top of the tree in one workgroup, in build order,
// with the boxes of their clusters [c_lo, nc) in LDS (a child from an earlier
// pass gives its box from global memory); a barrier between passes, and the
// range's boxes written back at the end.  nc - c_lo <= kLdsClusters.
__global__ __launch_bounds__(kLdsThreads) void k_refit_merge_lds(const uint2* __restrict__ passes, uint32_t k0,
                                                                 uint32_t k1, uint32_t c_lo, uint32_t nc,
                                                                 const uint32_t* __restrict__ cleft,
                                                                 const uint32_t* __restrict__ cright,
                                                                 rtq_aabb* __restrict__ cbox) {
    constexpr int kSlots = kLdsClusters / kLdsThreads;
    __shared__ rtq_aabb s_box[kLdsClusters];
    const uint32_t t = threadIdx.x;
    const uint32_t R = nc - c_lo;
    uint32_t L[kSlots], Rc[kSlots];
#pragma unroll
    for (int q = 0; q < kSlots; ++q) {
        const uint32_t idx = t + kLdsThreads * q;
        L[q] = idx < R ? cleft[c_lo + idx] : 0u;
        Rc[q] = idx < R ? cright[c_lo + idx] : 0u;
    }
    for (uint32_t k = k0; k < k1; ++k) {
        const uint2 p = passes[k];
#pragma unroll
        for (int q = 0; q < kSlots; ++q) {
            const uint32_t idx = t + kLdsThreads * q, c = c_lo + idx;
            if (idx >= R || c < p.x || c >= p.x + p.y) continue;
            const rtq_aabb l = L[q] >= c_lo ? s_box[L[q] - c_lo] : cbox[L[q]];
            const rtq_aabb r = Rc[q] >= c_lo ? s_box[Rc[q] - c_lo] : cbox[Rc[q]];
            rtq_aabb m{};
            for (int d = 0; d < 3; ++d) {   // the merge of refit_merge_one
                m.min[d] = smin(l.min[d], r.min[d]);
                m.max[d] = smax(l.max[d], r.max[d]);
            }
            s_box[idx] = m;
        }
        __syncthreads();
    }
#pragma unroll
    for (int q = 0; q < kSlots; ++q) {
        const uint32_t idx = t + kLdsThreads * q;
        if (idx < R) cbox[c_lo + idx] = s_box[idx];
    }
}

// ------------------------------------------------------------- top of the tree

// One thread: BFS from the root over internal-node refs, at most kTopNodes
// nodes.  A child's slot is its position in the BFS queue, known when it is
// appended, so each node is read once and written with its refs into the set
// already rewritten to kTopBit | slot (leaf refs and refs leaving the set
// unchanged).
// One wave, level by level: lane l holds BFS slot l.  A level's lanes load
// their nodes together, then number their internal children in (slot, left
// before right) order with a wave prefix sum, which is the order a FIFO
// breadth-first walk appends them; children past kTopNodes keep their refs.
static_assert(kTopNodes <= 64, "one lane per top-of-tree slot");
__global__ __launch_bounds__(64) void k_top_nodes(const float4* __restrict__ nodes, float4* __restrict__ top,
                                                  uint32_t* __restrict__ n_top) {
    __shared__ uint32_t s_q[64];
    const uint32_t lane = threadIdx.x;
    uint32_t node = 0;                 // node index of slot `lane` once its level is reached
    uint32_t begin = 0, end = 1;       // the current level's slots
    while (begin < end) {
        const bool mine = lane >= begin && lane < end;
        float4 a{}, b{}, c{}, w{};
        uint32_t r0 = kLeafBit, r1 = kLeafBit;
        if (mine) {
            const float4* src = nodes + 4ull * node;
            a = src[0];
            b = src[1];
            c = src[2];
            w = src[3];
            r0 = __float_as_uint(w.x);
            r1 = __float_as_uint(w.y);
        }
        const uint32_t k0 = (r0 & kLeafBit) ? 0u : 1u, k1 = (r1 & kLeafBit) ? 0u : 1u;
        uint32_t incl = k0 + k1;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const uint32_t o = __shfl_up(incl, d, 64);
            if (lane >= (uint32_t)d) incl += o;
        }
        const uint32_t total = __shfl(incl, 63, 64);
        const uint32_t s0 = end + incl - (k0 + k1), s1 = s0 + k0;
        if (k0 && s0 < kTopNodes) {
            s_q[s0] = r0;
            r0 = kTopBit | s0;
        }
        if (k1 && s1 < kTopNodes) {
            s_q[s1] = r1;
            r1 = kTopBit | s1;
        }
        if (mine) {
            w.x = __uint_as_float(r0);
            w.y = __uint_as_float(r1);
            float4* dst = top + 4ull * lane;
            dst[0] = a;
            dst[1] = b;
            dst[2] = c;
            dst[3] = w;
        }
        __syncthreads();
        const uint32_t next = min(end + total, kTopNodes);
        if (lane >= end && lane < next) node = s_q[lane];
        __syncthreads();
        begin = end;
        end = next;
    }
    if (lane == 0) *n_top = end;
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
    // in range even after a failed sort (which fails the build)
    const uint32_t ti = min(sorted_idx[i], n - 1);
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

// ------------------------------------------------------- fused PLOC passes
//
// The single-launch PLOC iteration of PLOC++ (Benthin, Meister, Barczak,
// Mohanty & Wald, HPG 2022), replacing the neighbour search, flag, scan (two
// or three launches), merge and state-update launches of a pass.  Workgroup w
// owns the T = kPassThreads - 2r positions [wT, wT + T) of the active list. It
// stages the boxes of [wT - 2r, wT + T + 2r) in LDS and finds the nearest
// neighbour of every position of [wT - r, wT + T + r), so the partner of each
// owned position has its own neighbour at hand.  It then sets the merge and
// stay flags of its positions, scans them inside the workgroup and gets the
// counts before its tile by decoupled look-back over per-tile status words
// (Merrill & Garland, "Single-pass Parallel Prefix Scan with Decoupled
// Look-back", 2016).  Tiles are numbered in the order workgroups start (an
// atomic counter), so every tile a workgroup waits on is already running, and
// the wait is bounded (kSpinLimit) so a broken invariant ends in an error
// flag instead of a hang.  A status word carries its counts, so it is stored
// and loaded with relaxed device-scope atomics: acquire / release would add an
// L2 write-back or invalidate per poll.  A tile whose wait gave up publishes
// kTilePoison; a tile whose look-back reaches one poisons itself too and
// neither writes its clusters.  Each position goes through the arithmetic of
// k_ploc_nn, k_ploc_flags and k_ploc_merge, so the clusters are identical.
//
// Passes with at most kTailMax clusters (33 of the 61 passes of a 2M-triangle
// terrain) run in one workgroup, k_ploc_tail, the whole list held in LDS.
//
// Device state per pass slot: {clusters, next merge id, list buffer 0/1, 0};
// a pass reads slot k & 1 and writes slot (k + 1) & 1, so no workgroup reads
// a word another one writes.

constexpr int kPassThreads = 512;
constexpr int kTailThreads = 1024;
constexpr uint32_t kTailMax = 4 * kTailThreads;   // 4 positions per tail thread
constexpr uint64_t kTileAgg = 1ull << 62;         // tile status: own counts published
constexpr uint64_t kTilePre = 2ull << 62;         // tile status: counts through this tile
constexpr uint64_t kTilePoison = 3ull << 62;      // tile status: a wait at or before it gave up
constexpr uint64_t kTileState = 3ull << 62;
constexpr uint32_t kSpinLimit = 1u << 24;         // ~0.5 s of s_sleep(1) polls

enum : uint32_t { kPlocErrSpin = 1, kPlocErrNoNeighbour = 2, kPlocErrBound = 4, kPlocErrNoMerge = 8 };

struct PlocLists {
    uint32_t* c[2];
    Box6* b[2];
};

struct PlocClusters {
    rtq_aabb* box;
    uint32_t *left, *right, *parent, *count, *tri;
};

__device__ __forceinline__ uint64_t tile_word(uint64_t st, uint32_t merges, uint32_t stays) {
    return st | ((uint64_t)merges << 32) | stays;
}

// Exclusive sum over the workgroup; `total` is the workgroup's sum.  Every
// thread calls it; s_w holds one word per wave.
template <int THREADS>
__device__ __forceinline__ uint32_t block_exclusive_sum(uint32_t v, uint32_t* s_w, uint32_t& total) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t incl = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t o = __shfl_up(incl, d, 64);
        if (lane >= d) incl += o;
    }
    if (lane == 63) s_w[wave] = incl;
    __syncthreads();
    uint32_t before = 0, all = 0;
#pragma unroll
    for (int w = 0; w < THREADS / 64; ++w) {
        const uint32_t x = s_w[w];
        before += (w < wave) ? x : 0u;
        all += x;
    }
    total = all;
    return before + incl - v;
}

// Merge of clusters cid (box b) and cj (box bj) into cluster id, as
// k_ploc_merge writes it; b becomes the merged box.
// Merged cluster id of (cid, cj) with triangle count `count` (the sum of theirs).
__device__ __forceinline__ void ploc_merge_store(const PlocClusters& cl, uint32_t id, uint32_t cid, uint32_t cj,
                                                 uint32_t count, Box6& b, const Box6& bj) {
    Box6 m;
    rtq_aabb a{};
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        m.lo[k] = smin(b.lo[k], bj.lo[k]);
        m.hi[k] = smax(b.hi[k], bj.hi[k]);
        a.min[k] = m.lo[k];
        a.max[k] = m.hi[k];
    }
    cl.box[id] = a;
    cl.left[id] = cid;
    cl.right[id] = cj;
    cl.parent[cid] = id;
    cl.parent[cj] = id;
    cl.count[id] = count;
    cl.tri[id] = 0;
    b = m;
}

__device__ __forceinline__ void ploc_merge_write(const PlocClusters& cl, uint32_t id, uint32_t cid, uint32_t cj,
                                                 Box6& b, const Box6& bj) {
    // the counts are read before any store: the arrays may alias as far as
    // the compiler knows, and a load after a store waits for the store
    const uint32_t count = cl.count[cid] + cl.count[cj];
    ploc_merge_store(cl, id, cid, cj, count, b, bj);
}

struct PlocPassArgs {
    const uint32_t* st_in;
    uint32_t* st_out;
    uint32_t bound;       // clusters at the start of the batch: the grid covers them
    uint32_t radius;
    uint32_t* tile_ctr;   // zeroed before the pass
    uint64_t* tile_status;// one word per tile, zeroed before the pass
    uint32_t* err;
    uint4* log;           // (first merge id, merges, clusters, 1 if left to the tail)
    PlocLists lists;
    PlocClusters cl;
};

__global__ __launch_bounds__(kPassThreads) void k_ploc_pass(PlocPassArgs a) {
    __shared__ Box6 s_box[kPassThreads + 64];
    __shared__ uint32_t s_nn[kPassThreads];
    __shared__ uint32_t s_w[kPassThreads / 64];
    __shared__ uint32_t s_hdr[5];  // tile, clusters, next merge id, list buffer, skip
    __shared__ uint32_t s_excl[3]; // merges and stays before the tile, abort
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int r = (int)a.radius;
    const uint32_t T = (uint32_t)(kPassThreads - 2 * r);
    if (tid == 0) {
        const uint32_t iter = a.st_in[0], n_total = a.st_in[1], cur = a.st_in[2] & 1u;
        uint32_t skip = 0;
        if (__hip_atomic_load(a.err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0 || iter <= kTailMax) {
            skip = 1;
        } else if (iter > a.bound) {
            atomicOr(a.err, kPlocErrBound);
            skip = 1;
        }
        if (skip && blockIdx.x == 0) {  // left to the tail kernel (or stopped): carry the state over
            a.st_out[0] = iter;
            a.st_out[1] = n_total;
            a.st_out[2] = cur;
            *a.log = make_uint4(n_total, 0u, iter, 1u);
        }
        // the grid covers the batch's first pass: only as many workgroups as
        // this pass has tiles take one (the others leave without touching the
        // counter, whose atomics serialise), so the tiles are 0 .. need - 1
        const uint32_t need = (iter + T - 1) / T;
        if (!skip && blockIdx.x >= need) skip = 1;
        const uint32_t tile = skip ? 0u : atomicAdd(a.tile_ctr, 1u);
        s_hdr[0] = tile;
        s_hdr[1] = iter;
        s_hdr[2] = n_total;
        s_hdr[3] = cur;
        s_hdr[4] = skip;
    }
    __syncthreads();
    if (s_hdr[4]) return;
    const uint32_t tile = s_hdr[0], iter = s_hdr[1], n_total = s_hdr[2], cur = s_hdr[3];
    const uint32_t* __restrict__ c_in = a.lists.c[cur];
    const Box6* __restrict__ b_in = a.lists.b[cur];
    const int base = (int)(tile * T);
    const int lo = base - 2 * r;
    for (int q = tid; q < kPassThreads + 2 * r; q += kPassThreads) {
        const int p = lo + q;
        if (p >= 0 && p < (int)iter) s_box[q] = b_in[p];
    }
    __syncthreads();

    // nearest neighbour of position i = base - r + tid (k_ploc_nn); its box
    // is s_box[tid + r] and the window [i - r, i + r] is s_box[tid .. tid + 2r]
    const int i = base - r + tid;
    {
        uint32_t arg = kNone;
        if (i >= 0 && i < (int)iter) {
            const Box6 bi = s_box[tid + r];
            const int start = max(i - r, 0), end = min(i + r + 1, (int)iter);
            float best = INFINITY;
            // branch-free body, unrolled: the LDS reads of several j overlap
#pragma unroll 4
            for (int j = start; j < end; ++j) {
                const float d = merged_sa(bi, s_box[j - lo]);
                const bool take = j != i && d < best;
                best = take ? d : best;
                arg = take ? (uint32_t)j : arg;
            }
            if (arg == kNone) atomicOr(a.err, kPlocErrNoNeighbour);
        }
        s_nn[tid] = arg;
    }
    __syncthreads();

    // flags of the owned positions (k_ploc_flags): the partner j lies in
    // [i - r, i + r], so its neighbour is s_nn[tid + (j - i)]
    const bool owned = tid >= r && tid < r + (int)T && i < (int)iter;
    uint32_t fx = 0, fy = 0, j = kNone;
    if (owned) {
        j = s_nn[tid];
        const bool mutual = j != kNone && s_nn[tid + ((int)j - i)] == (uint32_t)i;
        fx = (mutual && (uint32_t)i < j) ? 1u : 0u;
        fy = (mutual && j < (uint32_t)i) ? 0u : 1u;
    }
    uint32_t total = 0;
    const uint32_t v = (fx << 16) | fy;  // at most kPassThreads of each per tile
    const uint32_t local = block_exclusive_sum<kPassThreads>(v, s_w, total);

    // counts before this tile: decoupled look-back by wave 0
    if (wave == 0) {
        const uint32_t ax = total >> 16, ay = total & 0xFFFFu;
        uint32_t ex = 0, ey = 0;
        bool failed = false;
        if (tile == 0) {
            if (lane == 0)
                __hip_atomic_store(&a.tile_status[0], tile_word(kTilePre, ax, ay), __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);
        } else {
            if (lane == 0)
                __hip_atomic_store(&a.tile_status[tile], tile_word(kTileAgg, ax, ay), __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);
            int64_t top = (int64_t)tile - 1;
            uint32_t spins = 0;
            for (;;) {
                const int64_t idx = top - lane;  // lane 0 = nearest tile
                uint64_t w = kTilePre;           // before tile 0: an empty prefix
                if (idx >= 0)
                    w = __hip_atomic_load(&a.tile_status[idx], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                while (__ballot((w & kTileState) == 0) != 0) {
                    if (++spins > kSpinLimit) {
                        failed = true;
                        break;
                    }
                    __builtin_amdgcn_s_sleep(1);
                    if ((w & kTileState) == 0)
                        w = __hip_atomic_load(&a.tile_status[idx], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                }
                if (failed) break;
                // nearest tile with counts through it (a prefix, or poison)
                const unsigned long long pre = __ballot((w & kTileState) >= kTilePre);
                const int stop = pre ? __ffsll((long long)pre) - 1 : 63;
                if (pre && __shfl((uint32_t)(w >> 62), stop, 64) == 3u) {
                    failed = true;
                    break;
                }
                uint32_t cx = lane <= stop ? (uint32_t)(w >> 32) & 0x3FFFFFFFu : 0u;
                uint32_t cy = lane <= stop ? (uint32_t)w : 0u;
#pragma unroll
                for (int d = 32; d >= 1; d >>= 1) {
                    cx += __shfl_xor(cx, d, 64);
                    cy += __shfl_xor(cy, d, 64);
                }
                ex += cx;
                ey += cy;
                if (pre) break;
                top -= 64;
            }
            if (lane == 0) {
                if (failed) atomicOr(a.err, kPlocErrSpin);
                __hip_atomic_store(&a.tile_status[tile],
                                   failed ? kTilePoison : tile_word(kTilePre, ex + ax, ey + ay), __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);
            }
        }
        if (lane == 0) {
            s_excl[0] = ex;
            s_excl[1] = ey;
            s_excl[2] = failed ? 1u : 0u;
            if (!failed && tile == (iter - 1) / T) {  // the tile holding the last position ends the pass
                a.st_out[0] = ey + ay;
                a.st_out[1] = n_total + ex + ax;
                a.st_out[2] = cur ^ 1u;
                *a.log = make_uint4(n_total, ex + ax, iter, 0u);
            }
        }
    }
    __syncthreads();
    if (s_excl[2] || !owned || !fy) return;

    // merge and compaction (k_ploc_merge)
    const uint32_t ox = s_excl[0] + (local >> 16), oy = s_excl[1] + (local & 0xFFFFu);
    uint32_t cid = c_in[i];
    Box6 b = s_box[tid + r];
    if (fx) {
        ploc_merge_write(a.cl, n_total + ox, cid, c_in[j], b, s_box[(int)j - lo]);
        cid = n_total + ox;
    }
    a.lists.c[cur ^ 1u][oy] = cid;
    a.lists.b[cur ^ 1u][oy] = b;
}

// All remaining passes once at most kTailMax clusters are left, in one
// workgroup with the list in LDS (same arithmetic as k_ploc_pass).  Thread t
// owns positions 4t .. 4t+3 for flags, scan and merge.  Does nothing unless
// 1 < clusters <= kTailMax.
__global__ __launch_bounds__(kTailThreads) void k_ploc_tail(uint32_t* __restrict__ state, uint32_t radius,
                                                            uint32_t* __restrict__ err, uint4* __restrict__ log,
                                                            uint32_t* __restrict__ log_count, PlocLists lists,
                                                            PlocClusters cl) {
    __shared__ Box6 s_box[kTailMax];
    __shared__ uint32_t s_c[kTailMax];
    __shared__ uint32_t s_n[kTailMax];   // triangle counts: no global loads between the passes' barriers
    __shared__ uint16_t s_nn[kTailMax];
    __shared__ uint32_t s_w[kTailThreads / 64];
    __shared__ uint32_t s_hdr[4];
    constexpr uint16_t kNone16 = 0xFFFFu;
    const int tid = threadIdx.x;
    const int r = (int)radius;
    if (tid == 0) {
        const uint32_t m0 = state[0];
        s_hdr[0] = m0;
        s_hdr[1] = state[1];
        s_hdr[2] = state[2] & 1u;
        s_hdr[3] = (*err != 0 || m0 <= 1 || m0 > kTailMax) ? 1u : 0u;
        if (s_hdr[3]) *log_count = 0;
    }
    __syncthreads();
    if (s_hdr[3]) return;
    uint32_t m = s_hdr[0], n_total = s_hdr[1];
    const uint32_t cur = s_hdr[2];
    for (uint32_t q = tid; q < m; q += kTailThreads) {
        s_box[q] = lists.b[cur][q];
        const uint32_t c = lists.c[cur][q];
        s_c[q] = c;
        s_n[q] = cl.count[c];
    }
    __syncthreads();
    uint32_t passes = 0;
    while (m > 1) {
        for (int i = tid; i < (int)m; i += kTailThreads) {
            const Box6 bi = s_box[i];
            const int start = max(i - r, 0), end = min(i + r + 1, (int)m);
            float best = INFINITY;
            uint16_t arg = kNone16;
#pragma unroll 4
            for (int j = start; j < end; ++j) {
                const float d = merged_sa(bi, s_box[j]);
                const bool take = j != i && d < best;
                best = take ? d : best;
                arg = take ? (uint16_t)j : arg;
            }
            if (arg == kNone16) atomicOr(err, kPlocErrNoNeighbour);
            s_nn[i] = arg;
        }
        __syncthreads();
        uint32_t f[4], sum = 0;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const uint32_t i = 4u * tid + k;
            f[k] = 0;
            if (i < m) {
                const uint32_t j = s_nn[i];
                const bool mutual = j != kNone16 && s_nn[j] == i;
                f[k] = ((mutual && i < j) ? 1u << 16 : 0u) | ((mutual && j < i) ? 0u : 1u);
            }
            sum += f[k];
        }
        uint32_t total = 0;
        uint32_t run = block_exclusive_sum<kTailThreads>(sum, s_w, total);
        Box6 ob[4];
        uint32_t oc[4], on[4], opos[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const uint32_t i = 4u * tid + k;
            opos[k] = kNone;
            if (i < m && (f[k] & 0xFFFFu)) {
                uint32_t cid = s_c[i], cnt = s_n[i];
                Box6 b = s_box[i];
                if (f[k] >> 16) {
                    const uint32_t j = s_nn[i], id = n_total + (run >> 16);
                    cnt += s_n[j];
                    ploc_merge_store(cl, id, cid, s_c[j], cnt, b, s_box[j]);
                    cid = id;
                }
                ob[k] = b;
                oc[k] = cid;
                on[k] = cnt;
                opos[k] = run & 0xFFFFu;
            }
            run += f[k];
        }
        __syncthreads();  // every read of this pass's list is done
#pragma unroll
        for (int k = 0; k < 4; ++k)
            if (opos[k] != kNone) {
                s_box[opos[k]] = ob[k];
                s_c[opos[k]] = oc[k];
                s_n[opos[k]] = on[k];
            }
        const uint32_t merges = total >> 16, stays = total & 0xFFFFu;
        if (tid == 0) log[passes] = make_uint4(n_total, merges, m, 0u);
        if (merges == 0) {  // uniform: no progress, the build reports it
            if (tid == 0) atomicOr(err, kPlocErrNoMerge);
            break;
        }
        n_total += merges;
        m = stays;
        ++passes;
        __syncthreads();
    }
    if (tid == 0) {
        state[0] = m;
        state[1] = n_total;
        *log_count = passes;
    }
}

__global__ void k_root_init(uint32_t root, uint32_t* pos, uint32_t* fl, uint32_t* dep) {
    pos[root] = 0;
    fl[root] = 0;
    dep[root] = 0;
}

__device__ __forceinline__ void place_one(uint32_t p, const uint32_t* __restrict__ cleft,
                                          const uint32_t* __restrict__ cright, const uint32_t* __restrict__ ccount,
                                          uint32_t* __restrict__ pos, uint32_t* __restrict__ fl,
                                          uint32_t* __restrict__ dep) {
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
    if (k < count) place_one(first + k, cleft, cright, ccount, pos, fl, dep);
}

// The passes [k0, k1) at the top of the tree in one workgroup, last pass
// first, with the placement state (position, first leaf, depth) of their
// clusters [c_lo, nc) in LDS: a pass reads and writes LDS only, a barrier
// between passes.  Children from earlier passes get their state in global
// memory; the clusters of the range write theirs back at the end.  Each
// thread keeps the child ids and left-child triangle counts of its clusters
// c_lo + t + kLdsThreads * s in registers.  nc - c_lo <= kLdsClusters.
__global__ __launch_bounds__(kLdsThreads) void k_place_lds(const uint2* __restrict__ passes, uint32_t k0, uint32_t k1,
                                                           uint32_t c_lo, uint32_t nc,
                                                           const uint32_t* __restrict__ cleft,
                                                           const uint32_t* __restrict__ cright,
                                                           const uint32_t* __restrict__ ccount,
                                                           uint32_t* __restrict__ pos, uint32_t* __restrict__ fl,
                                                           uint32_t* __restrict__ dep) {
    constexpr int kSlots = kLdsClusters / kLdsThreads;
    __shared__ uint32_t s_pos[kLdsClusters], s_fl[kLdsClusters], s_dep[kLdsClusters];
    const uint32_t t = threadIdx.x;
    const uint32_t R = nc - c_lo;
    uint32_t L[kSlots], Rc[kSlots], LC[kSlots];
#pragma unroll
    for (int q = 0; q < kSlots; ++q) {
        const uint32_t idx = t + kLdsThreads * q;
        L[q] = idx < R ? cleft[c_lo + idx] : 0u;
        Rc[q] = idx < R ? cright[c_lo + idx] : 0u;
    }
#pragma unroll
    for (int q = 0; q < kSlots; ++q) LC[q] = t + kLdsThreads * q < R ? ccount[L[q]] : 0u;
    if (t == 0) {  // the root, the last cluster
        s_pos[R - 1] = pos[nc - 1];
        s_fl[R - 1] = fl[nc - 1];
        s_dep[R - 1] = dep[nc - 1];
    }
    __syncthreads();
    for (uint32_t k = k1; k > k0; --k) {
        const uint2 p = passes[k - 1];
#pragma unroll
        for (int q = 0; q < kSlots; ++q) {
            const uint32_t idx = t + kLdsThreads * q, c = c_lo + idx;
            if (idx >= R || c < p.x || c >= p.x + p.y) continue;
            const uint32_t pp = s_pos[idx], pf = s_fl[idx], pd = s_dep[idx];
            const uint32_t l = L[q], r = Rc[q], lc = LC[q];
            // left child right after the parent, right child after the left subtree
            if (l >= c_lo) {
                s_pos[l - c_lo] = pp + 1; s_fl[l - c_lo] = pf; s_dep[l - c_lo] = pd + 1;
            } else {
                pos[l] = pp + 1; fl[l] = pf; dep[l] = pd + 1;
            }
            if (r >= c_lo) {
                s_pos[r - c_lo] = pp + 2 * lc; s_fl[r - c_lo] = pf + lc; s_dep[r - c_lo] = pd + 1;
            } else {
                pos[r] = pp + 2 * lc; fl[r] = pf + lc; dep[r] = pd + 1;
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int q = 0; q < kSlots; ++q) {
        const uint32_t idx = t + kLdsThreads * q;
        if (idx < R) {
            pos[c_lo + idx] = s_pos[idx];
            fl[c_lo + idx] = s_fl[idx];
            dep[c_lo + idx] = s_dep[idx];
        }
    }
}

// Leaf-ordered traversal triangles, and the depth of each workgroup's deepest
// leaf in block_depth[blockIdx.x] (k_max_u32 takes the maximum: atomics of
// every wave on one word would serialise, tens of thousands of them).
__global__ __launch_bounds__(kThreads) void k_write_leaves(
    uint32_t n_clusters, const uint32_t* __restrict__ ctri, const uint32_t* __restrict__ cleft,
    const uint32_t* __restrict__ fl, const uint32_t* __restrict__ dep, const float4* __restrict__ tri_world,
    float4* __restrict__ tris, uint32_t* __restrict__ block_depth) {
    __shared__ uint32_t s_max[kThreads / 64];
    const uint32_t c = blockIdx.x * kThreads + threadIdx.x;
    uint32_t depth = 0;
    if (c < n_clusters && cleft[c] == kNone) {
        const uint32_t t = ctri[c];
        const uint64_t dst = 3ull * fl[c];
        tris[dst + 0] = tri_world[3ull * t + 0];
        tris[dst + 1] = tri_world[3ull * t + 1];
        tris[dst + 2] = tri_world[3ull * t + 2];
        depth = dep[c];
    }
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) depth = max(depth, (uint32_t)__shfl_xor(depth, d, 64));
    if ((threadIdx.x & 63) == 0) s_max[threadIdx.x / 64] = depth;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t m = 0;
        for (uint32_t w = 0; w < kThreads / 64; ++w) m = max(m, s_max[w]);
        block_depth[blockIdx.x] = m;
    }
}

// out[0] = max of v[0..n) (one workgroup of 1024 lanes).
__global__ __launch_bounds__(1024) void k_max_u32(const uint32_t* __restrict__ v, uint32_t n,
                                                  uint32_t* __restrict__ out) {
    __shared__ uint32_t s_max[16];
    uint32_t m = 0;
    for (uint32_t i = threadIdx.x; i < n; i += 1024) m = max(m, v[i]);
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) m = max(m, (uint32_t)__shfl_xor(m, d, 64));
    if ((threadIdx.x & 63) == 0) s_max[threadIdx.x / 64] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (uint32_t w = 1; w < 16; ++w) m = max(m, s_max[w]);
        out[0] = m;
    }
}

// Reference node array (scene.cpp:189-208): cluster c at its pre-order
// position, children as positions, leaves with their triangle.
__global__ __launch_bounds__(kThreads) void k_write_ref(
    uint32_t n_clusters, const rtq_aabb* __restrict__ cbox, const uint32_t* __restrict__ ctri,
    const uint32_t* __restrict__ cleft, const uint32_t* __restrict__ cright, const uint32_t* __restrict__ pos,
    rtq_bvh_node* __restrict__ ref) {
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


// PLOC with fused passes: batches of kPlocBatch k_ploc_pass launches (each a
// no-op once at most kTailMax clusters are left) and one k_ploc_tail launch
// (a no-op until then), then one read of the state and the pass logs.  The
// active list starts in (c0, b0); c1, b1 are the other buffers.  A batch's
// passes all take the grid of its first pass; 16 passes per batch bring a
// 2M-triangle terrain (61 passes) to the tail in two host round trips.
rtq_status ploc_fused(rtq_ctx* ctx, Scene* s, uint32_t radius, Temps& tmp, uint32_t* c0, uint32_t* c1,
                      Box6* b0, Box6* b1, std::vector<std::pair<uint32_t, uint32_t>>& passes, hipStream_t st) {
    constexpr uint32_t kPlocBatch = 16;
    const uint32_t n = s->n_tris;
    const uint32_t T = kPassThreads - 2 * radius;
    const uint32_t max_tiles = (n + T - 1) / T;
    // ctl: kPlocBatch tile counters, then kPlocBatch x max_tiles status words
    const size_t ctl_words = 2 * kPlocBatch + (size_t)kPlocBatch * max_tiles * 2;
    uint32_t *ctl = nullptr, *state = nullptr, *err = nullptr, *tail_count = nullptr;
    uint4 *log = nullptr, *tail_log = nullptr;
    RTQ_HIP(ctx, tmp.get(&ctl, ctl_words));
    // state (8 words), error and tail pass count (8), pass log: one block,
    // read back with one copy per batch
    constexpr uint32_t kHeadWords = 16 + 4 * kPlocBatch;
    RTQ_HIP(ctx, tmp.get(&state, kHeadWords));
    err = state + 8;
    log = reinterpret_cast<uint4*>(state + 16);
    RTQ_HIP(ctx, tmp.get(&tail_log, kTailMax));
    tail_count = err + 1;
    const uint32_t h_state0[8] = {n, n, 0u, 0u, 0u, 0u, 0u, 0u};
    RTQ_HIP(ctx, hipMemcpyAsync(state, h_state0, sizeof h_state0, hipMemcpyHostToDevice, st));
    RTQ_HIP(ctx, hipMemsetAsync(err, 0, 2 * sizeof(uint32_t), st));
    PlocLists lists{{c0, c1}, {b0, b1}};
    PlocClusters cl{s->cbox, s->cleft, s->cright, s->cparent, s->ccount, s->ctri};
    uint32_t iter = n, n_total = n;
    std::vector<uint4> h_tail;
    while (iter > 1) {
        const uint32_t bound = iter;
        if (iter > kTailMax) {
            const uint32_t tiles = (bound + T - 1) / T;
            RTQ_HIP(ctx, hipMemsetAsync(ctl, 0, sizeof(uint32_t) * (2ull * kPlocBatch + 2ull * kPlocBatch * tiles), st));
            uint64_t* status = reinterpret_cast<uint64_t*>(ctl + 2 * kPlocBatch);
            for (uint32_t k = 0; k < kPlocBatch; ++k) {
                PlocPassArgs a{state + 4 * (k & 1), state + 4 * ((k + 1) & 1), bound, radius, ctl + k,
                               status + (size_t)k * tiles, err, log + k, lists, cl};
                hipLaunchKernelGGL(k_ploc_pass, dim3(tiles), dim3(kPassThreads), 0, st, a);
                RTQ_HIP(ctx, hipGetLastError());
            }
        }
        // the state of the batch's last pass is in slot 0 (kPlocBatch is even)
        hipLaunchKernelGGL(k_ploc_tail, dim3(1), dim3(kTailThreads), 0, st, state, radius, err, tail_log, tail_count,
                           lists, cl);
        RTQ_HIP(ctx, hipGetLastError());
        // the head block and the first tail log entries in one read (entries
        // past the tail's pass count are not used)
        constexpr uint32_t kTailPeek = 128;
        uint32_t h_head[kHeadWords];
        uint4 h_peek[kTailPeek];
        HostReads rd(ctx, st);
        RTQ_HIP(ctx, rd.add(h_head, state, sizeof h_head));
        RTQ_HIP(ctx, rd.add(h_peek, tail_log, sizeof h_peek));
        RTQ_HIP(ctx, rd.wait());
        const uint32_t* h_state = h_head;
        const uint32_t* h_err = h_head + 8;
        uint4 h_log[kPlocBatch];
        memcpy(h_log, h_head + 16, sizeof h_log);
        if (h_err[0] & kPlocErrNoMerge) return fail(ctx, RTQ_EINTERNAL, "PLOC pass made no merge");
        if (h_err[0])
            return fail(ctx, RTQ_EINTERNAL, "fused PLOC pass failed (error bits 0x%x)", h_err[0]);
        if (iter > kTailMax)
            for (uint32_t k = 0; k < kPlocBatch; ++k) {
                const uint4 e = h_log[k];  // (first merge id, merges, clusters at the pass, left to the tail)
                if (e.w) break;
                if (e.y == 0) return fail(ctx, RTQ_EINTERNAL, "PLOC pass with %u clusters made no merge", e.z);
                passes.emplace_back(e.x, e.y);
            }
        const uint32_t nt = h_err[1];
        if (nt > kTailMax) return fail(ctx, RTQ_EINTERNAL, "PLOC tail ran %u passes", nt);
        if (nt) {
            h_tail.resize(nt);
            if (nt <= kTailPeek) {
                memcpy(h_tail.data(), h_peek, sizeof(uint4) * nt);
            } else {
                RTQ_HIP(ctx, rd.add(h_tail.data(), tail_log, sizeof(uint4) * nt));
                RTQ_HIP(ctx, rd.wait());
            }
            for (const uint4& e : h_tail) passes.emplace_back(e.x, e.y);
        }
        if (h_state[0] >= iter && h_state[0] > 1)
            return fail(ctx, RTQ_EINTERNAL, "PLOC made no progress at %u clusters", iter);
        iter = h_state[0];
        n_total = h_state[1];
    }
    if (n_total != 2 * n - 1) return fail(ctx, RTQ_EINTERNAL, "PLOC produced %u clusters, expected %u", n_total, 2 * n - 1);
    return RTQ_OK;
}

// PLOC with one launch per step of a pass: neighbour search, flags, scan,
// merge and state update (RTQ_BUILD_PLOC_SPLIT).  Passes are queued kPlocBatch
// at a time against a device-side cluster count; the host reads the count and
// the pass log once per batch (passes queued after the last merge find one
// cluster and change nothing).
rtq_status ploc_split(rtq_ctx* ctx, Scene* s, uint32_t radius, Temps& tmp, uint32_t* c_in, uint32_t* c_out,
                      Box6* b_in, Box6* b_out, std::vector<std::pair<uint32_t, uint32_t>>& pass_ranges,
                      hipStream_t st) {
    constexpr uint32_t kPlocBatch = 8;
    const uint32_t n = s->n_tris;
    uint32_t *nn = nullptr, *state = nullptr;
    uint2 *flags = nullptr, *offs = nullptr, *tot = nullptr, *scan_ws = nullptr;
    uint4* pass_log = nullptr;
    RTQ_HIP(ctx, tmp.get(&nn, n));
    RTQ_HIP(ctx, tmp.get(&flags, n));
    RTQ_HIP(ctx, tmp.get(&offs, n));
    RTQ_HIP(ctx, tmp.get(&tot, 1));
    RTQ_HIP(ctx, tmp.get(&state, 2));
    RTQ_HIP(ctx, tmp.get(&pass_log, kPlocBatch));
    RTQ_HIP(ctx, tmp.get(&scan_ws, std::max<uint64_t>(scan_u32x2_workspace(n), 1)));
    RTQ_HIP(ctx, hipMemsetAsync(nn, 0, sizeof(uint32_t) * n, st));
    const uint32_t h_state0[2] = {n, n};
    RTQ_HIP(ctx, hipMemcpyAsync(state, h_state0, sizeof h_state0, hipMemcpyHostToDevice, st));
    uint32_t iter = n, n_total = n;
    while (iter > 1) {
        const uint32_t bound = iter;
        for (uint32_t k = 0; k < kPlocBatch; ++k) {
            hipLaunchKernelGGL(k_ploc_nn, dim3(blocks_for(bound)), dim3(kThreads), 0, st, b_in, state,
                               radius, nn);
            RTQ_HIP(ctx, hipGetLastError());
            hipLaunchKernelGGL(k_ploc_flags, dim3(blocks_for(bound)), dim3(kThreads), 0, st, nn, state, bound, flags);
            RTQ_HIP(ctx, hipGetLastError());
            const rtq_status rs = scan_u32x2_ws(ctx, flags, offs, bound, tot, scan_ws, st);
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
        HostReads rd(ctx, st);
        RTQ_HIP(ctx, rd.add(h_log, pass_log, sizeof h_log));
        RTQ_HIP(ctx, rd.add(h_state, state, sizeof h_state));
        RTQ_HIP(ctx, rd.wait());
        for (uint32_t k = 0; k < kPlocBatch; ++k) {
            const uint4 e = h_log[k];  // (first merge id, merges, clusters at the pass)
            if (e.z <= 1) break;
            if (e.y == 0) return fail(ctx, RTQ_EINTERNAL, "PLOC pass with %u clusters made no merge", e.z);
            pass_ranges.emplace_back(e.x, e.y);
        }
        iter = h_state[0];
        n_total = h_state[1];
    }
    if (n_total != 2 * n - 1)
        return fail(ctx, RTQ_EINTERNAL, "PLOC produced %u clusters, expected %u", n_total, 2 * n - 1);
    return RTQ_OK;
}

// Steps after PLOC, shared by the build and the refit: pre-order placement,
// the reference node array and leaf-ordered triangles, the box pad from the
// scene extent, the binary traversal nodes and (node_width 8) the compressed
// wide nodes.  Buffers that exist already are written in place (a refit keeps
// their addresses, so captured batch graphs stay valid).
// Work queued on the context's second stream: fork(st) makes it wait for
// what st has queued, join(st) makes st wait for it.  Forked work that is
// never joined (an early return) is waited for on the host when the guard
// goes out of scope, before the temporaries it uses are given back.
class SideWork {
public:
    explicit SideWork(rtq_ctx* ctx) : ctx_(ctx) {}
    SideWork(const SideWork&) = delete;
    SideWork& operator=(const SideWork&) = delete;
    ~SideWork() {
        if (forked_ && !joined_) (void)hipStreamSynchronize(ctx_->side);
    }
    hipError_t fork(hipStream_t st) {
        hipError_t e = hipEventRecord(ctx_->fork, st);
        if (e == hipSuccess) e = hipStreamWaitEvent(ctx_->side, ctx_->fork, 0);
        forked_ = true;  // even on an error: the destructor then waits for the side stream
        return e;
    }
    hipError_t join(hipStream_t st) {
        if (!forked_ || joined_) return hipSuccess;
        hipError_t e = hipEventRecord(ctx_->join, ctx_->side);
        if (e == hipSuccess) e = hipStreamWaitEvent(st, ctx_->join, 0);
        if (e == hipSuccess) joined_ = true;
        return e;
    }

private:
    rtq_ctx* ctx_;
    bool forked_ = false, joined_ = false;
};

// refit: the tree is the one the last build wrote the layouts for, with new
// boxes: its pre-order placement is kept and the 8-wide nodes keep their
// children.  flag_dev (optional): one more device word to read back with the
// layout's own host read, into *flag_host.
rtq_status write_layouts(rtq_ctx* ctx, Scene* s, const rtq_aabb* d_bounds, hipStream_t st, bool refit,
                         const uint32_t* flag_dev = nullptr, uint32_t* flag_host = nullptr) {
    const uint32_t n = s->n_tris;
    const uint32_t nc = s->n_clusters;
    Temps tmp(ctx);
    uint32_t *maxd = nullptr, *flag = nullptr, *bidx = nullptr, *ntrav = nullptr;
    // pre-order positions, first leaf ranks and depths stay with the scene:
    // the reference node array is written from them on demand
    // (ensure_ref_nodes), and a refit keeps them
    if (!s->preorder) RTQ_HIP(ctx, dev_alloc(&s->preorder, nc));
    if (!s->first_leaf) RTQ_HIP(ctx, dev_alloc(&s->first_leaf, nc));
    if (!s->node_depth) RTQ_HIP(ctx, dev_alloc(&s->node_depth, nc));
    // so is every cluster's traversal node index: a refit needs no new marks
    if (!s->trav_index) RTQ_HIP(ctx, dev_alloc(&s->trav_index, nc));
    uint32_t* const pos = s->preorder;
    uint32_t* const fl = s->first_leaf;
    uint32_t* const dep = s->node_depth;
    bidx = s->trav_index;
    s->ref_ready = false;
    RTQ_HIP(ctx, tmp.get(&maxd, 2));
    uint32_t* bdep = nullptr;
    RTQ_HIP(ctx, tmp.get(&bdep, blocks_for(nc)));
    RTQ_HIP(ctx, tmp.get(&flag, nc));
    RTQ_HIP(ctx, tmp.get(&ntrav, 1));
    const uint32_t root = nc - 1;
    // The 8-wide collapse's cost table reads the clusters only: it runs on the
    // second stream beside the placement and the binary nodes, its launches
    // queued between the placement's so that both streams start at once.
    float* wcost = nullptr;
    uint32_t* wchoice = nullptr;
    SideWork side(ctx);
    const uint32_t dp_steps = s->params.node_width == 8 && !refit ? wide_cost_steps(s) : 0u;
    uint32_t dp_next = 0;
    auto queue_cost = [&](uint32_t upto) -> rtq_status {
        upto = std::min(upto, dp_steps);
        if (dp_next >= upto) return RTQ_OK;
        const rtq_status r = wide_cost_table(ctx, s, wcost, wchoice, dp_next, upto, ctx->side);
        dp_next = upto;
        return r;
    };
    if (dp_steps) {
        RTQ_HIP(ctx, tmp.get(&wcost, (size_t)(nc - n) * 8));
        RTQ_HIP(ctx, tmp.get(&wchoice, nc - n));
        RTQ_HIP(ctx, side.fork(st));
    }
    rtq_status rs = RTQ_OK;
    if (!refit) {
        hipLaunchKernelGGL(k_root_init, dim3(1), dim3(1), 0, st, root, pos, fl, dep);
        RTQ_HIP(ctx, hipGetLastError());
        // the passes at the top of the tree in one launch (their clusters'
        // state in LDS), then one launch per larger pass, last pass first
        const uint32_t n_passes = (uint32_t)s->passes.size();
        if (s->lds_from < n_passes) {
            hipLaunchKernelGGL(k_place_lds, dim3(1), dim3(kLdsThreads), 0, st, s->d_passes, s->lds_from, n_passes,
                               s->passes[s->lds_from].first, nc, s->cleft, s->cright, s->ccount, pos, fl, dep);
            RTQ_HIP(ctx, hipGetLastError());
        }
        rs = queue_cost(2);
        if (rs != RTQ_OK) return rs;
        for (uint32_t k = s->lds_from; k > 0; --k) {
            const auto& pr = s->passes[k - 1];
            hipLaunchKernelGGL(k_place_pass, dim3(blocks_for(pr.second)), dim3(kThreads), 0, st, pr.first,
                               pr.second, s->cleft, s->cright, s->ccount, pos, fl, dep);
            RTQ_HIP(ctx, hipGetLastError());
            rs = queue_cost(dp_next + 1);
            if (rs != RTQ_OK) return rs;
        }
    }
    hipLaunchKernelGGL(k_write_leaves, dim3(blocks_for(nc)), dim3(kThreads), 0, st, nc, s->ctri, s->cleft, fl,
                       dep, s->tri_world, s->tris, bdep);
    RTQ_HIP(ctx, hipGetLastError());
    // the depth and the traversal node indices depend on the topology only:
    // a refit keeps them
    if (!refit) {
        hipLaunchKernelGGL(k_max_u32, dim3(1), dim3(1024), 0, st, bdep, blocks_for(nc), maxd);
        RTQ_HIP(ctx, hipGetLastError());
        if (n > 1) {
            hipLaunchKernelGGL(k_mark_trav, dim3(blocks_for(nc)), dim3(kThreads), 0, st, nc, root,
                               s->params.max_leaf_size, s->cleft, s->ccount, pos, flag);
            RTQ_HIP(ctx, hipGetLastError());
            rs = scan_u32(ctx, flag, bidx, nc, ntrav, st);
            if (rs != RTQ_OK) return rs;
        }
    }

    // one host read: scene bounds (box pad), and after a build the tree depth
    // and the traversal node count
    rtq_aabb h_bounds[2];
    uint32_t h_depth = s->depth, h_ntrav = refit ? s->n_trav : 1u;
    HostReads rd(ctx, st);
    RTQ_HIP(ctx, rd.add(h_bounds, d_bounds, refit ? sizeof(rtq_aabb) : sizeof h_bounds));   // a refit: the root box
    if (!refit) RTQ_HIP(ctx, rd.add(&h_depth, maxd, sizeof h_depth));
    if (!refit && n > 1) RTQ_HIP(ctx, rd.add(&h_ntrav, ntrav, sizeof h_ntrav));
    if (flag_dev) RTQ_HIP(ctx, rd.add(flag_host, flag_dev, sizeof(uint32_t)));
    rs = queue_cost(dp_steps);
    if (rs != RTQ_OK) return rs;
    RTQ_HIP(ctx, rd.wait());
    float extent = 0.0f;
    for (int k = 0; k < 3; ++k)
        extent = std::max(extent, std::max(std::fabs(h_bounds[0].min[k]), std::fabs(h_bounds[0].max[k])));
    s->box_pad = std::max(std::ldexp(extent, -18), 1e-30f);
    s->depth = h_depth;

    const uint32_t max_leaf = s->params.max_leaf_size;
    if (s->nodes && h_ntrav != s->n_trav)
        return fail(ctx, RTQ_EINTERNAL, "refit: %u traversal nodes, the tree has %u", h_ntrav, s->n_trav);
    s->n_trav = h_ntrav;
    if (!s->nodes) RTQ_HIP(ctx, take(s->kept_nodes, &s->nodes, 4ull * h_ntrav));
    if (!s->node_aux) RTQ_HIP(ctx, take(s->kept_aux, &s->node_aux, h_ntrav));
    if (n == 1) {
        hipLaunchKernelGGL(k_write_single, dim3(1), dim3(1), 0, st, s->cbox, s->box_pad, s->nodes, s->node_aux);
        RTQ_HIP(ctx, hipGetLastError());
    } else {
        hipLaunchKernelGGL(k_write_trav, dim3(blocks_for(nc)), dim3(kThreads), 0, st, nc, root, max_leaf,
                           s->box_pad, s->cbox, s->cleft, s->cright, s->ccount, pos, fl, bidx, dep, s->nodes,
                           s->node_aux);
        RTQ_HIP(ctx, hipGetLastError());
    }
    if (s->params.node_width == 8) {
        if (refit) {
            rs = refit_wide(ctx, s, st);
        } else {
            RTQ_HIP(ctx, side.join(st));
            rs = build_wide(ctx, s, wchoice, fl, pos, n == 1 ? nullptr : bidx, st);
        }
        if (rs != RTQ_OK) return rs;
    }
    rs = write_top(ctx, s, st);
    if (rs != RTQ_OK) return rs;
    RTQ_HIP(ctx, hipStreamSynchronize(st));
    return RTQ_OK;
}

}  // namespace

rtq_status ensure_ref_nodes(rtq_ctx* ctx, Scene* s, hipStream_t st) {
    if (s->ref_ready) return RTQ_OK;
    const uint32_t nc = s->n_clusters;
    if (!s->built || !s->preorder || nc == 0)
        return fail(ctx, RTQ_EINTERNAL, "reference nodes: scene has no placed tree");
    if (!s->ref_nodes) RTQ_HIP(ctx, dev_alloc(&s->ref_nodes, nc));
    if (!s->ref_nodes) return fail(ctx, RTQ_ENOMEM, "reference nodes: allocation failed");
    hipLaunchKernelGGL(k_write_ref, dim3(blocks_for(nc)), dim3(kThreads), 0, st, nc, s->cbox, s->ctri, s->cleft,
                       s->cright, s->preorder, s->ref_nodes);
    RTQ_HIP(ctx, hipGetLastError());
    s->ref_ready = true;
    return RTQ_OK;
}

rtq_status build_scene(rtq_ctx* ctx, Scene* s, const rtq_build_params& prm) {
    hipStream_t st = ctx->stream;
    const uint32_t n = s->n_tris;
    const uint32_t nc = 2 * n - 1;
    // work queued earlier may still read the buffers a rebuild overwrites
    RTQ_HIP(ctx, hipDeviceSynchronize());
    reset_build(s);
    s->params = prm;
    s->n_clusters = nc;
    // size the context's scratch for this build's temporaries up front, so a
    // first build does not allocate them one by one.  Alive together, per
    // triangle: sort keys and values and the PLOC lists 64 B, the traversal
    // node marks and indices 16 B; with 8-wide nodes also the collapse cost
    // table 36 B and the collapse's two level lists 8 B per binary traversal
    // node (one per triangle with leaves of one triangle); the fused PLOC
    // passes' tile status words (16 passes x 8 B per tile of 480), each
    // placement workgroup's deepest leaf and the collapse's tile status
    // words, under n / 2 B; 1 MiB for scans and 256-B rounding
    const size_t per_tri = prm.node_width == 8 ? 64 + 16 + 36 + 8 : 64 + 16;
    ctx->arena.peak = std::max<size_t>(ctx->arena.peak, per_tri * n + n / 2 + ((size_t)1 << 20));
    rtq_status rs = arena_reserve(ctx);
    if (rs != RTQ_OK) return rs;

    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    RTQ_HIP(ctx, hipEventCreate(&ev0));
    RTQ_HIP(ctx, hipEventCreate(&ev1));
    struct EvGuard {
        hipEvent_t a, b;
        ~EvGuard() { (void)hipEventDestroy(a); (void)hipEventDestroy(b); }
    } evg{ev0, ev1};
    RTQ_HIP(ctx, hipEventRecord(ev0, st));

    // persistent products, sized by the triangle count: a rebuild rewrites them
    if (!s->cbox) RTQ_HIP(ctx, dev_alloc(&s->cbox, nc));
    if (!s->ctri) RTQ_HIP(ctx, dev_alloc(&s->ctri, nc));
    if (!s->cleft) RTQ_HIP(ctx, dev_alloc(&s->cleft, nc));
    if (!s->cright) RTQ_HIP(ctx, dev_alloc(&s->cright, nc));
    if (!s->cparent) RTQ_HIP(ctx, dev_alloc(&s->cparent, nc));
    if (!s->ccount) RTQ_HIP(ctx, dev_alloc(&s->ccount, nc));
    if (!s->sorted_idx) RTQ_HIP(ctx, dev_alloc(&s->sorted_idx, n));
    if (!s->sorted_codes) RTQ_HIP(ctx, dev_alloc(&s->sorted_codes, n));
    if (!s->tris) RTQ_HIP(ctx, dev_alloc(&s->tris, 3ull * n));

    Temps tmp(ctx);
    rtq_aabb* d_bounds = nullptr;
    uint32_t *tk = nullptr, *tv = nullptr, *c_in = nullptr, *c_out = nullptr;
    Box6 *b_in = nullptr, *b_out = nullptr;
    RTQ_HIP(ctx, tmp.get(&d_bounds, kBoundsBoxes));
    RTQ_HIP(ctx, tmp.get(&tk, n));
    RTQ_HIP(ctx, tmp.get(&tv, n));
    RTQ_HIP(ctx, tmp.get(&c_in, n));
    RTQ_HIP(ctx, tmp.get(&c_out, n));
    RTQ_HIP(ctx, tmp.get(&b_in, n));
    RTQ_HIP(ctx, tmp.get(&b_out, n));
    uint32_t* sort_err = nullptr;
    RTQ_HIP(ctx, tmp.get(&sort_err, 1));
    RTQ_HIP(ctx, hipMemsetAsync(sort_err, 0, sizeof(uint32_t), st));

    // 1. bounds, cube, Morton codes, sort
    rs = compute_bounds(ctx, s, prm.flags & RTQ_BUILD_REF_PADDED_BOUNDS, d_bounds, st);
    if (rs != RTQ_OK) return rs;
    hipLaunchKernelGGL(k_morton, dim3(blocks_for(n)), dim3(kThreads), 0, st, s->tris_ref, s->models, n,
                       d_bounds, s->sorted_codes, s->sorted_idx);
    RTQ_HIP(ctx, hipGetLastError());
    rs = radix_sort_pairs(ctx, s->sorted_codes, s->sorted_idx, tk, tv, n, 30, sort_err, st);
    if (rs != RTQ_OK) return rs;

    // 2. PLOC
    hipLaunchKernelGGL(k_cluster_init, dim3(blocks_for(nc)), dim3(kThreads), 0, st, n, s->cparent,
                       s->cleft, s->cright, s->ctri);
    RTQ_HIP(ctx, hipGetLastError());
    hipLaunchKernelGGL(k_leaf_init, dim3(blocks_for(n)), dim3(kThreads), 0, st, s->tris_ref, s->models,
                       s->sorted_idx, n, s->cbox, s->ctri, s->ccount, c_in, b_in);
    RTQ_HIP(ctx, hipGetLastError());
    std::vector<std::pair<uint32_t, uint32_t>> pass_ranges;
    rs = (prm.flags & RTQ_BUILD_PLOC_SPLIT)
             ? ploc_split(ctx, s, prm.search_radius, tmp, c_in, c_out, b_in, b_out, pass_ranges, st)
             : ploc_fused(ctx, s, prm.search_radius, tmp, c_in, c_out, b_in, b_out, pass_ranges, st);
    if (rs != RTQ_OK) return rs;
    s->ploc_passes = (uint32_t)pass_ranges.size();
    s->passes = pass_ranges;
    // device copy of the pass ranges for the single-workgroup launches of the
    // top passes (placement, collapse costs, refit merges): the passes from
    // lds_from on create at most kLdsClusters clusters in all
    s->lds_from = s->ploc_passes;
    while (s->lds_from > 0 && nc - s->passes[s->lds_from - 1].first <= kLdsClusters) --s->lds_from;
    if (s->ploc_passes > 0) {
        s->h_passes.resize(s->ploc_passes);
        for (uint32_t k = 0; k < s->ploc_passes; ++k)
            s->h_passes[k] = make_uint2(s->passes[k].first, s->passes[k].second);
        RTQ_HIP(ctx, take(s->kept_passes, &s->d_passes, s->ploc_passes));
        RTQ_HIP(ctx, hipMemcpyAsync(s->d_passes, s->h_passes.data(), sizeof(uint2) * s->ploc_passes,
                                    hipMemcpyHostToDevice, st));
    }

    // 3-4. node arrays of every layout
    uint32_t h_sort_err = 0;
    rs = write_layouts(ctx, s, d_bounds, st, false, sort_err, &h_sort_err);
    if (rs != RTQ_OK) return rs;
    if (h_sort_err) return fail(ctx, RTQ_EINTERNAL, "build: a radix sort pass's look-back gave up");
    RTQ_HIP(ctx, hipEventRecord(ev1, st));
    RTQ_HIP(ctx, hipEventSynchronize(ev1));
    float ms = 0.0f;
    RTQ_HIP(ctx, hipEventElapsedTime(&ms, ev0, ev1));
    s->build_ms = ms;
    s->built = true;
    s->generation = ++ctx->builds;
    // temporaries that did not fit the scratch reserved above raised its peak
    s->scratch_spill = ctx->arena.peak > ctx->arena.cap ? ctx->arena.peak - ctx->arena.cap : 0;
    return RTQ_OK;
}

rtq_status write_top(rtq_ctx* ctx, Scene* s, hipStream_t st) {
    if (kTopNodes == 0) return RTQ_OK;
    // the set depends on the topology only, so a refit or an order update
    // rewrites the same buffer with the same node count
    // one float4 past the nodes holds the node count
    if (!s->top) RTQ_HIP(ctx, dev_alloc(&s->top, 4ull * kTopNodes + 1));
    uint32_t* d_n = reinterpret_cast<uint32_t*>(s->top + 4ull * kTopNodes);
    hipLaunchKernelGGL(k_top_nodes, dim3(1), dim3(64), 0, st, s->nodes, s->top, d_n);
    hipError_t e = hipGetLastError();
    uint32_t n = 0;
    HostReads rd(ctx, st);
    if (e == hipSuccess) e = rd.add(&n, d_n, sizeof n);
    if (e == hipSuccess) e = rd.wait();
    if (e != hipSuccess) return hip_fail(ctx, e, "write_top");
    s->n_top = n;
    return RTQ_OK;
}

// New boxes for the kept cluster tree (rtq_bvh_refit): leaf boxes from the
// current vertices, then the merges of each PLOC pass in build order (a
// merge's children come from earlier passes), then the leaf triangles and the
// binary nodes again and new boxes in the kept 8-wide nodes, written into the
// buffers they already have.
rtq_status refit_scene(rtq_ctx* ctx, Scene* s) {
    hipStream_t st = ctx->stream;
    const uint32_t n = s->n_tris;
    const rtq_status ra = arena_reserve(ctx);
    if (ra != RTQ_OK) return ra;
    hipLaunchKernelGGL(k_refit_leaves, dim3(blocks_for(n)), dim3(kThreads), 0, st, s->tris_ref, s->models, s->ctri,
                       n, s->cbox);
    RTQ_HIP(ctx, hipGetLastError());
    const uint32_t n_passes = (uint32_t)s->passes.size();
    for (uint32_t k = 0; k < s->lds_from; ++k) {
        const auto& pr = s->passes[k];
        hipLaunchKernelGGL(k_refit_merge, dim3(blocks_for(pr.second)), dim3(kThreads), 0, st, pr.first, pr.second,
                           s->cleft, s->cright, s->cbox);
        RTQ_HIP(ctx, hipGetLastError());
    }
    if (s->lds_from < n_passes) {
        hipLaunchKernelGGL(k_refit_merge_lds, dim3(1), dim3(kLdsThreads), 0, st, s->d_passes, s->lds_from, n_passes,
                           s->passes[s->lds_from].first, s->n_clusters, s->cleft, s->cright, s->cbox);
        RTQ_HIP(ctx, hipGetLastError());
    }
    // the box pad needs the largest coordinate magnitude of the scene bounds:
    // the root cluster's new box holds the same minima and maxima (merges are
    // exact), and the reference quirk of adding the origin cannot raise it
    return write_layouts(ctx, s, s->cbox + (s->n_clusters - 1), st, true);
}

rtq_status morton_codes(rtq_ctx* ctx, Scene* s, uint32_t flags, uint32_t* host_codes) {
    hipStream_t st = ctx->stream;
    RTQ_HIP(ctx, hipDeviceSynchronize());  // nothing queued still uses the arena
    const rtq_status ra = arena_reserve(ctx);
    if (ra != RTQ_OK) return ra;
    Temps tmp(ctx);
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
