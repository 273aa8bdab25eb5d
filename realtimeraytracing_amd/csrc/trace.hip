// BVH traversal + ray/triangle intersection on gfx950.
//
// Replaces getClosestHitBVH / intersectBVH / rayTriangleIntersection
// (raytracer.glsl:102-147, 182-295).  Design:
//  - persistent waves: a wave refills its idle lanes from a pool of
//    consecutive rays taken from a work counter (one atomic per chunk, lanes
//    ranked with mbcnt over the ballot) whenever at least `refill` lanes are
//    idle, so finished lanes do not wait for the slowest ray of the wave; the
//    ray range is split into parts owned by XCDs, drained by their owner first;
//  - two-child 64-B nodes (both child boxes in one 64-B fetch), nearer child
//    first, farther child pushed; traversal stack of child refs in LDS for the
//    top kStackLds entries ([slot][lane] so lanes hit distinct banks), deeper
//    entries in a per-thread global spill column sized from the tree depth;
//  - the first kTopNodes nodes of the binary tree (BFS from the root) are
//    staged in LDS once per workgroup and read from there: every ray passes
//    through them, and an LDS read replaces an L2 round trip on the dependent
//    chain of node fetches;
//  - conservative slab test: boxes are stored widened outward by a pad of
//    2^-18 of the scene's largest coordinate (errors that scale with the scene:
//    rounding of box coordinates, of vertex translation by the origin, and the
//    leaf test's own rounding of t and of its edge functions), and the far
//    distance of every box is widened by |t_far| * 2^-19 before the compare
//    (errors that scale with the distance: the approximate reciprocal, the
//    rounding of o * inv and of the FMA, and the leaf test's translation of
//    far-away vertices); the interval is [tmin, t_best] inclusive so a later,
//    lower-id triangle at the same t is still reached.  The slack covers the
//    watertight test for any ray origin; the RefMT test's barycentrics carry
//    an error that grows with |o - p0| / cos(theta) across the ray, which no
//    box widening bounds, so RefMT rays keep this guarantee only while that
//    error stays below the pad;
//  - leaf test: watertight FP32 (Woop, Benthin & Wald, JCGT 2013) with the FP64
//    re-evaluation of edge functions that are exactly zero, or the reference's
//    Moller-Trumbore (RefMT); closest hit keeps the smallest (t, id).
// Float expressions of the leaf tests are evaluated in the order of the CPU
// restatement (built with -ffp-contract=off), so t/u/v are bit-identical.
#include <algorithm>
#include <mutex>

#include "rtq_internal.hpp"

namespace rtq {
namespace {

constexpr int kTraceThreads = 256;
// Work distribution: every ray set (item) of a launch is split into kXcds
// contiguous parts and part j of each item belongs to XCD j.  Workgroups are
// dispatched to the XCDs round-robin, so a wave of block b drains XCD b %
// kXcds's part of every item first (its rays then share nodes in that XCD's
// L2), then steals parts of the other XCDs.  Part (item i, j) has counter
// p = i * kXcds + j, 64 B apart; counter[kDry..] is a bit mask of exhausted
// parts; counter[kArrive] counts finished blocks.
constexpr uint32_t kXcds = 8;
constexpr uint32_t kMaxBatchItems = 64;
constexpr uint32_t kMaxParts = kMaxBatchItems * kXcds;
constexpr uint32_t kPartStride = 16;
constexpr uint32_t kDry = kMaxParts * kPartStride;
constexpr uint32_t kArrive = kDry + kMaxParts / 32;
constexpr uint32_t kCounterWords = kArrive + kPartStride;

// Relative widening of a box's far distance (tn <= tf + |tf| * kSlabRel):
// 32 units of 2^-24, twice the sum of the distance-proportional errors of the
// box test (reciprocal 2, o * inv 1, FMA 1, 8-wide node base 1) and of the
// watertight test (vertex translation 1, t = T / det about 6), plus the
// rounding of the widening itself.
constexpr float kSlabRel = 0x1p-19f;
// Newest stack entries held in registers (0: none) before LDS: a pop then
// hands over its node at once, and the register refill from LDS completes
// behind the next node fetch.
#ifndef RTQ_STACK_REG
#define RTQ_STACK_REG 0
#endif
constexpr int kStackReg = RTQ_STACK_REG;
// Optional VGPR budget of the trace kernels (occupancy is 512 / budget waves
// per SIMD), or an occupancy the compiler must fit (waves per SIMD).
#ifdef RTQ_TRACE_VGPRS
#define RTQ_TRACE_OCCUPANCY __attribute__((amdgpu_num_vgpr(RTQ_TRACE_VGPRS)))
#elif defined(RTQ_TRACE_WAVES)
#define RTQ_TRACE_OCCUPANCY __attribute__((amdgpu_waves_per_eu(RTQ_TRACE_WAVES)))
#else
#define RTQ_TRACE_OCCUPANCY
#endif

// One ray set of a batched launch: rays [begin, begin + n) of the launch's
// global index space trace against one scene.
struct BatchItem {
    const float4* nodes;
    const float4* tris;
    const rtq_ray* rays;
    rtq_hit* hits;
    uint32_t begin;
    uint32_t n;
    uint32_t any;
    uint32_t learned;  // take the learned child order where the nodes hold it
};
static_assert(sizeof(BatchItem) == 48, "batch descriptor is 48 B");

// Item descriptors of a batched launch, passed by value (3 KB of kernel
// arguments) so a captured graph keeps them; empty for single-scene launches.
template <bool MULTI>
struct ItemTable {
    BatchItem item[kMaxBatchItems];
};
template <>
struct ItemTable<false> {
    uint32_t unused;
};

struct TraceArgs {
    const float4* __restrict__ nodes;
    const float4* __restrict__ tris;
    const rtq_ray* __restrict__ rays;
    rtq_hit* __restrict__ hits;
    uint32_t n_items;        // batched launches: items in the ItemTable argument
    uint32_t n;
    uint32_t* counter;
    uint32_t* spill;
    uint32_t spill_stride;
    uint32_t refill;     // refill once this many lanes of a wave are idle
    uint32_t chunk;      // rays taken from a work counter per atomic
    uint32_t slack;      // leaf tests start once at most this many lanes still descend
    uint32_t leaf_slack; // node visits resume once at most this many lanes test triangles
    uint32_t* stat_nodes;
    uint32_t* stat_tris;
    const float4* top;   // top-of-tree copy (kTopBit refs) of the binary nodes, or null
    uint32_t n_top;      // its node count (0: traversal starts at global node 0)
};

// Rays are read once and hit records written once per launch: non-temporal
// (streaming) accesses keep them from displacing nodes and triangles in the
// caches.
typedef float stream4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ float4 load_stream(const float4* p) {
    const stream4 v = __builtin_nontemporal_load(reinterpret_cast<const stream4*>(p));
    return make_float4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ void store_stream(rtq_hit* p, const rtq_hit& h) {
    const stream4 v = {__int_as_float(h.id), h.t, h.u, h.v};
    __builtin_nontemporal_store(v, reinterpret_cast<stream4*>(p));
}

__device__ __forceinline__ float sel3(float4 v, int k) { return k == 0 ? v.x : (k == 1 ? v.y : v.z); }
__device__ __forceinline__ float sel3f(float x, float y, float z, int k) {
    return k == 0 ? x : (k == 1 ? y : z);
}

// Watertight test; on a hit with tmin <= t <= tmax returns true and t, V, W, det.
__device__ __forceinline__ bool wt_test(float4 p0, float4 p1, float4 p2, int kx, int ky, int kz, float opx,
                                        float opy, float opz, float sx, float sy, float sz, float tmin,
                                        float tmax, float& t, float& V_out, float& W_out, float& det_out) {
    const float Akx = sel3(p0, kx) - opx, Aky = sel3(p0, ky) - opy, Akz = sel3(p0, kz) - opz;
    const float Bkx = sel3(p1, kx) - opx, Bky = sel3(p1, ky) - opy, Bkz = sel3(p1, kz) - opz;
    const float Ckx = sel3(p2, kx) - opx, Cky = sel3(p2, ky) - opy, Ckz = sel3(p2, kz) - opz;
    const float Ax = Akx - sx * Akz;
    const float Ay = Aky - sy * Akz;
    const float Bx = Bkx - sx * Bkz;
    const float By = Bky - sy * Bkz;
    const float Cx = Ckx - sx * Ckz;
    const float Cy = Cky - sy * Ckz;
    float U = Cx * By - Cy * Bx;
    float V = Ax * Cy - Ay * Cx;
    float W = Bx * Ay - By * Ax;
    if (U == 0.0f || V == 0.0f || W == 0.0f) {
        const double CxBy = (double)Cx * (double)By, CyBx = (double)Cy * (double)Bx;
        U = (float)(CxBy - CyBx);
        const double AxCy = (double)Ax * (double)Cy, AyCx = (double)Ay * (double)Cx;
        V = (float)(AxCy - AyCx);
        const double BxAy = (double)Bx * (double)Ay, ByAx = (double)By * (double)Ax;
        W = (float)(BxAy - ByAx);
    }
    if ((U < 0.0f || V < 0.0f || W < 0.0f) && (U > 0.0f || V > 0.0f || W > 0.0f)) return false;
    const float det = (U + V) + W;
    if (det == 0.0f) return false;
    const float Az = sz * Akz;
    const float Bz = sz * Bkz;
    const float Cz = sz * Ckz;
    const float T = (U * Az + V * Bz) + W * Cz;
    const float tt = T / det;
    if (!(tt >= tmin && tt <= tmax)) return false;
    t = tt;
    V_out = V;
    W_out = W;
    det_out = det;
    return true;
}

// raytracer.glsl:102-147 (shader vertex order: its p1 is P2, its p2 is P1).
__device__ __forceinline__ bool refmt_test(float4 q0, float4 q1, float4 q2, float ox, float oy, float oz,
                                           float dx, float dy, float dz, float tmin, float tmax, float& t,
                                           float& b0_out, float& b1_out) {
    const float p0x = q0.x, p0y = q0.y, p0z = q0.z;
    const float e0x = q2.x - p0x, e0y = q2.y - p0y, e0z = q2.z - p0z;  // shader p1 - p0
    const float e1x = q1.x - p0x, e1y = q1.y - p0y, e1z = q1.z - p0z;  // shader p2 - p0
    // n = cross(e1, e0)
    const float nx = e1y * e0z - e0y * e1z;
    const float ny = e1z * e0x - e0z * e1x;
    const float nz = e1x * e0y - e0x * e1y;
    // q = cross(d, e1)
    const float qx = dy * e1z - e1y * dz;
    const float qy = dz * e1x - e1z * dx;
    const float qz = dx * e1y - e1x * dy;
    const float a = (e0x * qx + e0y * qy) + e0z * qz;
    if ((nx * dx + ny * dy) + nz * dz >= 0.0f || fabsf(a) < 1e-4f) return false;
    const float sx = (ox - p0x) / a, sy = (oy - p0y) / a, sz = (oz - p0z) / a;
    // r = cross(s, e0)
    const float rx = sy * e0z - e0y * sz;
    const float ry = sz * e0x - e0z * sx;
    const float rz = sx * e0y - e0x * sy;
    const float b0 = (sx * qx + sy * qy) + sz * qz;
    const float b1 = (rx * dx + ry * dy) + rz * dz;
    const float b2 = (1.0f - b0) - b1;
    if (b0 < 0.0f || b1 < 0.0f || b2 < 0.0f) return false;
    const float tt = (e1x * rx + e1y * ry) + e1z * rz;
    if (tt < 0.0f) return false;
    if (!(tt >= tmin && tt <= tmax)) return false;
    t = tt;
    b0_out = b0;
    b1_out = b1;
    return true;
}

__device__ __forceinline__ uint4 bits4(float4 v) {
    return make_uint4(__float_as_uint(v.x), __float_as_uint(v.y), __float_as_uint(v.z), __float_as_uint(v.w));
}

__device__ __forceinline__ void cswap(uint32_t& a, uint32_t& b) {
    const uint32_t lo = min(a, b), hi = max(a, b);
    a = lo;
    b = hi;
}

// 19-comparator sorting network for 8 keys (optimal size, depth 6).
__device__ __forceinline__ void sort8(uint32_t* k) {
    cswap(k[0], k[2]); cswap(k[1], k[3]); cswap(k[4], k[6]); cswap(k[5], k[7]);
    cswap(k[0], k[4]); cswap(k[1], k[5]); cswap(k[2], k[6]); cswap(k[3], k[7]);
    cswap(k[0], k[1]); cswap(k[2], k[3]); cswap(k[4], k[5]); cswap(k[6], k[7]);
    cswap(k[2], k[4]); cswap(k[3], k[5]);
    cswap(k[1], k[4]); cswap(k[3], k[6]);
    cswap(k[1], k[2]); cswap(k[3], k[4]); cswap(k[5], k[6]);
}

// child ref of slot s (0..7) from the two ref quads, as a select tree
__device__ __forceinline__ uint32_t slot_ref(uint4 a, uint4 b, uint32_t s) {
    const uint4 q = (s & 4u) ? b : a;
    const uint32_t lo = (s & 1u) ? q.y : q.x;
    const uint32_t hi = (s & 1u) ? q.w : q.z;
    return (s & 2u) ? hi : lo;
}

__device__ __forceinline__ bool is_leaf_ref(uint32_t r) { return (r & kLeafBit) && r != kEmpty; }
__device__ __forceinline__ bool is_node_ref(uint32_t r) { return !(r & kLeafBit); }

// Node layouts a launch traces.
enum Layout : int {
    kBinary = 0,   // 64-B two-child nodes, nearer child first
    kWide = 1,         // 128-B compressed 8-wide nodes
    kOrdered = 2,      // 64-B two-child nodes, learned child order where both are hit
    kWideOrdered = 3,  // 8-wide nodes, the learned first child where it is hit
};

__host__ __device__ constexpr bool wide_layout(int layout) { return layout == kWide || layout == kWideOrdered; }
__host__ __device__ constexpr bool learned_layout(int layout) { return layout == kOrdered || layout == kWideOrdered; }

// MODE: 0 closest hit, 1 any hit, 2 per batch item.  MULTI: 0 one ray set;
// 1 several items (ray sets over any scenes), each lane takes its scene from
// its item; 2 several items over one scene, whose nodes and triangles are
// kernel arguments (uniform registers instead of two per-lane pointers).  A
// batched lane keeps its item index and looks up the hit buffer when it
// writes the hit.
// The body of the trace kernels below (k_trace, and k_trace_occ with an
// occupancy hint).
template <int LEAF, int MODE, bool STATS, int LAYOUT, int MULTI, int STACK>
__device__ __forceinline__ void trace_waves(const TraceArgs& a, const ItemTable<(MULTI != 0)>& table) {
    constexpr uint32_t kStackLds = STACK;
    __shared__ uint32_t s_stack[kStackLds * kTraceThreads];
    __shared__ BatchItem s_items[MULTI ? kMaxBatchItems : 1];
    // the top of the binary tree, read from LDS instead of L2 by every ray
    constexpr uint32_t kTopSlots = (wide_layout(LAYOUT) || kTopNodes == 0) ? 1u : kTopNodes;
    __shared__ float4 s_top[4 * kTopSlots];
    const uint32_t tid = threadIdx.x;
    const uint32_t lane = tid & 63u;
    uint32_t* const lds = s_stack + tid;  // entry k at lds[k * kTraceThreads]: lanes on distinct banks
    uint32_t* const spill = a.spill + (size_t)blockIdx.x * kTraceThreads + tid;
    const size_t stride = a.spill_stride;
    const uint32_t n_top = (wide_layout(LAYOUT) || kTopNodes == 0) ? 0u : min(a.n_top, kTopSlots);
    const uint32_t root = n_top ? kTopBit : 0u;
    if constexpr (MULTI)
        for (uint32_t i = tid; i < a.n_items; i += kTraceThreads) s_items[i] = table.item[i];
    for (uint32_t i = tid; i < 4 * n_top; i += kTraceThreads) s_top[i] = a.top[i];
    if (MULTI || n_top) __syncthreads();
    // the lane's scene and ray set (kernel arguments unless batched)
    const float4* nodes_p = a.nodes;
    const float4* tris_p = a.tris;
    rtq_hit* hits_p = a.hits;
    uint32_t local = 0;              // ray index within its ray set (one ray set)
    uint32_t lane_item = 0;          // batch item of the lane's ray
    bool lane_any = MODE == 1;
    // learned order: bit (octant + 8 * any) of a node's order word set = child 0
    // first, bit 16 + that = decision used; order_mask clears the word for
    // lanes whose ray set does not use the learned order
    uint32_t order_bit = 0, order_mask = 0xFFFFFFFFu;

    int rid = -1;
    // wave-uniform ray pool [pool_next, pool_end) of item pool_item, taken
    // from the seq-th part of the wave's order; dry once every part is
    // exhausted
    uint32_t pool_next = 0, pool_end = 0, pool_item = 0;
    const uint32_t xcd = blockIdx.x % kXcds;
    const uint32_t n_items = MULTI ? a.n_items : 1u;
    const uint32_t n_parts = n_items * kXcds;
    uint32_t seq = 0;
    bool dry = false;
    float ox = 0, oy = 0, oz = 0, dx = 0, dy = 0, dz = 0;
    float ix = 0, iy = 0, iz = 0, oix = 0, oiy = 0, oiz = 0;
    float tmin = 0, tbest = 0;
    int kx = 0, ky = 1, kz = 2;
    float sx = 0, sy = 0, sz = 0, opx = 0, opy = 0, opz = 0;
    int best_id = -1;
    float best_t = 0, best_a = 0, best_b = 0, best_c = 0;
    // cur: next node to visit (or a leaf waiting behind the postponed one);
    // leaf: postponed leaf; kEmpty in both once the stack ran out
    uint32_t cur = kEmpty, leaf = kEmpty, sp = 0;
    uint32_t leaf_k = 0;  // next triangle of `leaf`
    uint32_t n_nodes = 0, n_tests = 0;

    // entry k of the stack's memory part: LDS, then the HBM spill column
    auto mem_store = [&](uint32_t k, uint32_t v) {
        if (k < kStackLds) {
            lds[k * kTraceThreads] = v;
        } else {
            __hip_atomic_store(spill + (size_t)(k - kStackLds) * stride, v, __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_WORKGROUP);
        }
    };
    auto mem_load = [&](uint32_t k) -> uint32_t {
        if (k < kStackLds) return lds[k * kTraceThreads];
        return __hip_atomic_load(spill + (size_t)(k - kStackLds) * stride, __ATOMIC_RELAXED,
                                 __HIP_MEMORY_SCOPE_WORKGROUP);
    };
    // sp entries in all; with a register top the newest min(sp, kStackReg)
    // are reg[0] (newest) .. reg[kStackReg - 1], the older ones memory
    // entries 0 .. sp - kStackReg - 1
    uint32_t reg[kStackReg > 0 ? kStackReg : 1] = {};
    auto push = [&](uint32_t v) {
        if constexpr (kStackReg > 0) {
            if (sp >= (uint32_t)kStackReg) mem_store(sp - kStackReg, reg[kStackReg - 1]);
#pragma unroll
            for (int i = kStackReg - 1; i > 0; --i) reg[i] = reg[i - 1];
            reg[0] = v;
        } else {
            mem_store(sp, v);
        }
        ++sp;
    };
    auto pop = [&]() -> uint32_t {
        if (sp == 0) return kEmpty;
        --sp;
        if constexpr (kStackReg > 0) {
            const uint32_t v = reg[0];
#pragma unroll
            for (int i = 0; i < kStackReg - 1; ++i) reg[i] = reg[i + 1];
            if (sp >= (uint32_t)kStackReg) reg[kStackReg - 1] = mem_load(sp - kStackReg);
            return v;
        } else {
            return mem_load(sp);
        }
    };
    auto finish = [&]() {
        rtq_hit h;
        if (best_id >= 0) {
            h.id = best_id;
            h.t = best_t;
            if (LEAF == 0) {
                h.u = best_a / best_c;
                h.v = best_b / best_c;
            } else {
                h.u = best_a;
                h.v = best_b;
            }
        } else {
            h.id = -1;
            h.t = INFINITY;
            h.u = 0.0f;
            h.v = 0.0f;
        }
        if constexpr (MULTI != 0) {
            const BatchItem& it = s_items[lane_item];
            store_stream(it.hits + ((uint32_t)rid - it.begin), h);
        } else {
            store_stream(hits_p + local, h);
        }
        if (STATS) {
            if (a.stat_nodes) a.stat_nodes[rid] = n_nodes;
            if (a.stat_tris) a.stat_tris[rid] = n_tests;
        }
        rid = -1;
    };
    // One internal-node visit: cur <- nearest hit child (farther ones pushed) or a pop.
    auto visit = [&]() {
        if constexpr (wide_layout(LAYOUT)) {
            const float4* nd = nodes_p + 8ull * cur;
            const float4 h = nd[0];
            const uint4 qx = bits4(nd[1]), qy = bits4(nd[2]), qz = bits4(nd[3]);
            const uint4 c03 = bits4(nd[4]), c47 = bits4(nd[5]);
            if (STATS) ++n_nodes;
            const uint32_t meta = __float_as_uint(h.w);
            // child plane distance = q * (2^e * inv) + (origin - o) * inv
            const float scx = __uint_as_float((meta & 0xFFu) << 23) * ix;
            const float scy = __uint_as_float(((meta >> 8) & 0xFFu) << 23) * iy;
            const float scz = __uint_as_float(((meta >> 16) & 0xFFu) << 23) * iz;
            const float bx = __builtin_fmaf(h.x, ix, -oix);
            const float by = __builtin_fmaf(h.y, iy, -oiy);
            const float bz = __builtin_fmaf(h.z, iz, -oiz);
            // learned first child for the ray's octant and mode (table entry
            // 8 | slot; lf = 8: none); dropped from the key list when hit
            uint32_t lf = 8u;
            if constexpr (LAYOUT == kWideOrdered) {
                const uint2 tab = *reinterpret_cast<const uint2*>(nd + 6);
                const uint32_t e = (((order_bit & 8u) ? tab.y : tab.x) & order_mask) >> (4u * (order_bit & 7u));
                lf = (e & 8u) ? (e & 7u) : 8u;
            }
            bool lf_hit = false;
            // near / far planes by the direction's sign: with ix >= 0 the lower
            // plane is the nearer (q_lo <= q_hi and the FMA is monotonic, so
            // this equals min / max of the two distances exactly)
            const bool px = ix >= 0.0f, py = iy >= 0.0f, pz = iz >= 0.0f;
            const uint32_t nx0 = px ? qx.x : qx.z, nx1 = px ? qx.y : qx.w, fx0 = px ? qx.z : qx.x, fx1 = px ? qx.w : qx.y;
            const uint32_t ny0 = py ? qy.x : qy.z, ny1 = py ? qy.y : qy.w, fy0 = py ? qy.z : qy.x, fy1 = py ? qy.w : qy.y;
            const uint32_t nz0 = pz ? qz.x : qz.z, nz1 = pz ? qz.y : qz.w, fz0 = pz ? qz.z : qz.x, fz1 = pz ? qz.w : qz.y;
            uint32_t key[8];
            bool hv[8];   // slot hit, the learned first child left out (kept as lane masks: no VALU to test)
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                const int w = j >> 2, sh = 8 * (j & 3);
                const float tnx = __builtin_fmaf((float)(((w ? nx1 : nx0) >> sh) & 0xFFu), scx, bx);
                const float tfx = __builtin_fmaf((float)(((w ? fx1 : fx0) >> sh) & 0xFFu), scx, bx);
                const float tny = __builtin_fmaf((float)(((w ? ny1 : ny0) >> sh) & 0xFFu), scy, by);
                const float tfy = __builtin_fmaf((float)(((w ? fy1 : fy0) >> sh) & 0xFFu), scy, by);
                const float tnz = __builtin_fmaf((float)(((w ? nz1 : nz0) >> sh) & 0xFFu), scz, bz);
                const float tfz = __builtin_fmaf((float)(((w ? fz1 : fz0) >> sh) & 0xFFu), scz, bz);
                const float tn = fmaxf(fmaxf(tnx, tny), fmaxf(tnz, tmin));
                const float tf = fminf(fminf(tfx, tfy), fminf(tfz, tbest));
                // empty slots hold inverted boxes and a harmless leaf ref (one
                // real triangle), so the valid bits need no test here
                const bool hit = tn <= __builtin_fmaf(fabsf(tf), kSlabRel, tf);
                const bool learned_first = LAYOUT == kWideOrdered && hit && lf == (uint32_t)j;
                if (LAYOUT == kWideOrdered) lf_hit = lf_hit || learned_first;
                // order key: entry distance (non-negative floats order as uints), slot in the low bits
                key[j] = (hit && !learned_first)
                             ? (((uint32_t)max((int32_t)__float_as_uint(tn), 0) & ~7u) | (uint32_t)j)
                             : 0xFFFFFFFFu;
                hv[j] = hit && !learned_first;
            }
            if (MODE == 1 || (MODE == 2 && lane_any)) {
                // any hit: the learned child first, else the lowest hit slot; the others in slot order,
                // popped outward from the first child's side (slots follow the binary subtree's
                // left-to-right frontier, so neighbouring slots are neighbouring subtrees); an
                // entry-distance order cost more VALU than it saved visits, occluded rays included
                if (!lf_hit && !(hv[0] || hv[1] || hv[2] || hv[3] || hv[4] || hv[5] || hv[6] || hv[7])) {
                    cur = pop();
                    return;
                }
                if (!lf_hit) {
                    // the highest pushed first, so pops go up; the lowest hit slot is
                    // kept back as the next node (its ref without a select tree)
                    const uint32_t refs[8] = {c03.x, c03.y, c03.z, c03.w, c47.x, c47.y, c47.z, c47.w};
                    uint32_t next = kEmpty;
                    bool have = false;   // a lane mask: no compare of `next`
#pragma unroll
                    for (int j = 7; j >= 0; --j)
                        if (hv[j]) {
                            if (have) push(next);
                            next = refs[j];
                            have = true;
                        }
                    cur = next;
                    return;
                }
                const uint32_t first = lf;
                if (lf < 4u) {
#pragma unroll
                    for (int j = 7; j >= 0; --j)   // the highest pushed first: pops go up
                        if (hv[j] && (uint32_t)j != first) push(slot_ref(c03, c47, (uint32_t)j));
                } else {
#pragma unroll
                    for (int j = 0; j < 8; ++j)    // pops go down from a learned child in the upper half
                        if (hv[j] && (uint32_t)j != first) push(slot_ref(c03, c47, (uint32_t)j));
                }
                cur = slot_ref(c03, c47, first);
            } else {
                // the learned child (not in the keys) goes first, the others nearest first
                sort8(key);
                if (!lf_hit && key[0] == 0xFFFFFFFFu) {
                    cur = pop();
                    return;
                }
                // farthest first onto the stack, so the next pop is the second nearest
#pragma unroll
                for (int j = 7; j >= 1; --j)
                    if (key[j] != 0xFFFFFFFFu) push(slot_ref(c03, c47, key[j] & 7u));
                if (lf_hit) {
                    if (key[0] != 0xFFFFFFFFu) push(slot_ref(c03, c47, key[0] & 7u));
                    cur = slot_ref(c03, c47, lf);
                } else {
                    cur = slot_ref(c03, c47, key[0] & 7u);
                }
            }
        } else {
            float4 n0, n1, n2, n3;
            if (cur & kTopBit) {
                const float4* nd = s_top + 4 * (cur & ~kTopBit);
                n0 = nd[0]; n1 = nd[1]; n2 = nd[2]; n3 = nd[3];
            } else {
                const float4* nd = nodes_p + 4ull * cur;
                n0 = nd[0]; n1 = nd[1]; n2 = nd[2]; n3 = nd[3];
            }
            if (STATS) ++n_nodes;
            const float ax0 = __builtin_fmaf(n0.x, ix, -oix), bx0 = __builtin_fmaf(n0.y, ix, -oix);
            const float ay0 = __builtin_fmaf(n0.z, iy, -oiy), by0 = __builtin_fmaf(n0.w, iy, -oiy);
            const float az0 = __builtin_fmaf(n1.x, iz, -oiz), bz0 = __builtin_fmaf(n1.y, iz, -oiz);
            const float ax1 = __builtin_fmaf(n1.z, ix, -oix), bx1 = __builtin_fmaf(n1.w, ix, -oix);
            const float ay1 = __builtin_fmaf(n2.x, iy, -oiy), by1 = __builtin_fmaf(n2.y, iy, -oiy);
            const float az1 = __builtin_fmaf(n2.z, iz, -oiz), bz1 = __builtin_fmaf(n2.w, iz, -oiz);
            const float tn0 = fmaxf(fmaxf(fminf(ax0, bx0), fminf(ay0, by0)), fmaxf(fminf(az0, bz0), tmin));
            const float tf0 = fminf(fminf(fmaxf(ax0, bx0), fmaxf(ay0, by0)), fminf(fmaxf(az0, bz0), tbest));
            const float tn1 = fmaxf(fmaxf(fminf(ax1, bx1), fminf(ay1, by1)), fmaxf(fminf(az1, bz1), tmin));
            const float tf1 = fminf(fminf(fmaxf(ax1, bx1), fmaxf(ay1, by1)), fminf(fmaxf(az1, bz1), tbest));
            const bool h0 = tn0 <= __builtin_fmaf(fabsf(tf0), kSlabRel, tf0);
            const bool h1 = tn1 <= __builtin_fmaf(fabsf(tf1), kSlabRel, tf1);
            const uint32_t r0 = __float_as_uint(n3.x), r1 = __float_as_uint(n3.y);
            if (h0 && h1) {
                bool first0;
                if constexpr (LAYOUT == kOrdered) {
                    // learned decision where its confidence bit is set, else the nearer child
                    const uint32_t ow = (__float_as_uint(n3.z) & order_mask) >> order_bit;
                    first0 = ((ow >> 16) & 1u) ? (ow & 1u) != 0 : tn0 <= tn1;
                } else {
                    first0 = tn0 <= tn1;
                }
                push(first0 ? r1 : r0);
                cur = first0 ? r0 : r1;
            } else if (h0 || h1) {
                cur = h0 ? r0 : r1;
            } else {
                cur = pop();
            }
        }
    };

    for (;;) {
        // ---- refill idle lanes from the wave's pool of consecutive rays; the
        // pool is topped up with one atomic on a part counter per chunk
        const bool need = (rid < 0) && !dry;
        const uint64_t need_mask = __ballot(need);
        if (need_mask) {
            if (pool_next == pool_end) {
                const uint32_t leader = (uint32_t)__ffsll((unsigned long long)need_mask) - 1u;
                uint32_t lo = 0, hi = 0, item = 0;
                if (lane == leader) {
                    // the current part, then the next ones of the wave's order
                    // (own XCD's part of each item, then the next XCD's);
                    // parts marked exhausted are skipped without an atomic
                    const unsigned long long* dry_mask =
                        reinterpret_cast<const unsigned long long*>(a.counter + kDry);
                    unsigned long long gone = 0;
                    uint32_t gone_word = kNone;
                    for (const uint32_t first = seq; seq < n_parts; ++seq) {
                        item = seq % n_items;
                        const uint32_t j = (xcd + seq / n_items) % kXcds;
                        const uint32_t p = item * kXcds + j;
                        // up to 16 parts are probed with their own atomics; more
                        // (large batches) consult the exhausted-part mask so a
                        // wave's end-of-launch probing stays short
                        if (n_parts > 16 && seq != first) {
                            if (p / 64 != gone_word) {
                                gone_word = p / 64;
                                gone = __hip_atomic_load(dry_mask + gone_word, __ATOMIC_RELAXED,
                                                         __HIP_MEMORY_SCOPE_AGENT);
                            }
                            if ((gone >> (p % 64)) & 1ull) continue;
                        }
                        uint32_t ib = 0, in = a.n;
                        if constexpr (MULTI) {
                            ib = s_items[item].begin;
                            in = s_items[item].n;
                        }
                        const uint32_t pb = ib + (uint32_t)(((uint64_t)in * j) / kXcds);
                        const uint32_t pe = ib + (uint32_t)(((uint64_t)in * (j + 1)) / kXcds);
                        const uint32_t base = atomicAdd(a.counter + p * kPartStride, a.chunk);
                        if (base < pe - pb) {
                            lo = pb + base;
                            hi = min(lo + a.chunk, pe);
                            break;
                        }
                        if (n_parts > 16)
                            atomicOr(const_cast<unsigned long long*>(dry_mask) + p / 64, 1ull << (p % 64));
                    }
                }
                lo = __shfl(lo, (int)leader);
                hi = __shfl(hi, (int)leader);
                item = __shfl(item, (int)leader);
                seq = __shfl(seq, (int)leader);
                if (lo == hi) {
                    dry = true;
                } else {
                    pool_next = lo;
                    pool_end = hi;
                    pool_item = item;
                }
            }
            if (!dry) {
                const uint32_t rank = __builtin_amdgcn_mbcnt_hi((uint32_t)(need_mask >> 32),
                                                                __builtin_amdgcn_mbcnt_lo((uint32_t)need_mask, 0u));
                const uint32_t avail = pool_end - pool_next;
                const uint32_t r = pool_next + rank;
                pool_next += min((uint32_t)__popcll(need_mask), avail);
                if (need && rank < avail) {
                    rid = (int)r;
                    const rtq_ray* rays_p = a.rays;
                    local = r;
                    if constexpr (MULTI != 0) {
                        const BatchItem& it = s_items[pool_item];
                        if constexpr (MULTI == 1) {
                            nodes_p = it.nodes;
                            tris_p = it.tris;
                        }
                        rays_p = it.rays;
                        lane_item = pool_item;
                        local = r - it.begin;
                        lane_any = it.any != 0;
                        order_mask = it.learned ? 0xFFFFFFFFu : 0u;
                    }
                    const float4* rp = reinterpret_cast<const float4*>(rays_p + local);
                    const float4 r0 = load_stream(rp), r1 = load_stream(rp + 1);
                    ox = r0.x; oy = r0.y; oz = r0.z; tmin = r0.w;
                    dx = r1.x; dy = r1.y; dz = r1.z; tbest = r1.w;
                    if constexpr (learned_layout(LAYOUT)) {
                        const bool any_ray = MODE == 1 || (MODE == 2 && lane_any);
                        order_bit = (dx < 0.0f ? 1u : 0u) | (dy < 0.0f ? 2u : 0u) | (dz < 0.0f ? 4u : 0u) |
                                    (any_ray ? 8u : 0u);
                    }
                    best_id = -1;
                    n_nodes = 0;
                    n_tests = 0;
                    sp = 0;
                    cur = root;
                    leaf = kEmpty;
                    leaf_k = 0;
                    // box-test reciprocals (zero components replaced by +-1e-30)
                    const float ex = fabsf(dx) < 1e-30f ? copysignf(1e-30f, dx) : dx;
                    const float ey = fabsf(dy) < 1e-30f ? copysignf(1e-30f, dy) : dy;
                    const float ez = fabsf(dz) < 1e-30f ? copysignf(1e-30f, dz) : dz;
                    ix = __builtin_amdgcn_rcpf(ex);
                    iy = __builtin_amdgcn_rcpf(ey);
                    iz = __builtin_amdgcn_rcpf(ez);
                    oix = ox * ix; oiy = oy * iy; oiz = oz * iz;
                    if (LEAF == 0) {
                        // watertight ray set-up: kz = first axis of max |d|
                        const float ax = fabsf(dx), ay = fabsf(dy), az = fabsf(dz);
                        int z = 0;
                        float m = ax;
                        if (ay > m) { z = 1; m = ay; }
                        if (az > m) { z = 2; }
                        int x = z + 1; if (x == 3) x = 0;
                        int y = x + 1; if (y == 3) y = 0;
                        const float dkz = sel3f(dx, dy, dz, z);
                        if (dkz < 0.0f) { const int tmp = x; x = y; y = tmp; }
                        kx = x; ky = y; kz = z;
                        sx = sel3f(dx, dy, dz, x) / dkz;
                        sy = sel3f(dx, dy, dz, y) / dkz;
                        sz = 1.0f / dkz;
                        opx = sel3f(ox, oy, oz, x);
                        opy = sel3f(ox, oy, oz, y);
                        opz = sel3f(ox, oy, oz, z);
                        if (dkz == 0.0f) cur = kEmpty;  // zero direction: nothing can be hit
                    }
                    if (!(tmin <= tbest)) cur = kEmpty;  // empty interval
                }
            }
        }
        // no lane holds a ray only once the counter has run dry
        if (__ballot(rid >= 0) == 0) break;

        // ---- traversal steps until enough lanes are idle
        for (;;) {
            if (rid >= 0) {
                // Internal nodes until (nearly) every lane of the wave holds a
                // leaf: a lane that finds one postpones it and keeps descending
                // (speculative traversal, Aila & Laine, HPG 2009); the wave
                // moves on to leaf tests once at most `slack` lanes still look
                // for their first leaf.
                bool searching = leaf == kEmpty;
                while (is_node_ref(cur)) {
                    visit();
                    if (is_leaf_ref(cur) && leaf == kEmpty) {
                        leaf = cur;
                        cur = pop();
                        searching = false;
                    }
                    if ((uint32_t)__popcll(__ballot(searching)) <= a.slack) break;
                }
                // postponed leaf, then any leaf that is waiting in cur: one
                // triangle per lane per iteration, so a lane whose leaf ends
                // early goes on with its next leaf instead of waiting for the
                // longest leaf of the wave; once at most `leaf_slack` lanes
                // still test triangles the wave returns to node visits and
                // those lanes resume their leaf (leaf, leaf_k) next time
                while (is_leaf_ref(leaf)) {
                    const uint32_t first = (leaf >> 3) & 0x0FFFFFFFu;
                    const uint32_t count = (leaf & 7u) + 1u;
                    const float4* tp = tris_p + 3ull * (first + leaf_k);
                    const float4 p0 = tp[0], p1 = tp[1], p2 = tp[2];
                    if (STATS) ++n_tests;
                    const int id = (int)__float_as_uint(p0.w);
                    float t, ua, ub, uc = 0.0f;
                    bool h;
                    if (LEAF == 0)
                        h = wt_test(p0, p1, p2, kx, ky, kz, opx, opy, opz, sx, sy, sz, tmin, tbest, t, ua, ub, uc);
                    else
                        h = refmt_test(p0, p1, p2, ox, oy, oz, dx, dy, dz, tmin, tbest, t, ua, ub);
                    if (h && (best_id < 0 || t < best_t || (t == best_t && id < best_id))) {
                        best_id = id;
                        best_t = t;
                        best_a = ua;
                        best_b = ub;
                        best_c = uc;
                        tbest = t;
                        if (MODE == 1 || (MODE == 2 && lane_any)) {
                            cur = kEmpty;
                            leaf = kEmpty;
                            break;
                        }
                    }
                    if (++leaf_k == count) {
                        leaf_k = 0;
                        if (is_leaf_ref(cur)) {
                            leaf = cur;
                            cur = pop();
                        } else {
                            leaf = kEmpty;
                        }
                    }
                    if ((uint32_t)__popcll(__ballot(is_leaf_ref(leaf))) <= a.leaf_slack) break;
                }
                if (cur == kEmpty && leaf == kEmpty) finish();
            }
            const uint64_t idle = __ballot(rid < 0);
            const uint64_t active = __ballot(rid >= 0);
            if (active == 0 || (!dry && (uint32_t)__popcll(idle) >= a.refill)) break;
        }
    }
    // A wave leaves the loop only after every part ran dry for it, so when
    // every block has arrived here no fetch is outstanding: the last block
    // resets the part counters, the exhausted-part mask and the arrival counter
    // for the next launch on the stream.
    __shared__ uint32_t s_last;
    __syncthreads();
    if (tid == 0) {
        __threadfence();
        s_last = atomicAdd(a.counter + kArrive, 1u) == gridDim.x - 1;
    }
    __syncthreads();
    if (s_last) {
        for (uint32_t p = tid; p < n_parts; p += kTraceThreads) atomicExch(a.counter + p * kPartStride, 0u);
        if (tid < (n_parts + 31) / 32) atomicExch(a.counter + kDry + tid, 0u);
        if (tid == 0) atomicExch(a.counter + kArrive, 0u);
    }
}

template <int LEAF, int MODE, bool STATS, int LAYOUT, int MULTI, int STACK>
__global__ __launch_bounds__(kTraceThreads) RTQ_TRACE_OCCUPANCY void k_trace(TraceArgs a,
                                                                            ItemTable<(MULTI != 0)> table) {
    trace_waves<LEAF, MODE, STATS, LAYOUT, MULTI, STACK>(a, table);
}

// The same kernel with an occupancy (waves per SIMD) the register allocation
// must fit.
template <int LEAF, int MODE, bool STATS, int LAYOUT, int MULTI, int WAVES, int STACK>
__global__ __launch_bounds__(kTraceThreads) __attribute__((amdgpu_waves_per_eu(WAVES))) void k_trace_occ(
    TraceArgs a, ItemTable<(MULTI != 0)> table) {
    trace_waves<LEAF, MODE, STATS, LAYOUT, MULTI, STACK>(a, table);
}

// Occupancy hints by layout (0 = the compiler's own choice), from timings on
// the 2M-triangle terrain: the batched binary kernels at least 7 waves per
// SIMD (no spill), the single-scene 8-wide kernels at 7 instead of 6 (74 -> 72
// VGPRs, no spill), the one-scene batched 8-wide kernel at 6 instead of 5
// (80 VGPRs, no spill: +10% on the batched form); the single-scene binary
// kernels were slower when pushed from 7 to 8 waves, and the multi-scene
// batched 8-wide kernel (89 VGPRs, per-lane scene pointers) spills at 6.
#ifndef RTQ_OCC_BATCH
#define RTQ_OCC_BATCH 7
#endif
#ifndef RTQ_OCC_WIDE
#define RTQ_OCC_WIDE 7
#endif
#ifndef RTQ_OCC_BATCH_WIDE
#define RTQ_OCC_BATCH_WIDE 0
#endif
#ifndef RTQ_OCC_BATCH_WIDE_ONE
#define RTQ_OCC_BATCH_WIDE_ONE 6
#endif
// The learned 8-wide kernels hold the first-child decode as well: 6 waves per
// SIMD keep them free of spills.
#ifndef RTQ_OCC_WIDE_ORDERED
#define RTQ_OCC_WIDE_ORDERED 6
#endif
#ifndef RTQ_OCC_BATCH_WIDE_ORDERED_ONE
#define RTQ_OCC_BATCH_WIDE_ORDERED_ONE 6
#endif
template <int LAYOUT, int MULTI>
constexpr int occupancy_hint() {
    if constexpr (LAYOUT == kWideOrdered && MULTI == 0) return RTQ_OCC_WIDE_ORDERED;
    if constexpr (LAYOUT == kWideOrdered && MULTI == 2) return RTQ_OCC_BATCH_WIDE_ORDERED_ONE;
    if constexpr (MULTI == 0) return wide_layout(LAYOUT) ? RTQ_OCC_WIDE : 0;
    if constexpr (wide_layout(LAYOUT)) return MULTI == 2 ? RTQ_OCC_BATCH_WIDE_ONE : RTQ_OCC_BATCH_WIDE;
    return RTQ_OCC_BATCH;  // binary, with or without the learned order
}

template <int LEAF, int MODE, bool STATS, int LAYOUT, int MULTI, int STACK>
auto trace_kernel() {
    constexpr int w = occupancy_hint<LAYOUT, MULTI>();
    if constexpr (w > 0)
        return &k_trace_occ<LEAF, MODE, STATS, LAYOUT, MULTI, w, STACK>;
    else
        return &k_trace<LEAF, MODE, STATS, LAYOUT, MULTI, STACK>;
}

// LDS stack entries per lane: choice 0, 1 (the default) or 2 of the layout's family.
template <int LAYOUT>
constexpr int stack_choice(int k) {
    return (int)(wide_layout(LAYOUT) ? kStackChoicesWide[k] : kStackChoicesBinary[k]);
}

// Persistent grid sized from occupancy, spill column sized from the stack
// bound `depth_entries`, one launch (or only the scratch sizing).
template <int LEAF, int MODE, bool STATS, int LAYOUT, int MULTI, int STACK = stack_choice<LAYOUT>(1)>
rtq_status launch(rtq_ctx* ctx, TraceScratch& sc, uint64_t depth_entries, const TraceArgs& base,
                  const ItemTable<(MULTI != 0)>& table, uint32_t n, bool size_only, hipStream_t st) {
    const auto kern = trace_kernel<LEAF, MODE, STATS, LAYOUT, MULTI, STACK>();
    int per_cu = 0;
    hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, kTraceThreads, 0);
    if (e != hipSuccess || per_cu <= 0) per_cu = 4;
    uint64_t blocks = std::min<uint64_t>((n + kTraceThreads - 1) / kTraceThreads,
                                         (uint64_t)per_cu * (uint64_t)std::max(ctx->n_cu, 1));
    if (blocks == 0) blocks = 1;
    constexpr uint32_t kStackLds = STACK;
    const uint64_t column = depth_entries > kStackLds ? depth_entries - kStackLds : 0;  // spill words per thread
    if (sc.fixed && column > 0) {
        // a spill buffer that cannot grow: as many workgroups as it holds
        const uint64_t fit = sc.spill_words / (column * kTraceThreads);
        if (fit == 0)
            return fail(ctx, RTQ_ECAPACITY,
                        "trace: a workgroup's traversal stacks (%llu spill entries per ray) exceed the fixed "
                        "scratch of %llu bytes", (unsigned long long)column,
                        (unsigned long long)(sc.spill_words * sizeof(uint32_t)));
        blocks = std::min(blocks, fit);
    }
    const uint64_t threads = blocks * kTraceThreads;
    const uint64_t spill_words = column ? column * threads : 1;
    if (spill_words > sc.spill_words) {
        if (sc.fixed) return fail(ctx, RTQ_EINTERNAL, "trace: fixed scratch sizing");
        hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
        RTQ_HIP(ctx, hipStreamIsCapturing(st, &cs));
        if (cs != hipStreamCaptureStatusNone)
            return fail(ctx, RTQ_EINVAL, "trace: traversal scratch must be sized before graph capture");
        if (sc.spill) {
            // earlier launches on any stream may still use it
            RTQ_HIP(ctx, hipDeviceSynchronize());
            (void)hipFree(sc.spill);
            sc.spill = nullptr;
            sc.spill_words = 0;
        }
        RTQ_HIP(ctx, dev_alloc(&sc.spill, spill_words));
        sc.spill_words = spill_words;
    }
    if (size_only) return RTQ_OK;
    TraceArgs args = base;
    args.counter = sc.counter;
    args.spill = sc.spill;
    args.spill_stride = (uint32_t)threads;
    args.refill = ctx->refill;
    args.chunk = ctx->chunk;
    args.slack = ctx->spec_slack;
    args.leaf_slack = ctx->leaf_slack;
    hipLaunchKernelGGL(kern, dim3((uint32_t)blocks), dim3(kTraceThreads), 0, st, args, table);
    RTQ_HIP(ctx, hipGetLastError());
    return RTQ_OK;
}

// Stack entries a traversal can hold: binary nodes push one entry per tree
// level below the root, a wide node at most 7 per level.  The wide bound uses
// the binary depth (never less than the wide depth): a refit keeps the binary
// topology but may collapse the wide tree differently, and graphs captured
// before it keep their spill buffers.
uint64_t stack_bound(const Scene* s, bool wide) {
    return wide ? 7ull * s->depth + 2 : (uint64_t)s->depth + 2;
}

template <int LEAF, int MODE, bool STATS>
rtq_status launch_single(rtq_ctx* ctx, const Scene* s, TraceArgs a, uint32_t n, int layout, hipStream_t st,
                         TraceScratch* fixed = nullptr) {
    const ItemTable<false> none{};
    const bool wide = wide_layout(layout);
    a.nodes = wide ? s->wide : s->nodes;
    a.top = wide ? nullptr : s->top;
    a.n_top = wide ? 0u : s->n_top;
    TraceScratch& sc = fixed ? *fixed : ctx->scratch;
    // a fixed scratch is sized per call, so it can take the current 8-wide
    // tree's own bound (at most 7 pushes per level) instead of the binary one
    const uint64_t depth = (fixed && wide) ? 7ull * s->wide_depth + 2 : stack_bound(s, wide);
    if (layout == kWide) return launch<LEAF, MODE, STATS, kWide, 0>(ctx, sc, depth, a, none, n, false, st);
    if (layout == kWideOrdered)
        return launch<LEAF, MODE, STATS, kWideOrdered, 0>(ctx, sc, depth, a, none, n, false, st);
    if (layout == kOrdered) return launch<LEAF, MODE, STATS, kOrdered, 0>(ctx, sc, depth, a, none, n, false, st);
    return launch<LEAF, MODE, STATS, kBinary, 0>(ctx, sc, depth, a, none, n, false, st);
}

// Layout traced for a scene and trace flags; kNone-like -1 with an error set
// when RTQ_TRACE_LEARNED asks for order bits the scene does not hold.
// Learned decisions for the ray's mode: the binary kernel with learned order.
bool learned_decisions(const Scene* s, uint32_t flags) {
    return (flags & RTQ_TRACE_LEARNED) && s->ordered && ((flags & RTQ_TRACE_ANY) ? s->order_any : s->order_closest);
}

// Node layout of a trace.  A learned-order ray without decisions for its mode
// takes the nearer child everywhere, which is the plain traversal, on the
// layout plain rays use; allow_wide false keeps it (and plain rays) on binary
// nodes, for batches that also hold learned-order rays.
int pick_layout(rtq_ctx* ctx, const Scene* s, uint32_t flags, bool allow_wide = true) {
    if (flags & RTQ_TRACE_LEARNED) {
        if (!s->ordered) {
            fail(ctx, RTQ_ENOTBUILT, "trace: RTQ_TRACE_LEARNED needs order bits (rtq_order_apply)");
            return -1;
        }
        if (learned_decisions(s, flags))
            return (allow_wide && s->wide != nullptr && s->wide_ordered && !(flags & RTQ_TRACE_BINARY) &&
                    ctx->prefer_wide)
                       ? kWideOrdered
                       : kOrdered;
    }
    return (allow_wide && s->wide != nullptr && !(flags & RTQ_TRACE_BINARY) && ctx->prefer_wide) ? kWide : kBinary;
}

}  // namespace

rtq_status scratch_alloc(rtq_ctx* ctx, TraceScratch* sc) {
    RTQ_HIP(ctx, dev_alloc(&sc->counter, kCounterWords));
    RTQ_HIP(ctx, hipMemset(sc->counter, 0, kCounterWords * sizeof(uint32_t)));
    return RTQ_OK;
}

void scratch_free(TraceScratch* sc) {
    if (sc->counter) (void)hipFree(sc->counter);
    if (sc->spill) (void)hipFree(sc->spill);
    *sc = TraceScratch{};
}

// ---- allocation-free trace ------------------------------------------------
namespace {

// Scratch of rtq_trace_noalloc: the work counters (kCounterWords words), then
// the traversal stacks' spill columns.  A file-scope device array: each GPU
// the library runs on holds its own copy until the process ends.  In host code
// `scratch` names a host-side stand-in, never device memory: kernels and
// copies take the per-device address from noalloc_setup instead.
constexpr size_t kNoAllocWords = RTQ_NOALLOC_SCRATCH_BYTES / sizeof(uint32_t);
constexpr size_t kNoAllocSpillOffset = (kCounterWords + 63) & ~size_t(63);
static __device__ __align__(256) uint32_t scratch[kNoAllocWords];

constexpr int kMaxDevices = 64;
std::mutex g_noalloc_mutex;
uint32_t* g_noalloc_ptr[kMaxDevices] = {};

__global__ void k_zero_u32(uint32_t* __restrict__ p, uint32_t n) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = 0u;
}

}  // namespace

rtq_status noalloc_setup(rtq_ctx* ctx) {
    if (ctx->device < 0 || ctx->device >= kMaxDevices)
        return fail(ctx, RTQ_EINVAL, "device %d: at most %d devices", ctx->device, kMaxDevices);
    std::lock_guard<std::mutex> lock(g_noalloc_mutex);
    if (g_noalloc_ptr[ctx->device]) return RTQ_OK;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || p == nullptr) return hip_fail(ctx, e, "hipGetSymbolAddress(scratch)");
    g_noalloc_ptr[ctx->device] = static_cast<uint32_t*>(p);
    return RTQ_OK;
}

rtq_status trace_noalloc(rtq_ctx* ctx, const Scene* s, const rtq_ray* rays, uint64_t n, uint32_t flags,
                         rtq_hit* hits, hipStream_t st) {
    if (n == 0) return RTQ_OK;
    if (n > 0x80000000ull) return fail(ctx, RTQ_ECAPACITY, "trace: n_rays must be <= 2^31");
    uint32_t* base = nullptr;
    {
        std::lock_guard<std::mutex> lock(g_noalloc_mutex);
        base = (ctx->device >= 0 && ctx->device < kMaxDevices) ? g_noalloc_ptr[ctx->device] : nullptr;
    }
    if (!base) return fail(ctx, RTQ_EINTERNAL, "trace_noalloc: scratch address not set up for device %d", ctx->device);
    const int layout = pick_layout(ctx, s, flags);
    if (layout < 0) return RTQ_ENOTBUILT;
    TraceScratch sc;
    sc.counter = base;
    sc.spill = base + kNoAllocSpillOffset;
    sc.spill_words = kNoAllocWords - kNoAllocSpillOffset;
    sc.fixed = true;
    // the counters start at zero (a launch's last block also resets them)
    hipLaunchKernelGGL(k_zero_u32, dim3((kCounterWords + 255) / 256), dim3(256), 0, st, sc.counter,
                       (uint32_t)kCounterWords);
    RTQ_HIP(ctx, hipGetLastError());
    TraceArgs args{};
    args.tris = s->tris;
    args.rays = rays;
    args.hits = hits;
    args.n = (uint32_t)n;
    const bool any = flags & RTQ_TRACE_ANY;
    const uint32_t nn = (uint32_t)n;
    if (!(flags & RTQ_TRACE_LEAF_REFMT))
        return any ? launch_single<0, 1, false>(ctx, s, args, nn, layout, st, &sc)
                   : launch_single<0, 0, false>(ctx, s, args, nn, layout, st, &sc);
    return any ? launch_single<1, 1, false>(ctx, s, args, nn, layout, st, &sc)
               : launch_single<1, 0, false>(ctx, s, args, nn, layout, st, &sc);
}

rtq_status trace(rtq_ctx* ctx, const Scene* s, const rtq_ray* rays, uint64_t n, uint32_t flags, rtq_hit* hits,
                 uint32_t* stat_nodes, uint32_t* stat_tris, bool stats, hipStream_t st) {
    if (n == 0) return RTQ_OK;
    // the counter runs past n by at most one chunk per wave: n <= 2^31 keeps it in range
    if (n > 0x80000000ull) return fail(ctx, RTQ_ECAPACITY, "trace: n_rays must be <= 2^31");
    TraceArgs args{};
    args.tris = s->tris;
    args.rays = rays;
    args.hits = hits;
    args.n = (uint32_t)n;
    args.stat_nodes = stat_nodes;
    args.stat_tris = stat_tris;
    const bool any = flags & RTQ_TRACE_ANY;
    const bool refmt = flags & RTQ_TRACE_LEAF_REFMT;
    const int layout = pick_layout(ctx, s, flags);
    if (layout < 0) return RTQ_ENOTBUILT;
    const uint32_t nn = (uint32_t)n;
    if (!stats) {
        if (!refmt)
            return any ? launch_single<0, 1, false>(ctx, s, args, nn, layout, st)
                       : launch_single<0, 0, false>(ctx, s, args, nn, layout, st);
        return any ? launch_single<1, 1, false>(ctx, s, args, nn, layout, st)
                   : launch_single<1, 0, false>(ctx, s, args, nn, layout, st);
    }
    if (!refmt)
        return any ? launch_single<0, 1, true>(ctx, s, args, nn, layout, st)
                   : launch_single<0, 0, true>(ctx, s, args, nn, layout, st);
    return any ? launch_single<1, 1, true>(ctx, s, args, nn, layout, st)
               : launch_single<1, 0, true>(ctx, s, args, nn, layout, st);
}

// A batched launch at the context's LDS stack depth (stack_lds 0 or a depth
// of the other layout family: the default).
template <int LEAF, int LAYOUT, int MULTI>
rtq_status launch_batch_stack(rtq_ctx* ctx, TraceScratch& sc, uint64_t depth, const TraceArgs& args,
                              const ItemTable<true>& table, uint32_t nt, bool size_only, hipStream_t st) {
    const int want = (int)ctx->stack_lds;
    if (want == stack_choice<LAYOUT>(0))
        return launch<LEAF, 2, false, LAYOUT, MULTI, stack_choice<LAYOUT>(0)>(ctx, sc, depth, args, table, nt,
                                                                             size_only, st);
    if (want == stack_choice<LAYOUT>(2))
        return launch<LEAF, 2, false, LAYOUT, MULTI, stack_choice<LAYOUT>(2)>(ctx, sc, depth, args, table, nt,
                                                                             size_only, st);
    return launch<LEAF, 2, false, LAYOUT, MULTI>(ctx, sc, depth, args, table, nt, size_only, st);
}

template <int MULTI>
rtq_status launch_batch(rtq_ctx* ctx, TraceScratch& sc, uint64_t depth, const TraceArgs& args,
                        const ItemTable<true>& table, uint32_t nt, bool size_only, hipStream_t st, bool refmt,
                        int layout) {
    if (refmt) {
        if (layout == kWide) return launch_batch_stack<1, kWide, MULTI>(ctx, sc, depth, args, table, nt, size_only, st);
        if (layout == kWideOrdered)
            return launch_batch_stack<1, kWideOrdered, MULTI>(ctx, sc, depth, args, table, nt, size_only, st);
        if (layout == kOrdered)
            return launch_batch_stack<1, kOrdered, MULTI>(ctx, sc, depth, args, table, nt, size_only, st);
        return launch_batch_stack<1, kBinary, MULTI>(ctx, sc, depth, args, table, nt, size_only, st);
    }
    if (layout == kWide) return launch_batch_stack<0, kWide, MULTI>(ctx, sc, depth, args, table, nt, size_only, st);
    if (layout == kWideOrdered)
        return launch_batch_stack<0, kWideOrdered, MULTI>(ctx, sc, depth, args, table, nt, size_only, st);
    if (layout == kOrdered)
        return launch_batch_stack<0, kOrdered, MULTI>(ctx, sc, depth, args, table, nt, size_only, st);
    return launch_batch_stack<0, kBinary, MULTI>(ctx, sc, depth, args, table, nt, size_only, st);
}

// Checks a batch's items and picks the node layout they trace together.
rtq_status batch_layout(rtq_ctx* ctx, const rtq_batch_item* items, uint32_t n_items, int* out) {
    *out = -1;
    if (n_items == 0) return RTQ_OK;
    if (!items) return fail(ctx, RTQ_EINVAL, "trace_batch: null item array");
    if (n_items > kMaxBatchItems) return fail(ctx, RTQ_ECAPACITY, "trace_batch: at most %u items per launch", kMaxBatchItems);
    const uint32_t leaf_flags = items[0].flags & RTQ_TRACE_LEAF_REFMT;
    int layout = -1;
    // Items with learned decisions trace a learned layout, and so does the
    // rest of the batch (a per-item order mask turns the decisions off for
    // the others): 8-wide nodes when every item can trace them (8-wide nodes
    // in its scene, their first-child tables for learned items, no
    // RTQ_TRACE_BINARY, 8-wide nodes preferred), binary nodes otherwise.
    bool any_learned = false, all_wide = true;
    for (uint32_t i = 0; i < n_items; ++i) {
        const Scene* s = find_scene(ctx, items[i].scene_id);
        if (!s || !s->built) continue;   // reported below
        const bool learned = learned_decisions(s, items[i].flags);
        any_learned = any_learned || learned;
        all_wide = all_wide && s->wide != nullptr && !(items[i].flags & RTQ_TRACE_BINARY) && ctx->prefer_wide &&
                   (!learned || s->wide_ordered);
    }
    for (uint32_t i = 0; i < n_items; ++i) {
        const rtq_batch_item& it = items[i];
        const Scene* s = find_scene(ctx, it.scene_id);
        if (!s) return fail(ctx, RTQ_EINVAL, "trace_batch: item %u: unknown scene %u", i, it.scene_id);
        if (!s->built) return fail(ctx, RTQ_ENOTBUILT, "trace_batch: item %u: scene %u has no BVH", i, it.scene_id);
        if (it.flags & ~(RTQ_TRACE_ANY | RTQ_TRACE_LEAF_REFMT | RTQ_TRACE_BINARY | RTQ_TRACE_LEARNED))
            return fail(ctx, RTQ_EINVAL, "trace_batch: item %u: unknown flags 0x%x", i, it.flags);
        if ((it.flags & RTQ_TRACE_LEAF_REFMT) != leaf_flags)
            return fail(ctx, RTQ_EINVAL, "trace_batch: all items must use the same leaf test");
        if (it.n_rays && (!it.rays_dev || !it.hits_dev))
            return fail(ctx, RTQ_EINVAL, "trace_batch: item %u: null buffer", i);
        int lay = pick_layout(ctx, s, it.flags);   // refuses RTQ_TRACE_LEARNED without order bits
        if (lay < 0) return RTQ_ENOTBUILT;
        if (any_learned) lay = all_wide ? kWideOrdered : kOrdered;
        if (layout < 0) layout = lay;
        if (layout != lay) return fail(ctx, RTQ_EINVAL, "trace_batch: all items must trace the same node layout");
    }
    *out = layout;
    return RTQ_OK;
}

// Up to kMaxBatchItems ray sets (any scenes, same leaf test and layout) in one
// persistent launch: waves take chunks of the concatenated ray index space
// (their XCD's part first, then any other part), so uneven sets balance across
// the whole chip.
rtq_status trace_batch(rtq_ctx* ctx, const rtq_batch_item* items, uint32_t n_items, TraceScratch* scratch,
                       bool size_only, hipStream_t st) {
    int layout = -1;
    const rtq_status rs = batch_layout(ctx, items, n_items, &layout);
    if (rs != RTQ_OK || n_items == 0) return rs;
    ItemTable<true> table{};
    uint64_t total = 0, depth = 0;
    const uint32_t leaf_flags = items[0].flags & RTQ_TRACE_LEAF_REFMT;
    const bool wide = wide_layout(layout);
    for (uint32_t i = 0; i < n_items; ++i) {
        const rtq_batch_item& it = items[i];
        const Scene* s = find_scene(ctx, it.scene_id);
        BatchItem& b = table.item[i];
        b.nodes = wide ? s->wide : s->nodes;
        b.tris = s->tris;
        b.rays = it.rays_dev;
        b.hits = it.hits_dev;
        b.begin = (uint32_t)total;
        b.n = (uint32_t)it.n_rays;
        b.any = (it.flags & RTQ_TRACE_ANY) ? 1u : 0u;
        b.learned = learned_decisions(s, it.flags) ? 1u : 0u;
        total += it.n_rays;
        if (total > 0x80000000ull) return fail(ctx, RTQ_ECAPACITY, "trace_batch: more than 2^31 rays in one launch");
        depth = std::max(depth, stack_bound(s, wide));
    }
    if (total == 0) return RTQ_OK;
    TraceArgs args{};
    args.n = (uint32_t)total;
    args.n_items = n_items;
    // one scene for every item: its nodes and triangles become kernel
    // arguments (MULTI 2), and binary layouts stage its top-of-tree copy
    const Scene* s0 = find_scene(ctx, items[0].scene_id);
    bool one_scene = true;
    for (uint32_t i = 1; i < n_items && one_scene; ++i)
        one_scene = table.item[i].nodes == table.item[0].nodes && table.item[i].tris == table.item[0].tris;
    if (one_scene) {
        args.nodes = table.item[0].nodes;
        args.tris = table.item[0].tris;
        if (!wide_layout(layout)) {
            args.top = s0->top;
            args.n_top = s0->n_top;
        }
    }
    TraceScratch& sc = scratch ? *scratch : ctx->scratch;
    const uint32_t nt = (uint32_t)total;
    if (one_scene) return launch_batch<2>(ctx, sc, depth, args, table, nt, size_only, st, leaf_flags != 0, layout);
    return launch_batch<1>(ctx, sc, depth, args, table, nt, size_only, st, leaf_flags != 0, layout);
}

}  // namespace rtq
