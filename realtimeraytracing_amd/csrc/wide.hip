// Compressed 8-wide traversal nodes built from the binary PLOC tree.
//
// Collapse: top-down, one level per launch.  A wide node starts from the two
// children of its binary node and repeatedly replaces the child with the
// largest surface area that is still "expandable" (internal with more than
// max_leaf triangles) by that child's two children, until it has 8 children
// or none is expandable (the greedy collapse of Wald et al., "Getting Rid of
// Packets", 2008).  Expandable children become wide nodes of the next level;
// their indices come from a scan, so numbering is deterministic and the
// children of a node are contiguous.  Other children are leaves: the
// subtree's triangles are already contiguous in leaf order.
//
// Node (128 B, 8 float4; only the first 96 B are read by the traversal):
//   [ox oy oz meta] meta = ex | ey << 8 | ez << 16 | valid << 24
//   [qlo_x[0..7] qhi_x[0..7]] [qlo_y.. qhi_y..] [qlo_z.. qhi_z..]   (bytes)
//   [ref0..ref3] [ref4..ref7] [0 0 0 0] [0 0 0 0]
// A child box is o + q * 2^(e - 127) per axis.  Boxes are widened by the scene
// pad and quantised outward in exact (double) arithmetic, so the decoded box
// always contains the padded child box: no hit is lost to quantisation
// (Ylitie, Karras & Laine, "Efficient Incoherent Ray Traversal on GPUs Through
// Compressed Wide BVHs", HPG 2017, for the quantised-box idea).
#include <cmath>

#include "rtq_internal.hpp"

namespace rtq {
namespace {

constexpr int kThreads = 256;
constexpr int kWidth = 8;

__device__ __forceinline__ bool expandable(uint32_t c, uint32_t max_leaf, const uint32_t* cleft,
                                           const uint32_t* ccount) {
    return cleft[c] != kNone && ccount[c] > max_leaf;
}

__device__ __forceinline__ float box_area(const rtq_aabb& b) {
    const float dx = b.max[0] - b.min[0], dy = b.max[1] - b.min[1], dz = b.max[2] - b.min[2];
    return 2.0f * ((dx * dy + dy * dz) + dz * dx);
}

__global__ __launch_bounds__(kThreads) void k_wide_collapse(uint32_t n_items, const uint32_t* __restrict__ items,
                                                            uint32_t max_leaf, const uint32_t* __restrict__ cleft,
                                                            const uint32_t* __restrict__ cright,
                                                            const uint32_t* __restrict__ ccount,
                                                            const rtq_aabb* __restrict__ cbox,
                                                            uint32_t* __restrict__ children,
                                                            uint32_t* __restrict__ n_children,
                                                            uint32_t* __restrict__ n_expand) {
    const uint32_t k = blockIdx.x * kThreads + threadIdx.x;
    if (k >= n_items) return;
    const uint32_t c = items[k];
    uint32_t ch[kWidth];
    uint32_t n = 0;
    if (cleft[c] == kNone) {
        ch[n++] = c;  // single-triangle scene: the root itself is the only (leaf) child
    } else {
        ch[n++] = cleft[c];
        ch[n++] = cright[c];
        while (n < kWidth) {
            int best = -1;
            float best_a = -1.0f;
            for (uint32_t i = 0; i < n; ++i) {
                if (!expandable(ch[i], max_leaf, cleft, ccount)) continue;
                const float a = box_area(cbox[ch[i]]);
                if (a > best_a) { best_a = a; best = (int)i; }
            }
            if (best < 0) break;
            const uint32_t e = ch[best];
            ch[best] = cleft[e];
            ch[n++] = cright[e];
        }
        // slots left over once every child is leaf-sized: split the largest
        // leaf-sized clusters into smaller leaves (a visit tests all 8 slots
        // anyway, and smaller leaf boxes cull more triangle tests)
        while (n < kWidth) {
            int best = -1;
            float best_a = -1.0f;
            for (uint32_t i = 0; i < n; ++i) {
                if (cleft[ch[i]] == kNone) continue;
                const float a = box_area(cbox[ch[i]]);
                if (a > best_a) { best_a = a; best = (int)i; }
            }
            if (best < 0) break;
            const uint32_t e = ch[best];
            ch[best] = cleft[e];
            ch[n++] = cright[e];
        }
    }
    uint32_t ne = 0;
    for (uint32_t i = 0; i < kWidth; ++i) {
        children[(size_t)k * kWidth + i] = i < n ? ch[i] : kNone;
        if (i < n && expandable(ch[i], max_leaf, cleft, ccount)) ++ne;
    }
    n_children[k] = n;
    n_expand[k] = ne;
}

__device__ int exponent_for(double extent) {
    // smallest e with 255 * 2^e >= extent, within the normal float range
    int e = extent > 0.0 ? ilogb(extent) - 7 : -126;
    if (e < -126) e = -126;
    while (e < 127 && ldexp(255.0, e) < extent) ++e;
    while (e > -126 && ldexp(255.0, e - 1) >= extent) --e;
    return e;
}

__global__ __launch_bounds__(kThreads) void k_wide_write(
    uint32_t n_items, uint32_t level_base, uint32_t next_base, uint32_t max_leaf, float pad,
    const uint32_t* __restrict__ children, const uint32_t* __restrict__ n_children,
    const uint32_t* __restrict__ child_offs, const uint32_t* __restrict__ cleft, const uint32_t* __restrict__ ccount,
    const rtq_aabb* __restrict__ cbox, const uint32_t* __restrict__ fl, float4* __restrict__ nodes,
    uint32_t* __restrict__ next_items) {
    const uint32_t k = blockIdx.x * kThreads + threadIdx.x;
    if (k >= n_items) return;
    const uint32_t n = n_children[k];
    float lo[kWidth][3], hi[kWidth][3];
    float org[3] = {INFINITY, INFINITY, INFINITY};
    double top[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (uint32_t i = 0; i < n; ++i) {
        const rtq_aabb b = cbox[children[(size_t)k * kWidth + i]];
        for (int a = 0; a < 3; ++a) {
            lo[i][a] = b.min[a] - pad;
            hi[i][a] = b.max[a] + pad;
            org[a] = fminf(org[a], lo[i][a]);
            top[a] = fmax(top[a], (double)hi[i][a]);
        }
    }
    int ex[3];
    double scale[3];
    for (int a = 0; a < 3; ++a) {
        ex[a] = exponent_for(top[a] - (double)org[a]);
        scale[a] = ldexp(1.0, ex[a]);
    }
    uint32_t qlo[3][2] = {{0, 0}, {0, 0}, {0, 0}}, qhi[3][2] = {{0, 0}, {0, 0}, {0, 0}};
    uint32_t ref[kWidth];
    uint32_t valid = 0, rank = 0;
    for (uint32_t i = 0; i < kWidth; ++i) {
        ref[i] = 0u;
        if (i >= n) continue;
        valid |= 1u << i;
        for (int a = 0; a < 3; ++a) {
            // exact: float differences and division by a power of two are exact in double
            double l = floor(((double)lo[i][a] - (double)org[a]) / scale[a]);
            double h = ceil(((double)hi[i][a] - (double)org[a]) / scale[a]);
            l = fmin(fmax(l, 0.0), 255.0);
            h = fmin(fmax(h, 0.0), 255.0);
            qlo[a][i >> 2] |= (uint32_t)l << (8 * (i & 3));
            qhi[a][i >> 2] |= (uint32_t)h << (8 * (i & 3));
        }
        const uint32_t c = children[(size_t)k * kWidth + i];
        if (expandable(c, max_leaf, cleft, ccount)) {
            const uint32_t slot = child_offs[k] + rank++;
            ref[i] = next_base + slot;
            next_items[slot] = c;
        } else {
            ref[i] = leaf_ref(fl[c], ccount[c]);
        }
    }
    const uint32_t meta = (uint32_t)(ex[0] + 127) | ((uint32_t)(ex[1] + 127) << 8) |
                          ((uint32_t)(ex[2] + 127) << 16) | (valid << 24);
    float4* o = nodes + 8ull * (level_base + k);
    o[0] = make_float4(org[0], org[1], org[2], __uint_as_float(meta));
    o[1] = make_float4(__uint_as_float(qlo[0][0]), __uint_as_float(qlo[0][1]), __uint_as_float(qhi[0][0]),
                       __uint_as_float(qhi[0][1]));
    o[2] = make_float4(__uint_as_float(qlo[1][0]), __uint_as_float(qlo[1][1]), __uint_as_float(qhi[1][0]),
                       __uint_as_float(qhi[1][1]));
    o[3] = make_float4(__uint_as_float(qlo[2][0]), __uint_as_float(qlo[2][1]), __uint_as_float(qhi[2][0]),
                       __uint_as_float(qhi[2][1]));
    o[4] = make_float4(__uint_as_float(ref[0]), __uint_as_float(ref[1]), __uint_as_float(ref[2]),
                       __uint_as_float(ref[3]));
    o[5] = make_float4(__uint_as_float(ref[4]), __uint_as_float(ref[5]), __uint_as_float(ref[6]),
                       __uint_as_float(ref[7]));
    o[6] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    o[7] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
}

__global__ void k_set_u32(uint32_t* p, uint32_t v) { *p = v; }

inline uint32_t blocks_for(uint64_t n) { return (uint32_t)((n + kThreads - 1) / kThreads); }

}  // namespace

// Builds s->wide (s->n_wide nodes, s->wide_depth levels) from the cluster
// arrays; fl = first leaf rank of every cluster (device).
rtq_status build_wide(rtq_ctx* ctx, Scene* s, const uint32_t* fl, hipStream_t st) {
    const uint32_t nc = s->n_clusters;
    const uint32_t root = nc - 1;
    const uint32_t cap = std::max<uint32_t>(s->n_trav, 1);  // wide nodes are a subset of the binary ones
    uint32_t *items = nullptr, *next = nullptr, *children = nullptr, *nchild = nullptr, *nexp = nullptr,
             *offs = nullptr, *total = nullptr;
    Temps tmp(ctx);
    RTQ_HIP(ctx, tmp.get(&items, cap));
    RTQ_HIP(ctx, tmp.get(&next, cap));
    RTQ_HIP(ctx, tmp.get(&children, (size_t)cap * kWidth));
    RTQ_HIP(ctx, tmp.get(&nchild, cap));
    RTQ_HIP(ctx, tmp.get(&nexp, cap));
    RTQ_HIP(ctx, tmp.get(&offs, cap));
    RTQ_HIP(ctx, tmp.get(&total, 1));
    // a refit writes the same buffer again (its size depends on the topology only)
    if (!s->wide) RTQ_HIP(ctx, dev_alloc(&s->wide, 8ull * cap));
    hipLaunchKernelGGL(k_set_u32, dim3(1), dim3(1), 0, st, items, root);
    RTQ_HIP(ctx, hipGetLastError());
    uint32_t level_items = 1, level_base = 0, levels = 0;
    while (level_items > 0) {
        if (level_base + level_items > cap) return fail(ctx, RTQ_EINTERNAL, "wide collapse overflow");
        hipLaunchKernelGGL(k_wide_collapse, dim3(blocks_for(level_items)), dim3(kThreads), 0, st, level_items, items,
                           s->params.max_leaf_size, s->cleft, s->cright, s->ccount, s->cbox, children, nchild, nexp);
        RTQ_HIP(ctx, hipGetLastError());
        rtq_status rs = scan_u32(ctx, nexp, offs, level_items, total, st);
        if (rs != RTQ_OK) return rs;
        uint32_t h_total = 0;
        RTQ_HIP(ctx, hipMemcpyAsync(&h_total, total, sizeof h_total, hipMemcpyDeviceToHost, st));
        RTQ_HIP(ctx, hipStreamSynchronize(st));
        const uint32_t next_base = level_base + level_items;
        hipLaunchKernelGGL(k_wide_write, dim3(blocks_for(level_items)), dim3(kThreads), 0, st, level_items,
                           level_base, next_base, s->params.max_leaf_size, s->box_pad, children, nchild, offs,
                           s->cleft, s->ccount, s->cbox, fl, s->wide, next);
        RTQ_HIP(ctx, hipGetLastError());
        std::swap(items, next);
        level_base = next_base;
        level_items = h_total;
        ++levels;
    }
    RTQ_HIP(ctx, hipStreamSynchronize(st));
    s->n_wide = level_base;
    s->wide_depth = levels;
    return RTQ_OK;
}

}  // namespace rtq
