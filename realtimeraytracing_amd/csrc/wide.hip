// Compressed 8-wide traversal nodes built from the binary PLOC tree.
//
// Collapse: top-down, one level per launch.  Each internal cluster's 8 slots
// are given out as an SAH cost table says (see k_dp_pass below): a child that
// is internal with more than max_leaf triangles becomes a wide node of the
// next level, any other child a leaf (the subtree's triangles are already
// contiguous in leaf order).  Next-level node indices come from a scan, so
// numbering is deterministic and the children of a node are contiguous.
//
// Node (128 B, 8 float4; the plain traversal reads the first 96 B):
//   [ox oy oz meta] meta = ex | ey << 8 | ez << 16 | valid << 24
//   [qlo_x[0..7] qhi_x[0..7]] [qlo_y.. qhi_y..] [qlo_z.. qhi_z..]   (bytes)
//   [ref0..ref3] [ref4..ref7]
//   [learned first child: 16 x 4 bits, lo hi 0 0]  (zero until an order model
//       is applied; entry octant + 8 * any_hit = 8 | slot when decided)
//   [0 0 0 0]
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

// SAH-optimal collapse (Ylitie, Karras & Laine, "Efficient incoherent ray
// traversal on GPUs through compressed wide BVHs", HPG 2017), restricted to
// this layout: cost[x][j-1] is the least SAH cost of cluster x's subtree when
// it may fill j slots of its parent: a cluster of more than max_leaf triangles
// given one slot is an 8-wide node (node cost + the best split of 8 slots
// between its children), a smaller one a leaf (triangle cost x count); j > 1
// slots go either to the cluster with fewer slots or split between its
// children.  choice[x] keeps, per j = 2..8, 0 (fewer slots) or the left
// child's share (bits 3(j-2)), and the node's own split (bits 21..23).
// Rows are computed bottom-up, one PLOC pass at a time (a merge's children
// come from earlier passes); single triangles have a constant row.
constexpr float kNodeCost = 1.0f;
constexpr float kTriCost = 0.6f;  // best of 0.15-4 on the 2M-triangle terrain

__device__ __forceinline__ void dp_row(uint32_t x, uint32_t nt, const float* __restrict__ cost,
                                       const rtq_aabb* __restrict__ cbox, float* row) {
    if (x < nt) {
        const float a = kTriCost * box_area(cbox[x]);
#pragma unroll
        for (int j = 0; j < kWidth; ++j) row[j] = a;
    } else {
#pragma unroll
        for (int j = 0; j < kWidth; ++j) row[j] = cost[(size_t)(x - nt) * kWidth + j];
    }
}

__device__ __forceinline__ void dp_cluster(uint32_t c, uint32_t nt, uint32_t max_leaf,
                                           const uint32_t* __restrict__ cleft, const uint32_t* __restrict__ cright,
                                           const uint32_t* __restrict__ ccount, const rtq_aabb* __restrict__ cbox,
                                           float* __restrict__ cost, uint32_t* __restrict__ choice) {
    float dl[kWidth], dr[kWidth], row[kWidth];
    dp_row(cleft[c], nt, cost, cbox, dl);
    dp_row(cright[c], nt, cost, cbox, dr);
    float bn = INFINITY;
    uint32_t kn = 1;
    for (uint32_t k = 1; k < kWidth; ++k) {
        const float v = dl[k - 1] + dr[kWidth - k - 1];
        if (v < bn) { bn = v; kn = k; }
    }
    const float a = box_area(cbox[c]);
    const uint32_t cnt = ccount[c];
    row[0] = cnt <= max_leaf ? kTriCost * a * (float)cnt : kNodeCost * a + bn;
    uint32_t ch = (kn - 1) << 21;
    for (uint32_t j = 2; j <= kWidth; ++j) {
        float best = row[j - 2];
        uint32_t pick = 0;
        for (uint32_t k = 1; k < j; ++k) {
            const float v = dl[k - 1] + dr[j - k - 1];
            if (v < best) { best = v; pick = k; }
        }
        row[j - 1] = best;
        ch |= pick << (3 * (j - 2));
    }
    float* out = cost + (size_t)(c - nt) * kWidth;
#pragma unroll
    for (int j = 0; j < kWidth; ++j) out[j] = row[j];
    choice[c - nt] = ch;
}

__global__ __launch_bounds__(kThreads) void k_dp_pass(uint32_t first, uint32_t count, uint32_t nt, uint32_t max_leaf,
                                                      const uint32_t* __restrict__ cleft,
                                                      const uint32_t* __restrict__ cright,
                                                      const uint32_t* __restrict__ ccount,
                                                      const rtq_aabb* __restrict__ cbox, float* __restrict__ cost,
                                                      uint32_t* __restrict__ choice) {
    const uint32_t i = blockIdx.x * kThreads + threadIdx.x;
    if (i < count) dp_cluster(first + i, nt, max_leaf, cleft, cright, ccount, cbox, cost, choice);
}

// Passes [k0, k1) in one workgroup, a barrier between passes.
constexpr int kDpSmallThreads = 1024;
__global__ __launch_bounds__(kDpSmallThreads) void k_dp_small(const uint2* __restrict__ passes, uint32_t k0,
                                                              uint32_t k1, uint32_t nt, uint32_t max_leaf,
                                                              const uint32_t* __restrict__ cleft,
                                                              const uint32_t* __restrict__ cright,
                                                              const uint32_t* __restrict__ ccount,
                                                              const rtq_aabb* __restrict__ cbox,
                                                              float* __restrict__ cost, uint32_t* __restrict__ choice) {
    for (uint32_t k = k0; k < k1; ++k) {
        const uint2 p = passes[k];
        for (uint32_t i = threadIdx.x; i < p.y; i += kDpSmallThreads)
            dp_cluster(p.x + i, nt, max_leaf, cleft, cright, ccount, cbox, cost, choice);
        __syncthreads();
    }
}

// A level's node count and first node index are device words (lvl[0], lvl[1])
// so several levels queue without a host read; the grid covers an upper bound
// `ub` of the count.
__global__ __launch_bounds__(kThreads) void k_wide_collapse(uint32_t ub, const uint32_t* __restrict__ lvl,
                                                            const uint32_t* __restrict__ items,
                                                            uint32_t max_leaf, const uint32_t* __restrict__ cleft,
                                                            const uint32_t* __restrict__ cright,
                                                            const uint32_t* __restrict__ ccount,
                                                            const rtq_aabb* __restrict__ cbox,
                                                            uint32_t nt, const uint32_t* __restrict__ choice,
                                                            uint32_t* __restrict__ children,
                                                            uint32_t* __restrict__ n_children,
                                                            uint32_t* __restrict__ n_expand) {
    const uint32_t k = blockIdx.x * kThreads + threadIdx.x;
    if (k >= ub) return;
    if (k >= lvl[0]) {
        n_expand[k] = 0;   // the scan covers the bound
        return;
    }
    const uint32_t c = items[k];
    uint32_t ch[kWidth];
    uint32_t n = 0;
    if (cleft[c] == kNone) {
        ch[n++] = c;  // single-triangle scene: the root itself is the only (leaf) child
    } else {
        // the slots as the cost table splits them, children left to right
        const uint32_t kn = ((choice[c - nt] >> 21) & 7u) + 1u;
        uint32_t sx[kWidth], sj[kWidth];
        int sp = 0;
        sx[sp] = cright[c]; sj[sp++] = kWidth - kn;
        sx[sp] = cleft[c]; sj[sp++] = kn;
        while (sp > 0) {
            --sp;
            const uint32_t x = sx[sp], j = sj[sp];
            if (j == 1 || cleft[x] == kNone) {
                ch[n++] = x;
                continue;
            }
            const uint32_t pick = (choice[x - nt] >> (3 * (j - 2))) & 7u;
            if (pick == 0) {
                sx[sp] = x; sj[sp++] = j - 1;
            } else {
                sx[sp] = cright[x]; sj[sp++] = j - pick;
                sx[sp] = cleft[x]; sj[sp++] = pick;
            }
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
    uint32_t ub, const uint32_t* __restrict__ lvl, uint32_t cap, uint32_t* __restrict__ err, uint32_t max_leaf,
    float pad,
    const uint32_t* __restrict__ items, const uint32_t* __restrict__ children, const uint32_t* __restrict__ n_children,
    const uint32_t* __restrict__ child_offs, const uint32_t* __restrict__ cleft, const uint32_t* __restrict__ ccount,
    const rtq_aabb* __restrict__ cbox, const uint32_t* __restrict__ fl, const uint32_t* __restrict__ pos,
    const uint32_t* __restrict__ bidx, float4* __restrict__ nodes, uint32_t* __restrict__ wide_trav,
    uint32_t* __restrict__ next_items) {
    const uint32_t k = blockIdx.x * kThreads + threadIdx.x;
    const uint32_t n_items = lvl[0], level_base = lvl[1], next_base = level_base + n_items;
    if (k >= ub || k >= n_items) return;
    // the binary traversal node of this node's root cluster (learned order)
    wide_trav[level_base + k] = bidx ? bidx[pos[items[k]]] : kNone;
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
            if (next_base + slot < cap) {
                next_items[slot] = c;
            } else {
                atomicOr(err, 1u);   // more wide nodes than binary ones: a broken invariant
            }
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

// The next level's count (the scan total) and first node index; a level that
// would pass the node capacity ends the tree with an error.
__global__ void k_level_next(const uint32_t* __restrict__ lvl, const uint32_t* __restrict__ total, uint32_t cap,
                             uint32_t* __restrict__ err, uint32_t* __restrict__ next) {
    const uint32_t base = lvl[1] + lvl[0];
    uint32_t count = *total;
    if ((uint64_t)base + count > cap) {
        *err = 1u;
        count = 0;
    }
    next[0] = count;
    next[1] = base;
}

inline uint32_t blocks_for(uint64_t n) { return (uint32_t)((n + kThreads - 1) / kThreads); }

}  // namespace

// Builds s->wide (s->n_wide nodes, s->wide_depth levels) from the cluster
// arrays; fl = first leaf rank of every cluster (device).
rtq_status build_wide(rtq_ctx* ctx, Scene* s, const uint32_t* fl, const uint32_t* pos, const uint32_t* bidx,
                      hipStream_t st) {
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
    if (!s->wide_trav) RTQ_HIP(ctx, dev_alloc(&s->wide_trav, cap));
    s->wide_ordered = false;   // the nodes are rewritten with empty tables
    hipLaunchKernelGGL(k_set_u32, dim3(1), dim3(1), 0, st, items, root);
    RTQ_HIP(ctx, hipGetLastError());
    // cost table of the collapse, bottom-up over the PLOC passes
    const uint32_t nt = s->n_tris, max_leaf = s->params.max_leaf_size;
    uint32_t* choice = nullptr;
    if (nc > nt) {
        float* cost = nullptr;
        RTQ_HIP(ctx, tmp.get(&cost, (size_t)(nc - nt) * kWidth));
        RTQ_HIP(ctx, tmp.get(&choice, nc - nt));
        const uint32_t n_passes = (uint32_t)s->passes.size();
        for (uint32_t k = 0; k < s->small_from; ++k) {
            const auto& pr = s->passes[k];
            hipLaunchKernelGGL(k_dp_pass, dim3(blocks_for(pr.second)), dim3(kThreads), 0, st, pr.first, pr.second, nt,
                               max_leaf, s->cleft, s->cright, s->ccount, s->cbox, cost, choice);
            RTQ_HIP(ctx, hipGetLastError());
        }
        if (s->small_from < n_passes) {
            hipLaunchKernelGGL(k_dp_small, dim3(1), dim3(kDpSmallThreads), 0, st, s->d_passes, s->small_from, n_passes,
                               nt, max_leaf, s->cleft, s->cright, s->ccount, s->cbox, cost, choice);
            RTQ_HIP(ctx, hipGetLastError());
        }
    }
    // levels top-down, kLevelsPerRead at a time between host reads: a level's
    // count and first index stay on the device (lvl), its launches cover an
    // upper bound (8 children per node, at most cap nodes)
    constexpr uint32_t kLevelsPerRead = 4;
    uint32_t* lvl = nullptr;
    RTQ_HIP(ctx, tmp.get(&lvl, 2 * (kLevelsPerRead + 1) + 1 + kLevelsPerRead));
    uint32_t* err = lvl + 2 * (kLevelsPerRead + 1);
    uint32_t* totals = err + 1;
    uint32_t level_items = 1, level_base = 0, levels = 0;
    while (level_items > 0) {
        const uint32_t h_head[3] = {level_items, level_base, 0u};
        RTQ_HIP(ctx, hipMemcpyAsync(lvl, h_head, 2 * sizeof(uint32_t), hipMemcpyHostToDevice, st));
        RTQ_HIP(ctx, hipMemcpyAsync(err, h_head + 2, sizeof(uint32_t), hipMemcpyHostToDevice, st));
        uint64_t ub = level_items;
        for (uint32_t g = 0; g < kLevelsPerRead; ++g) {
            const uint32_t u = (uint32_t)ub;
            uint32_t* lv = lvl + 2 * g;
            hipLaunchKernelGGL(k_wide_collapse, dim3(blocks_for(u)), dim3(kThreads), 0, st, u, lv, items,
                               s->params.max_leaf_size, s->cleft, s->cright, s->ccount, s->cbox, nt, choice, children,
                               nchild, nexp);
            RTQ_HIP(ctx, hipGetLastError());
            rtq_status rs = scan_u32(ctx, nexp, offs, u, totals + g, st);
            if (rs != RTQ_OK) return rs;
            hipLaunchKernelGGL(k_wide_write, dim3(blocks_for(u)), dim3(kThreads), 0, st, u, lv, cap, err,
                               s->params.max_leaf_size, s->box_pad, items, children, nchild, offs, s->cleft,
                               s->ccount, s->cbox, fl, pos, bidx, s->wide, s->wide_trav, next);
            RTQ_HIP(ctx, hipGetLastError());
            hipLaunchKernelGGL(k_level_next, dim3(1), dim3(1), 0, st, lv, totals + g, cap, err, lv + 2);
            RTQ_HIP(ctx, hipGetLastError());
            std::swap(items, next);
            ub = std::min<uint64_t>(8 * ub, cap);
        }
        uint32_t h_lvl[2 * (kLevelsPerRead + 1) + 1];
        RTQ_HIP(ctx, hipMemcpyAsync(h_lvl, lvl, sizeof h_lvl, hipMemcpyDeviceToHost, st));
        RTQ_HIP(ctx, hipStreamSynchronize(st));
        if (h_lvl[2 * (kLevelsPerRead + 1)]) return fail(ctx, RTQ_EINTERNAL, "wide collapse overflow");
        for (uint32_t g = 0; g < kLevelsPerRead; ++g)
            if (h_lvl[2 * g] > 0) ++levels;
        level_items = h_lvl[2 * kLevelsPerRead];
        level_base = h_lvl[2 * kLevelsPerRead + 1];
    }
    s->n_wide = level_base;
    s->wide_depth = levels;
    return RTQ_OK;
}

}  // namespace rtq
