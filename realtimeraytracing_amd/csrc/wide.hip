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

// a child that becomes an 8-wide node of the next level: a merged cluster
// (ids from nt on) of more than max_leaf triangles
__device__ __forceinline__ bool expandable(uint32_t c, uint32_t max_leaf, uint32_t nt, const uint32_t* ccount) {
    return c >= nt && ccount[c] > max_leaf;
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

// Row and choice word of a cluster of `cnt` triangles and box area a from its
// children's rows dl, dr.
__device__ __forceinline__ uint32_t dp_combine(const float* dl, const float* dr, float a, uint32_t cnt,
                                               uint32_t max_leaf, float* row) {
    float bn = INFINITY;
    uint32_t kn = 1;
    for (uint32_t k = 1; k < kWidth; ++k) {
        const float v = dl[k - 1] + dr[kWidth - k - 1];
        if (v < bn) { bn = v; kn = k; }
    }
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
    return ch;
}

__device__ __forceinline__ void dp_cluster(uint32_t c, uint32_t nt, uint32_t max_leaf,
                                           const uint32_t* __restrict__ cleft, const uint32_t* __restrict__ cright,
                                           const uint32_t* __restrict__ ccount, const rtq_aabb* __restrict__ cbox,
                                           float* __restrict__ cost, uint32_t* __restrict__ choice) {
    float dl[kWidth], dr[kWidth], row[kWidth];
    dp_row(cleft[c], nt, cost, cbox, dl);
    dp_row(cright[c], nt, cost, cbox, dr);
    const uint32_t ch = dp_combine(dl, dr, box_area(cbox[c]), ccount[c], max_leaf, row);
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

// Passes [k0, k1) at the top of the tree in one workgroup, bottom-up, with
// the cost rows of their clusters [c_lo, nc) in LDS (a child from an earlier
// pass gives its row from global memory); a barrier between passes.  Each
// thread keeps the child ids, box area and triangle count of its clusters
// c_lo + t + kLdsThreads * s in registers.  nc - c_lo <= kLdsClusters.
__global__ __launch_bounds__(kLdsThreads) void k_dp_lds(const uint2* __restrict__ passes, uint32_t k0, uint32_t k1,
                                                        uint32_t c_lo, uint32_t nc, uint32_t nt, uint32_t max_leaf,
                                                        const uint32_t* __restrict__ cleft,
                                                        const uint32_t* __restrict__ cright,
                                                        const uint32_t* __restrict__ ccount,
                                                        const rtq_aabb* __restrict__ cbox, const float* __restrict__ cost,
                                                        uint32_t* __restrict__ choice) {
    constexpr int kSlots = kLdsClusters / kLdsThreads;
    __shared__ float s_row[kLdsClusters * kWidth];
    const uint32_t t = threadIdx.x;
    const uint32_t R = nc - c_lo;
    uint32_t L[kSlots], Rc[kSlots], N[kSlots];
    float A[kSlots];
#pragma unroll
    for (int q = 0; q < kSlots; ++q) {
        const uint32_t idx = t + kLdsThreads * q;
        const bool in = idx < R;
        L[q] = in ? cleft[c_lo + idx] : 0u;
        Rc[q] = in ? cright[c_lo + idx] : 0u;
        N[q] = in ? ccount[c_lo + idx] : 0u;
        A[q] = in ? box_area(cbox[c_lo + idx]) : 0.0f;
    }
    for (uint32_t k = k0; k < k1; ++k) {
        const uint2 p = passes[k];
#pragma unroll
        for (int q = 0; q < kSlots; ++q) {
            const uint32_t idx = t + kLdsThreads * q, c = c_lo + idx;
            if (idx >= R || c < p.x || c >= p.x + p.y) continue;
            float dl[kWidth], dr[kWidth], row[kWidth];
            if (L[q] >= c_lo) {
#pragma unroll
                for (int j = 0; j < kWidth; ++j) dl[j] = s_row[(L[q] - c_lo) * kWidth + j];
            } else {
                dp_row(L[q], nt, cost, cbox, dl);
            }
            if (Rc[q] >= c_lo) {
#pragma unroll
                for (int j = 0; j < kWidth; ++j) dr[j] = s_row[(Rc[q] - c_lo) * kWidth + j];
            } else {
                dp_row(Rc[q], nt, cost, cbox, dr);
            }
            choice[c - nt] = dp_combine(dl, dr, A[q], N[q], max_leaf, row);
#pragma unroll
            for (int j = 0; j < kWidth; ++j) s_row[idx * kWidth + j] = row[j];
        }
        __syncthreads();
    }
}

// The slots of node cluster c as the cost table splits them, children left
// to right; returns their count.  Clusters below nt are the single triangles
// (no children), so a step's three loads do not wait for one another.
__device__ __forceinline__ uint32_t collapse_node(uint32_t c, const uint32_t* __restrict__ cleft,
                                                  const uint32_t* __restrict__ cright, uint32_t nt,
                                                  const uint32_t* __restrict__ choice, uint32_t* ch) {
    uint32_t n = 0;
    if (c < nt) {
        ch[n++] = c;  // single-triangle scene: the root itself is the only (leaf) child
        return n;
    }
    const uint32_t kn = ((choice[c - nt] >> 21) & 7u) + 1u;
    uint32_t sx[kWidth], sj[kWidth];
    int sp = 0;
    sx[sp] = cright[c]; sj[sp++] = kWidth - kn;
    sx[sp] = cleft[c]; sj[sp++] = kn;
    while (sp > 0) {
        --sp;
        const uint32_t x = sx[sp], j = sj[sp];
        if (j == 1 || x < nt) {
            ch[n++] = x;
            continue;
        }
        const uint32_t cw = choice[x - nt], l = cleft[x], r = cright[x];
        const uint32_t pick = (cw >> (3 * (j - 2))) & 7u;
        if (pick == 0) {
            sx[sp] = x; sj[sp++] = j - 1;
        } else {
            sx[sp] = r; sj[sp++] = j - pick;
            sx[sp] = l; sj[sp++] = pick;
        }
    }
    return n;
}

__device__ int exponent_for(double extent) {
    // smallest e with 255 * 2^e >= extent, within the normal float range
    int e = extent > 0.0 ? ilogb(extent) - 7 : -126;
    if (e < -126) e = -126;
    while (e < 127 && ldexp(255.0, e) < extent) ++e;
    while (e > -126 && ldexp(255.0, e - 1) >= extent) --e;
    return e;
}

// The first four float4 of an 8-wide node with children ch[0..n) in slots
// 0..n-1: origin and meta (exponents, valid mask), then the x, y and z planes.
// Child boxes are widened by pad and quantised outward in exact (double)
// arithmetic, so the decoded box contains the padded child box.
__device__ void write_wide_boxes(const uint32_t* ch, uint32_t n, const rtq_aabb* __restrict__ cbox, float pad,
                                 float4* __restrict__ o) {
    float lo[kWidth][3], hi[kWidth][3];
    float org[3] = {INFINITY, INFINITY, INFINITY};
    double top[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (uint32_t i = 0; i < n; ++i) {
        const rtq_aabb b = cbox[ch[i]];
        for (int q = 0; q < 3; ++q) {
            lo[i][q] = b.min[q] - pad;
            hi[i][q] = b.max[q] + pad;
            org[q] = fminf(org[q], lo[i][q]);
            top[q] = fmax(top[q], (double)hi[i][q]);
        }
    }
    int ex[3];
    double scale[3];
    for (int q = 0; q < 3; ++q) {
        ex[q] = exponent_for(top[q] - (double)org[q]);
        scale[q] = ldexp(1.0, ex[q]);
    }
    uint32_t qlo[3][2] = {{0, 0}, {0, 0}, {0, 0}}, qhi[3][2] = {{0, 0}, {0, 0}, {0, 0}};
    uint32_t valid = 0;
    for (uint32_t i = 0; i < n; ++i) {
        valid |= 1u << i;
        for (int q = 0; q < 3; ++q) {
            // exact: float differences and division by a power of two are exact in double
            double l = floor(((double)lo[i][q] - (double)org[q]) / scale[q]);
            double h = ceil(((double)hi[i][q] - (double)org[q]) / scale[q]);
            l = fmin(fmax(l, 0.0), 255.0);
            h = fmin(fmax(h, 0.0), 255.0);
            qlo[q][i >> 2] |= (uint32_t)l << (8 * (i & 3));
            qhi[q][i >> 2] |= (uint32_t)h << (8 * (i & 3));
        }
    }
    // empty slots: an inverted box (lower planes 255, upper 0) that a ray
    // misses on every axis unless its origin lies more than 2^19 node extents
    // away; their child ref is then harmless (one triangle test, see the
    // level kernel), so traversal needs no test of the valid bits
    for (uint32_t i = n; i < kWidth; ++i)
        for (int q = 0; q < 3; ++q) qlo[q][i >> 2] |= 0xFFu << (8 * (i & 3));
    const uint32_t meta = (uint32_t)(ex[0] + 127) | ((uint32_t)(ex[1] + 127) << 8) |
                          ((uint32_t)(ex[2] + 127) << 16) | (valid << 24);
    o[0] = make_float4(org[0], org[1], org[2], __uint_as_float(meta));
    o[1] = make_float4(__uint_as_float(qlo[0][0]), __uint_as_float(qlo[0][1]), __uint_as_float(qhi[0][0]),
                       __uint_as_float(qhi[0][1]));
    o[2] = make_float4(__uint_as_float(qlo[1][0]), __uint_as_float(qlo[1][1]), __uint_as_float(qhi[1][0]),
                       __uint_as_float(qhi[1][1]));
    o[3] = make_float4(__uint_as_float(qlo[2][0]), __uint_as_float(qlo[2][1]), __uint_as_float(qhi[2][0]),
                       __uint_as_float(qhi[2][1]));
}

// Tile status of the level scan: value in the low 32 bits, state on top.
constexpr uint64_t kLvAgg = 1ull << 62;     // the tile's own count
constexpr uint64_t kLvPre = 2ull << 62;     // the counts through the tile
constexpr uint64_t kLvPoison = 3ull << 62;  // a wait at or before the tile gave up
constexpr uint64_t kLvState = 3ull << 62;
constexpr uint32_t kLvSpinLimit = 1u << 24;
enum : uint32_t { kWideErrOverflow = 1, kWideErrScan = 2 };

struct LevelArgs {
    uint32_t* lv;              // this level's words: count, first node, first status word, tile ticket
    uint64_t* status;          // tile status words, n_status of them, zero before the first level
    uint32_t n_status;
    uint32_t cap;              // node capacity
    uint32_t* err;
    uint32_t max_leaf;
    uint32_t nt;
    float pad;
    const uint32_t* items;     // this level's node clusters
    uint32_t* next_items;      // the next level's
    const uint32_t* cleft;
    const uint32_t* cright;
    const uint32_t* ccount;
    const rtq_aabb* cbox;
    const uint32_t* choice;
    const uint32_t* fl;
    const uint32_t* pos;
    const uint32_t* bidx;
    float4* nodes;
    uint32_t* wide_trav;
    uint32_t* kids;            // 8 child clusters per node (kNone past the last)
};

// One level of the collapse in one launch.  Each thread takes one node of the
// level and picks its slots; the next level's nodes are numbered by an
// exclusive scan of the nodes' expandable children in node order, across
// workgroups by decoupled look-back (tiles numbered in the order workgroups
// start, so every tile waited on is running; the wait is bounded and a tile
// that gives up poisons its status and sets an error bit); then the thread
// writes its node.  The last tile writes the next level's words (lv[4..7]).
// The grid covers an upper bound of the level's count; workgroups past the
// count, and every workgroup of an empty level, return at once.
__global__ __launch_bounds__(kThreads) void k_wide_level(LevelArgs a) {
    __shared__ uint32_t s_tile, s_excl, s_bad, s_w[kThreads / 64];
    const uint32_t count = a.lv[0];
    const uint32_t tiles = (count + kThreads - 1) / kThreads;
    if (blockIdx.x >= tiles) return;
    const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
    if (tid == 0) s_tile = atomicAdd(a.lv + 3, 1u);
    __syncthreads();
    const uint32_t tile = s_tile;
    const uint32_t k = tile * kThreads + tid;
    const uint32_t level_base = a.lv[1], sbase = a.lv[2];

    uint32_t ch[kWidth];
    uint32_t n = 0, ne = 0, expand = 0;
    if (k < count) {
        n = collapse_node(a.items[k], a.cleft, a.cright, a.nt, a.choice, ch);
        for (uint32_t i = 0; i < n; ++i)
            if (expandable(ch[i], a.max_leaf, a.nt, a.ccount)) {
                expand |= 1u << i;
                ++ne;
            }
    }
    // exclusive sum of ne over the workgroup
    uint32_t x = ne;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t y = __shfl_up(x, d, 64);
        if ((int)lane >= d) x += y;
    }
    if (lane == 63) s_w[wave] = x;
    __syncthreads();
    uint32_t wave_off = 0, total = 0;
#pragma unroll
    for (uint32_t w = 0; w < kThreads / 64; ++w) {
        if (w < wave) wave_off += s_w[w];
        total += s_w[w];
    }
    const uint32_t local = wave_off + x - ne;

    // counts before this tile: decoupled look-back by wave 0
    if (wave == 0) {
        bool failed = (uint64_t)sbase + tiles > a.n_status;
        uint32_t excl = 0;
        if (!failed) {
            uint64_t* const st = a.status + sbase;
            if (tile == 0) {
                if (lane == 0) __hip_atomic_store(&st[0], kLvPre | total, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            } else {
                if (lane == 0)
                    __hip_atomic_store(&st[tile], kLvAgg | total, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                int64_t top = (int64_t)tile - 1;
                uint32_t spins = 0;
                for (;;) {
                    const int64_t idx = top - (int64_t)lane;  // lane 0 = nearest tile
                    uint64_t w = kLvPre;                       // before tile 0: an empty prefix
                    if (idx >= 0) w = __hip_atomic_load(&st[idx], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    while (__ballot((w & kLvState) == 0) != 0) {
                        if (++spins > kLvSpinLimit) {
                            failed = true;
                            break;
                        }
                        __builtin_amdgcn_s_sleep(1);
                        if ((w & kLvState) == 0)
                            w = __hip_atomic_load(&st[idx], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    }
                    if (failed) break;
                    // nearest tile with counts through it (a prefix, or poison)
                    const unsigned long long pre = __ballot((w & kLvState) >= kLvPre);
                    const int stop = pre ? __ffsll((long long)pre) - 1 : 63;
                    if (pre && __shfl((uint32_t)(w >> 62), stop, 64) == 3u) {
                        failed = true;
                        break;
                    }
                    uint32_t c = (int)lane <= stop ? (uint32_t)w : 0u;
#pragma unroll
                    for (int d = 32; d >= 1; d >>= 1) c += __shfl_xor(c, d, 64);
                    excl += c;
                    if (pre) break;
                    top -= 64;
                }
                if (lane == 0)
                    __hip_atomic_store(&st[tile], failed ? kLvPoison : (kLvPre | (uint64_t)(excl + total)),
                                       __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
        }
        if (lane == 0) {
            s_excl = excl;
            s_bad = failed ? 1u : 0u;
            if (failed) atomicOr(a.err, kWideErrScan);
        }
    }
    __syncthreads();
    if (s_bad) return;
    const uint32_t excl = s_excl;
    const uint32_t next_base = level_base + count;
    if (tile == tiles - 1 && tid == 0) {
        uint32_t next_count = excl + total;
        if ((uint64_t)next_base + next_count > a.cap) {
            atomicOr(a.err, kWideErrOverflow);   // more wide nodes than binary ones: a broken invariant
            next_count = 0;
        }
        a.lv[4] = next_count;
        a.lv[5] = next_base;
        a.lv[6] = sbase + tiles;
        a.lv[7] = 0u;
    }
    if (k >= count) return;

    // the node: children boxes quantised outward against the node's origin
    // the binary traversal node of this node's root cluster (learned order)
    a.wide_trav[level_base + k] = a.bidx ? a.bidx[a.pos[a.items[k]]] : kNone;
    float4* o = a.nodes + 8ull * (level_base + k);
    write_wide_boxes(ch, n, a.cbox, a.pad, o);
    uint32_t ref[kWidth];
    uint32_t rank = 0;
    for (uint32_t i = 0; i < kWidth; ++i) {
        // an empty slot's box is inverted (write_wide_boxes); should a far ray
        // still enter it, its ref is the leaf of traversal triangle 0 alone:
        // a real triangle, so whatever that test reports is a genuine hit
        ref[i] = leaf_ref(0u, 1u);
        if (i >= n) continue;
        const uint32_t c = ch[i];
        if ((expand >> i) & 1u) {
            const uint32_t slot = excl + local + rank++;
            ref[i] = next_base + slot;
            if ((uint64_t)next_base + slot < a.cap) {
                a.next_items[slot] = c;
            } else {
                atomicOr(a.err, kWideErrOverflow);
            }
        } else {
            ref[i] = leaf_ref(a.fl[c], a.ccount[c]);
        }
    }
    o[4] = make_float4(__uint_as_float(ref[0]), __uint_as_float(ref[1]), __uint_as_float(ref[2]),
                       __uint_as_float(ref[3]));
    o[5] = make_float4(__uint_as_float(ref[4]), __uint_as_float(ref[5]), __uint_as_float(ref[6]),
                       __uint_as_float(ref[7]));
    o[6] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    o[7] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    // the node's child clusters, for a refit
    uint4* kid = reinterpret_cast<uint4*>(a.kids + 8ull * (level_base + k));
    kid[0] = make_uint4(ch[0], n > 1 ? ch[1] : kNone, n > 2 ? ch[2] : kNone, n > 3 ? ch[3] : kNone);
    kid[1] = make_uint4(n > 4 ? ch[4] : kNone, n > 5 ? ch[5] : kNone, n > 6 ? ch[6] : kNone, n > 7 ? ch[7] : kNone);
}

// New boxes for the kept 8-wide nodes after a refit: each node's child
// clusters (kids, kNone past the last) quantised again from their refitted
// boxes; child refs stay, learned tables are cleared.
__global__ __launch_bounds__(kThreads) void k_wide_refit(uint32_t n_wide, const uint32_t* __restrict__ kids,
                                                         const rtq_aabb* __restrict__ cbox, float pad,
                                                         float4* __restrict__ nodes) {
    const uint32_t k = blockIdx.x * kThreads + threadIdx.x;
    if (k >= n_wide) return;
    const uint4* kp = reinterpret_cast<const uint4*>(kids + 8ull * k);
    const uint4 k0 = kp[0], k1 = kp[1];
    const uint32_t ch[kWidth] = {k0.x, k0.y, k0.z, k0.w, k1.x, k1.y, k1.z, k1.w};
    uint32_t n = 0;
    while (n < kWidth && ch[n] != kNone) ++n;
    float4* o = nodes + 8ull * k;
    write_wide_boxes(ch, n, cbox, pad, o);
    o[6] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
}

__global__ void k_set_u32(uint32_t* p, uint32_t v) { *p = v; }

inline uint32_t blocks_for(uint64_t n) { return (uint32_t)((n + kThreads - 1) / kThreads); }

}  // namespace

uint32_t wide_cost_steps(const Scene* s) {
    if (s->n_clusters <= s->n_tris) return 0;
    return s->lds_from + (s->lds_from < (uint32_t)s->passes.size() ? 1u : 0u);
}

rtq_status wide_cost_table(rtq_ctx* ctx, const Scene* s, float* cost, uint32_t* choice, uint32_t first,
                           uint32_t last, hipStream_t st) {
    const uint32_t nt = s->n_tris, max_leaf = s->params.max_leaf_size;
    const uint32_t n_passes = (uint32_t)s->passes.size();
    for (uint32_t k = first; k < last; ++k) {
        if (k < s->lds_from) {
            const auto& pr = s->passes[k];
            hipLaunchKernelGGL(k_dp_pass, dim3(blocks_for(pr.second)), dim3(kThreads), 0, st, pr.first, pr.second,
                               nt, max_leaf, s->cleft, s->cright, s->ccount, s->cbox, cost, choice);
        } else {
            hipLaunchKernelGGL(k_dp_lds, dim3(1), dim3(kLdsThreads), 0, st, s->d_passes, s->lds_from, n_passes,
                               s->passes[s->lds_from].first, s->n_clusters, nt, max_leaf, s->cleft, s->cright,
                               s->ccount, s->cbox, cost, choice);
        }
        RTQ_HIP(ctx, hipGetLastError());
    }
    return RTQ_OK;
}

// Builds s->wide (s->n_wide nodes, s->wide_depth levels) from the cluster
// arrays and the cost table's choices; fl = first leaf rank of every cluster
// (device).
rtq_status build_wide(rtq_ctx* ctx, Scene* s, const uint32_t* choice, const uint32_t* fl, const uint32_t* pos,
                      const uint32_t* bidx, hipStream_t st) {
    const uint32_t nc = s->n_clusters;
    const uint32_t root = nc - 1;
    const uint32_t cap = std::max<uint32_t>(s->n_trav, 1);  // wide nodes are a subset of the binary ones
    if (nc > s->n_tris && !choice) return fail(ctx, RTQ_EINTERNAL, "wide collapse without a cost table");
    uint32_t *items = nullptr, *next = nullptr;
    Temps tmp(ctx);
    RTQ_HIP(ctx, tmp.get(&items, cap));
    RTQ_HIP(ctx, tmp.get(&next, cap));
    // tile status words of every level's scan: a level of c nodes takes
    // ceil(c / kThreads) of them, and there are at most depth + 1 levels
    const uint32_t n_status = cap / kThreads + s->depth + 2;
    uint64_t* status = nullptr;
    RTQ_HIP(ctx, tmp.get(&status, n_status));
    RTQ_HIP(ctx, hipMemsetAsync(status, 0, sizeof(uint64_t) * n_status, st));
    // a refit writes the same buffer again (its size depends on the topology only)
    if (!s->wide) RTQ_HIP(ctx, take(s->kept_wide, &s->wide, 8ull * cap));
    if (!s->wide_trav) RTQ_HIP(ctx, take(s->kept_wide_trav, &s->wide_trav, cap));
    if (!s->wide_kids) RTQ_HIP(ctx, take(s->kept_wide_kids, &s->wide_kids, 8ull * cap));
    s->wide_ordered = false;   // the nodes are rewritten with empty tables
    hipLaunchKernelGGL(k_set_u32, dim3(1), dim3(1), 0, st, items, root);
    RTQ_HIP(ctx, hipGetLastError());
    // levels top-down, one launch each, kLevelsPerRead of them between host
    // reads: a level's words (count, first node, first status word, ticket)
    // stay on the device, its launch covers an upper bound of its count (8
    // children per node, at most cap nodes)
    constexpr uint32_t kLevelsPerRead = 6;
    constexpr uint32_t kLvWords = 4 * (kLevelsPerRead + 1);
    uint32_t* lvl = nullptr;
    RTQ_HIP(ctx, tmp.get(&lvl, kLvWords + 1));
    uint32_t* err = lvl + kLvWords;
    LevelArgs la{};
    la.status = status;
    la.n_status = n_status;
    la.cap = cap;
    la.err = err;
    la.max_leaf = s->params.max_leaf_size;
    la.nt = s->n_tris;
    la.pad = s->box_pad;
    la.cleft = s->cleft;
    la.cright = s->cright;
    la.ccount = s->ccount;
    la.cbox = s->cbox;
    la.choice = choice;
    la.fl = fl;
    la.pos = pos;
    la.bidx = bidx;
    la.nodes = s->wide;
    la.wide_trav = s->wide_trav;
    la.kids = s->wide_kids;
    uint32_t level_items = 1, level_base = 0, level_status = 0, levels = 0, n_wide = 0;
    while (level_items > 0) {
        // the group's first level; the others (and the error word) zero until written
        uint32_t h_head[kLvWords + 1] = {};
        h_head[0] = level_items;
        h_head[1] = level_base;
        h_head[2] = level_status;
        RTQ_HIP(ctx, hipMemcpyAsync(lvl, h_head, sizeof h_head, hipMemcpyHostToDevice, st));
        uint64_t ub = level_items;
        for (uint32_t g = 0; g < kLevelsPerRead; ++g) {
            la.lv = lvl + 4 * g;
            la.items = items;
            la.next_items = next;
            hipLaunchKernelGGL(k_wide_level, dim3(blocks_for(ub)), dim3(kThreads), 0, st, la);
            RTQ_HIP(ctx, hipGetLastError());
            std::swap(items, next);
            ub = std::min<uint64_t>(8 * ub, cap);
        }
        uint32_t h_lvl[kLvWords + 1];
        HostReads rd(ctx, st);
        RTQ_HIP(ctx, rd.add(h_lvl, lvl, sizeof h_lvl));
        RTQ_HIP(ctx, rd.wait());
        if (h_lvl[kLvWords] & kWideErrScan) return fail(ctx, RTQ_EINTERNAL, "wide collapse: a level scan gave up");
        if (h_lvl[kLvWords]) return fail(ctx, RTQ_EINTERNAL, "wide collapse overflow");
        for (uint32_t g = 0; g < kLevelsPerRead; ++g)
            if (h_lvl[4 * g] > 0) {
                ++levels;
                n_wide = h_lvl[4 * g + 1] + h_lvl[4 * g];
            }
        level_items = h_lvl[4 * kLevelsPerRead];
        level_base = h_lvl[4 * kLevelsPerRead + 1];
        level_status = h_lvl[4 * kLevelsPerRead + 2];
    }
    s->n_wide = n_wide;
    s->wide_depth = levels;
    return RTQ_OK;
}

rtq_status refit_wide(rtq_ctx* ctx, Scene* s, hipStream_t st) {
    if (!s->wide || !s->wide_kids) return fail(ctx, RTQ_EINTERNAL, "refit: the scene has no 8-wide nodes");
    if (s->n_wide == 0) return RTQ_OK;
    hipLaunchKernelGGL(k_wide_refit, dim3(blocks_for(s->n_wide)), dim3(kThreads), 0, st, s->n_wide, s->wide_kids,
                       s->cbox, s->box_pad, s->wide);
    RTQ_HIP(ctx, hipGetLastError());
    s->wide_ordered = false;   // the tables were cleared
    return RTQ_OK;
}

}  // namespace rtq
