// Traversal layout constants shared by the trace kernels and the layout
// writers: child reference encoding, LDS stack depths, top-of-tree table.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace rtq {

constexpr uint32_t kNone = 0xFFFFFFFFu;
// Child / stack reference of the traversal layout: bit 31 clear = internal node
// index; bit 31 set = leaf, bits 30..3 first triangle (leaf order), bits 2..0
// count - 1.  kEmpty ends a traversal.
constexpr uint32_t kLeafBit = 0x80000000u;
constexpr uint32_t kEmpty = 0xFFFFFFFFu;
constexpr uint32_t kMaxTris = (1u << 28) - 16u;
constexpr uint32_t kMaxLeaf = 8u;

// Traversal stack entries per lane held in LDS (deeper entries spill to HBM):
// binary layouts share the workgroup's LDS with the top-of-tree copy, the
// 8-wide layout (up to 7 pushes per level, no top copy) keeps a deeper one.
// Batched launches take one of three depths per layout family (the default in
// the middle); the deepest keeps 6 (8-wide) or 7 (binary) workgroups of 256
// lanes within a CU's 160 KB of LDS next to the batch item table.
#ifndef RTQ_STACK_LDS
#define RTQ_STACK_LDS 12
#endif
#ifndef RTQ_STACK_LDS_WIDE
#define RTQ_STACK_LDS_WIDE 16
#endif
constexpr uint32_t kStackChoicesBinary[3] = {8, RTQ_STACK_LDS, 14};
constexpr uint32_t kStackChoicesWide[3] = {8, RTQ_STACK_LDS_WIDE, 20};
inline bool valid_stack_lds(uint32_t v) {
    if (v == 0) return true;
    for (uint32_t k = 0; k < 3; ++k)
        if (v == kStackChoicesBinary[k] || v == kStackChoicesWide[k]) return true;
    return false;
}

// Top of the binary traversal tree, staged in LDS by every trace workgroup:
// the first kTopNodes internal nodes in BFS order from the root, copied with
// the child refs that stay inside this set rewritten to kTopBit | slot
// (internal node indices stay below 2^28, so bit 30 is free in node refs).
#ifndef RTQ_TOP_NODES
#define RTQ_TOP_NODES 63
#endif
constexpr uint32_t kTopNodes = RTQ_TOP_NODES;
constexpr uint32_t kTopBit = 0x40000000u;

__host__ __device__ inline uint32_t leaf_ref(uint32_t first, uint32_t count) {
    return kLeafBit | (first << 3) | (count - 1u);
}

}  // namespace rtq
