// Shading of hit records into RGBA32F pixels and the BVH-depth overlay: the
// caller side of the trace path, restating getColor / main of the reference
// shader (raytracer.glsl:159-179, 299-331) and the box colouring of its
// traversal (raytracer.glsl:182-276).
#include "rtq_internal.hpp"

namespace rtq {
namespace {

constexpr int kShadeThreads = 256;
// BVH overlay stack: a node at depth k < display depth pushes two children, so
// at most display depth + 1 entries are live.
constexpr uint32_t kOverlayStack = 64;

// Constants of raytracer.glsl:69-72.
constexpr float kWireframeWidth = 0.02f;
constexpr float kBvhLineWidth = 0.05f;

// GLSL min / max (GLSL 4.60 section 8.3): min(x, y) = y < x ? y : x,
// max(x, y) = x < y ? y : x, so a NaN second operand returns the first.
__device__ __forceinline__ float gmin(float x, float y) { return y < x ? y : x; }
__device__ __forceinline__ float gmax(float x, float y) { return x < y ? y : x; }

__global__ __launch_bounds__(kShadeThreads) void k_shade(const rtq_hit* __restrict__ hits, uint64_t n,
                                                         const rtq_triangle* __restrict__ tris, uint32_t n_tris,
                                                         const DevModel* __restrict__ models,
                                                         const float4* __restrict__ materials, uint32_t n_materials,
                                                         const float4* __restrict__ overlay, uint32_t flags,
                                                         float4* __restrict__ out) {
    const uint64_t i = (uint64_t)blockIdx.x * kShadeThreads + threadIdx.x;
    if (i >= n) return;
    float4 c = overlay ? overlay[i] : make_float4(0.0f, 0.0f, 0.0f, 1.0f);
    const rtq_hit h = hits[i];
    if (h.id >= 0) {
        // an id past the scene (hits of another scene) has no material
        const uint32_t m = (uint32_t)h.id < n_tris ? models[tris[h.id].model_id].material_id : kNone;
        if (m < n_materials) {
            const float4 mc = materials[m];
            c.x = c.x + mc.x;
            c.y = c.y + mc.y;
            c.z = c.z + mc.z;
            c.w = c.w + mc.w;
        }
        if (flags & RTQ_SHADE_WIREFRAME) {
            const float b2 = (1.0f - h.u) - h.v;
            if (h.u < kWireframeWidth || h.v < kWireframeWidth || b2 < kWireframeWidth)
                c = make_float4(0.0f, 0.0f, 0.0f, 1.0f);
        }
    }
    out[i] = c;
}

// intersectBVH (raytracer.glsl:182-237): 0 miss, 1 hit, 2 hit with the entry
// point within `threshold` of two slab planes.
__device__ __forceinline__ uint32_t ref_box_hit(float ox, float oy, float oz, float dx, float dy, float dz,
                                                const rtq_aabb& b, float threshold) {
    const float inv_x = 1.0f / dx;
    const float tx1 = (b.min[0] - ox) * inv_x;
    const float tx2 = (b.max[0] - ox) * inv_x;
    float tmin = gmin(tx1, tx2);
    float tmax = gmax(tx1, tx2);
    if (tmax < 0.0f || tmin > tmax) return 0;
    const float inv_y = 1.0f / dy;
    const float ty1 = (b.min[1] - oy) * inv_y;
    const float ty2 = (b.max[1] - oy) * inv_y;
    tmin = gmax(tmin, gmin(ty1, ty2));
    tmax = gmin(tmax, gmax(ty1, ty2));
    if (tmax < 0.0f || tmin > tmax) return 0;
    const float inv_z = 1.0f / dz;
    const float tz1 = (b.min[2] - oz) * inv_z;
    const float tz2 = (b.max[2] - oz) * inv_z;
    tmin = gmax(tmin, gmin(tz1, tz2));
    tmax = gmin(tmax, gmax(tz1, tz2));
    if (!(tmax >= 0.0f && tmin <= tmax)) return 0;
    const float ex = ox + dx * tmin, ey = oy + dy * tmin, ez = oz + dz * tmin;
    const bool near_x = fabsf(ex - b.min[0]) < threshold || fabsf(ex - b.max[0]) < threshold;
    const bool near_y = fabsf(ey - b.min[1]) < threshold || fabsf(ey - b.max[1]) < threshold;
    const bool near_z = fabsf(ez - b.min[2]) < threshold || fabsf(ez - b.max[2]) < threshold;
    return ((near_x && near_y) || (near_x && near_z) || (near_y && near_z)) ? 2u : 1u;
}

// One ray per thread; the reference DFS (pop, test, push left then right) cut
// below the display depth, where nodes no longer change the colour.
__global__ __launch_bounds__(kShadeThreads) void k_bvh_overlay(const rtq_bvh_node* __restrict__ nodes,
                                                               const rtq_ray* __restrict__ rays, uint64_t n,
                                                               uint32_t depth, float threshold,
                                                               float4* __restrict__ out) {
    const uint64_t i = (uint64_t)blockIdx.x * kShadeThreads + threadIdx.x;
    if (i >= n) return;
    const float4* rp = reinterpret_cast<const float4*>(rays + i);
    const float4 r0 = rp[0], r1 = rp[1];
    uint32_t stack[kOverlayStack];
    uint8_t sdepth[kOverlayStack];
    uint32_t sp = 0;
    stack[sp] = 0;
    sdepth[sp] = 0;
    ++sp;
    float4 color = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    while (sp > 0) {
        --sp;
        const uint32_t cur = stack[sp];
        const uint32_t d = sdepth[sp];
        const rtq_bvh_node nd = nodes[cur];
        const uint32_t h = ref_box_hit(r0.x, r0.y, r0.z, r1.x, r1.y, r1.z, nd.box, threshold);
        if (h == 0) continue;
        if (d == depth) {
            color = h == 2 ? make_float4(0.7f, 0.0f, 0.7f, 0.1f) : make_float4(0.5f, 0.0f, 0.5f, 0.1f);
            continue;
        }
        if (nd.left_child == 0 && nd.right_child == 0) continue;
        // d < depth <= kOverlayStack - 2 (checked on the host), so sp + 2 <= d + 2 fits
        stack[sp] = nd.left_child;
        sdepth[sp] = (uint8_t)(d + 1);
        ++sp;
        stack[sp] = nd.right_child;
        sdepth[sp] = (uint8_t)(d + 1);
        ++sp;
    }
    out[i] = color;
}

__global__ __launch_bounds__(kShadeThreads) void k_fill4(float4* __restrict__ out, uint64_t n, float4 v) {
    const uint64_t i = (uint64_t)blockIdx.x * kShadeThreads + threadIdx.x;
    if (i < n) out[i] = v;
}

inline uint32_t shade_blocks(uint64_t n) { return (uint32_t)((n + kShadeThreads - 1) / kShadeThreads); }

}  // namespace
}  // namespace rtq

using namespace rtq;

extern "C" {

rtq_status rtq_shade(rtq_ctx* ctx, uint32_t scene_id, const rtq_hit* hits_dev, uint64_t n,
                     const rtq_material* materials_dev, uint32_t n_materials, const float* overlay_dev,
                     uint32_t flags, float* rgba_dev, void* stream) {
    if (!ctx) return RTQ_EINVAL;
    Scene* s = find_scene(ctx, scene_id);
    if (!s) return fail(ctx, RTQ_EINVAL, "shade: unknown scene %u", scene_id);
    if (flags & ~RTQ_SHADE_WIREFRAME) return fail(ctx, RTQ_EINVAL, "shade: unknown flags 0x%x", flags);
    if (n == 0) return RTQ_OK;
    if (!hits_dev || !rgba_dev) return fail(ctx, RTQ_EINVAL, "shade: null buffer");
    if (n_materials && !materials_dev) return fail(ctx, RTQ_EINVAL, "shade: null material array");
    if (n > 0x80000000ull) return fail(ctx, RTQ_ECAPACITY, "shade: n must be <= 2^31");
    RTQ_HIP(ctx, hipSetDevice(ctx->device));
    hipLaunchKernelGGL(k_shade, dim3(shade_blocks(n)), dim3(kShadeThreads), 0, pick_stream(ctx, stream), hits_dev, n,
                       s->tris_ref, s->n_tris, s->models, reinterpret_cast<const float4*>(materials_dev), n_materials,
                       reinterpret_cast<const float4*>(overlay_dev), flags, reinterpret_cast<float4*>(rgba_dev));
    RTQ_HIP(ctx, hipGetLastError());
    return RTQ_OK;
}

rtq_status rtq_bvh_overlay(rtq_ctx* ctx, uint32_t scene_id, const rtq_ray* rays_dev, uint64_t n, uint32_t depth,
                           float* rgba_dev, void* stream) {
    if (!ctx) return RTQ_EINVAL;
    Scene* s = find_scene(ctx, scene_id);
    if (!s) return fail(ctx, RTQ_EINVAL, "bvh_overlay: unknown scene %u", scene_id);
    if (!s->built) return fail(ctx, RTQ_ENOTBUILT, "bvh_overlay: scene %u has no BVH", scene_id);
    if (n == 0) return RTQ_OK;
    if (!rays_dev || !rgba_dev) return fail(ctx, RTQ_EINVAL, "bvh_overlay: null buffer");
    if (n > 0x80000000ull) return fail(ctx, RTQ_ECAPACITY, "bvh_overlay: n must be <= 2^31");
    RTQ_HIP(ctx, hipSetDevice(ctx->device));
    hipStream_t st = pick_stream(ctx, stream);
    float4* out = reinterpret_cast<float4*>(rgba_dev);
    if (depth > s->depth) {
        // no node lies at that depth: bvhColor keeps its initial value
        hipLaunchKernelGGL(k_fill4, dim3(shade_blocks(n)), dim3(kShadeThreads), 0, st, out, n,
                           make_float4(0.0f, 0.0f, 0.0f, 0.0f));
        RTQ_HIP(ctx, hipGetLastError());
        return RTQ_OK;
    }
    if (depth + 2 > kOverlayStack)
        return fail(ctx, RTQ_ECAPACITY, "bvh_overlay: depth %u exceeds %u", depth, kOverlayStack - 2);
    // BVH_LINE_WIDTH / (uDepthDisplayBVH + 1.f), raytracer.glsl:222
    const float threshold = kBvhLineWidth / ((float)depth + 1.0f);
    const rtq_status rs = ensure_ref_nodes(ctx, s, st);
    if (rs != RTQ_OK) return rs;
    hipLaunchKernelGGL(k_bvh_overlay, dim3(shade_blocks(n)), dim3(kShadeThreads), 0, st, s->ref_nodes, rays_dev, n,
                       depth, threshold, out);
    RTQ_HIP(ctx, hipGetLastError());
    return RTQ_OK;
}

}  // extern "C"
