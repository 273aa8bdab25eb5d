/*
 * rtq scene types: byte-for-byte the layouts the reference uploads to its GPU
 * buffers, so a caller that fills the reference structs can hand them to rtq
 * unchanged.  Plain C, no glm; usable from C, C++ and HIP host code.
 *
 *   rtq_vec4         glm::vec4                              (16 B)
 *   rtq_triangle     cr::TriangleGPU   triangle.hpp:9-14      (64 B, _ModelId @48)
 *   rtq_mesh_model   cr::MeshModelGPU  mesh.hpp:12-15         (68 B, column-major mat4 + u32)
 *   rtq_aabb         cr::AABB_GPU      bvh.hpp:22-26          (32 B, _Min @0, _Max @16)
 *   rtq_bvh_node     cr::BVH_NodeGPU   bvh.hpp:36-42          (48 B, ids @32/36/40)
 *   rtq_material     cr::MaterialGPU   material.hpp:9-11      (16 B, RGBA)
 *
 * Ray and hit records have no counterpart in the reference (its shader builds a
 * ray per pixel and shades straight into an image, raytracer.glsl:92-100,299-331);
 * they are the trace boundary's own 32-B / 16-B records.
 */
#ifndef RTQ_SCENE_TYPES_H
#define RTQ_SCENE_TYPES_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct rtq_vec4 {
    float x, y, z, w;
} rtq_vec4;

/* cr::TriangleGPU (triangle.hpp:9-14): three vec4 vertices (w = 1 for points),
 * then the model id, aligned to 16 so the record is 64 B. */
typedef struct rtq_triangle {
    rtq_vec4 p0;
    rtq_vec4 p1;
    rtq_vec4 p2;
    uint32_t model_id;
    uint32_t pad_[3];
} rtq_triangle;

/* cr::MeshModelGPU (mesh.hpp:12-15): glm::mat4 is column-major, m[c*4 + r] is
 * column c, row r.  68 B on the host (glm::mat4 has 4-B alignment). */
typedef struct rtq_mesh_model {
    float m[16];
    uint32_t material_id;
} rtq_mesh_model;

/* cr::AABB_GPU (bvh.hpp:22-26): vec3 min, then vec3 max aligned to 16. */
typedef struct rtq_aabb {
    float min[3];
    float pad0_;
    float max[3];
    float pad1_;
} rtq_aabb;

/* cr::BVH_NodeGPU (bvh.hpp:36-42): a node is a leaf iff left == right == 0;
 * internal nodes carry triangle_id 0 (bvh.cpp:415-420). */
typedef struct rtq_bvh_node {
    rtq_aabb box;
    uint32_t triangle_id;
    uint32_t left_child;
    uint32_t right_child;
    uint32_t pad_;
} rtq_bvh_node;

/* cr::MaterialGPU (material.hpp:9-11): one RGBA colour. */
typedef struct rtq_material {
    float color[4];
} rtq_material;

/* Ray: origin, interval start, direction (need not be unit length), interval end.
 * A hit is reported for tmin <= t <= tmax. */
typedef struct rtq_ray {
    float o[3];
    float tmin;
    float d[3];
    float tmax;
} rtq_ray;

/* Hit: triangle id (original, caller-side index) or -1 for a miss, then t and
 * two barycentric weights.  Closest-hit: ties in t go to the lowest id.
 * Any-hit: id >= 0 marks an occluded ray; t/u/v are those of the hit that ended
 * the traversal. Miss: (-1, +inf, 0, 0). */
typedef struct rtq_hit {
    int32_t id;
    float t;
    float u;
    float v;
} rtq_hit;

#ifdef __cplusplus
}  /* extern "C" */

static_assert(sizeof(rtq_vec4) == 16, "glm::vec4 is 16 B");
static_assert(sizeof(rtq_triangle) == 64, "cr::TriangleGPU is 64 B");
static_assert(offsetof(rtq_triangle, model_id) == 48, "_ModelId @48");
static_assert(sizeof(rtq_mesh_model) == 68, "cr::MeshModelGPU is 68 B");
static_assert(sizeof(rtq_aabb) == 32, "cr::AABB_GPU is 32 B");
static_assert(offsetof(rtq_aabb, max) == 16, "_Max @16");
static_assert(sizeof(rtq_bvh_node) == 48, "cr::BVH_NodeGPU is 48 B");
static_assert(offsetof(rtq_bvh_node, triangle_id) == 32, "_TriangleId @32");
static_assert(offsetof(rtq_bvh_node, left_child) == 36, "_LeftChild @36");
static_assert(offsetof(rtq_bvh_node, right_child) == 40, "_RightChild @40");
static_assert(sizeof(rtq_material) == 16, "cr::MaterialGPU is 16 B");
static_assert(sizeof(rtq_ray) == 32, "ray record is 32 B");
static_assert(sizeof(rtq_hit) == 16, "hit record is 16 B");
#else
_Static_assert(sizeof(rtq_triangle) == 64, "cr::TriangleGPU is 64 B");
_Static_assert(sizeof(rtq_mesh_model) == 68, "cr::MeshModelGPU is 68 B");
_Static_assert(sizeof(rtq_aabb) == 32, "cr::AABB_GPU is 32 B");
_Static_assert(sizeof(rtq_bvh_node) == 48, "cr::BVH_NodeGPU is 48 B");
_Static_assert(sizeof(rtq_material) == 16, "cr::MaterialGPU is 16 B");
_Static_assert(sizeof(rtq_ray) == 32, "ray record is 32 B");
_Static_assert(sizeof(rtq_hit) == 16, "hit record is 16 B");
#endif

#endif /* RTQ_SCENE_TYPES_H */
