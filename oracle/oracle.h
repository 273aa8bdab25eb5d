/*
 * ORACLE — test infrastructure only.
 *
 * A CPU restatement, in plain C, of the reference's hot path (scene bounds,
 * Morton codes, pair sort, PLOC, pre-order flatten, the GLSL traversal and its
 * Moller-Trumbore leaf test) plus the checker's own watertight leaf test,
 * brute force and ray generators.  Each function cites the reference
 * file:line it follows (paths relative to the reference checkout).
 *
 * Only the tests, the smoke check and the CPU-baseline leg of the benchmark
 * may load this library, and only as the checker / the timed CPU baseline;
 * the product (librtq) never links, loads or falls back to it.
 *
 * Pinning (how each function is checked independently of the GPU):
 *   orc_expand_bits / orc_morton3d   exhaustive check over 0..1023 against a
 *                                    bit-by-bit interleave + known answers
 *   orc_scene_aabb / cube / centroid closed forms on hand-made scenes
 *   orc_sort_pairs                   numpy lexsort
 *   orc_ploc_build                   brute-force NN on tiny N, mutual-pair
 *                                    invariants, 2N-1 clusters, box nesting
 *   orc_refit_boxes                  equals orc_ploc_build's boxes on its own
 *                                    tree; leaf = triangle box, parent = exact
 *                                    merge (invariants)
 *   orc_flatten_preorder             pre-order invariants (root 0, L = i+1,
 *                                    every triangle in exactly one leaf)
 *   orc_refmt / orc_watertight       closed-form hits (axis-aligned triangles),
 *                                    FP64 intersection, edge/vertex sharing
 *   orc_trace_ref / orc_trace_bvh    equality with orc_trace_brute (orc_trace_bvh
 *                                    also for ray origins 10^2-10^7 scene
 *                                    extents away)
 *   orc_gpu_box_check                its count is zero for hits it must keep
 *                                    and positive for the unwidened form on
 *                                    far-away origins (a planted mistake)
 *   orc_raygen_*                     closed forms (centre pixel, unit length,
 *                                    cosine-weighted moments)
 *   orc_mlp_grad                     central finite differences of
 *                                    orc_mlp_loss; all-zero network closed form
 *   orc_mlp_logits                   hand-built networks with known outputs
 *   orc_order_features               hand-computed features of a hand node
 *   orc_ref_box_hit                  hand boxes: inside / behind / border /
 *                                    corner cases with known answers
 *   orc_bvh_overlay                  cut and uncut walks agree; a numpy
 *                                    restatement (a depth-d node is reached iff
 *                                    it and all its ancestors are hit; the
 *                                    right-first DFS reaches the lowest
 *                                    pre-order index last)
 *   orc_shade                        hand hit records with known colours
 */
#ifndef RTQ_ORACLE_H
#define RTQ_ORACLE_H

#include <stdint.h>
#include "../include/rtq/scene_types.h"

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_LEAF_WATERTIGHT 0
#define ORC_LEAF_REFMT 1

#define ORC_MODE_CLOSEST 0
#define ORC_MODE_ANY 1

/* glm mat4 * vec4 (column-major), the product used at bvh.cpp:238-240. */
void orc_mat4_mul_vec4(const float m[16], const float v[4], float out[4]);

/* bvh.cpp:235-251; pad_origin adds the zero-padded triangles' origin when
 * n < 65,536 (the reference scans its MAX-sized vector).  Signed zeros in the
 * result are canonicalised to +0. */
rtq_aabb orc_scene_aabb(const rtq_triangle* tris, uint32_t n, const rtq_mesh_model* models,
                        int pad_origin);
/* bvh.cpp:253-303, including the axis choice of lines 263-272. */
rtq_aabb orc_circumscribed_cube(rtq_aabb box);
/* triangle.cpp:31-33 + bvh.cpp:305-328 for one triangle. */
void orc_normalized_centroid(const rtq_triangle* tri, const rtq_mesh_model* model,
                             const rtq_aabb* cube, float out[3]);
uint32_t orc_expand_bits(uint32_t v);                     /* bvh.cpp:350-356 */
uint32_t orc_morton3d(float x, float y, float z);         /* bvh.cpp:358-372 */
void orc_morton_codes(const rtq_triangle* tris, uint32_t n, const rtq_mesh_model* models,
                      int pad_origin, uint32_t* codes);   /* bvh.cpp:330-348 */
/* std::sort on (code, index) pairs, bvh.cpp:214-232. */
void orc_sort_pairs(uint32_t* codes, uint32_t* idx, uint32_t n);
/* bvh.cpp:383-399 */
rtq_aabb orc_triangle_box(const rtq_triangle* tri, const rtq_mesh_model* model);
/* bvh.cpp:378-381 */
float orc_surface_area(const rtq_aabb* b);
/* bvh.cpp:401-413 */
rtq_aabb orc_merge(const rtq_aabb* a, const rtq_aabb* b);

/* Serial PLOC, bvh.cpp:26-210.  Output arrays hold 2n-1 entries (sorted_idx and
 * sorted_codes n).  Cluster ids of merges are assigned in increasing position
 * order within a pass.  Missing children/parents are 0xFFFFFFFF; leaf clusters
 * carry their triangle id, internal clusters 0.  leaf_count: triangles below
 * each cluster.  Returns the number of PLOC passes, or -1 on a pass without
 * any merge (impossible with finite boxes). */
int orc_ploc_build(const rtq_triangle* tris, uint32_t n, const rtq_mesh_model* models,
                   int pad_origin, uint32_t radius,
                   rtq_aabb* boxes, uint32_t* tri_id, uint32_t* left, uint32_t* right,
                   uint32_t* parent, uint32_t* leaf_count,
                   uint32_t* sorted_idx, uint32_t* sorted_codes);
/* rtq_bvh_refit's boxes for a kept cluster tree (leaves from the triangles,
 * merges in increasing cluster id); boxes holds 2n-1 entries. */
void orc_refit_boxes(const rtq_triangle* tris, const rtq_mesh_model* models, uint32_t n, const uint32_t* tri_id,
                     const uint32_t* left, const uint32_t* right, rtq_aabb* boxes);
/* scene.cpp:189-208: pre-order from root 2n-2 (cluster 0 when n == 1). */
int orc_flatten_preorder(uint32_t n, const rtq_aabb* boxes, const uint32_t* tri_id,
                         const uint32_t* left, const uint32_t* right, rtq_bvh_node* out);
/* Convenience: bounds -> Morton -> sort -> PLOC -> flatten; out has 2n-1 nodes.
 * Returns the PLOC pass count (or -1). */
int orc_build_ref_nodes(const rtq_triangle* tris, uint32_t n, const rtq_mesh_model* models,
                        int pad_origin, uint32_t radius, rtq_bvh_node* out);
/* Depth (root = 0) of the deepest leaf of a pre-order node array. */
uint32_t orc_tree_depth(const rtq_bvh_node* nodes, uint32_t n_nodes);

/* Triangles in world space (model matrix applied once), 9 floats each. */
void orc_world_vertices(const rtq_triangle* tris, uint32_t n, const rtq_mesh_model* models,
                        float* verts9);

/* Leaf tests.  Both return 1 on a hit with (t, u, v) written.
 * orc_refmt restates raytracer.glsl:102-147 with the GLSL vertex order (C++ P2
 * plays the shader's _P1), the culling sign taken from dot(cross(e1, e0), d)
 * without the normalize, and the ray interval [tmin, tmax] added; u, v are the
 * shader's b0, b1.
 * orc_watertight is the watertight test of Woop, Benthin & Wald (JCGT 2013),
 * two-sided, with the FP64 re-evaluation of edge functions that are exactly 0;
 * u, v are the weights of P1 and P2.  Accepts tmin <= t <= tmax. */
int orc_refmt(const rtq_ray* ray, const float* v9, float tmax, float* t, float* u, float* v);
int orc_watertight(const rtq_ray* ray, const float* v9, float tmax, float* t, float* u, float* v);
/* FP64 ray/plane distance of the ray to the triangle's plane (for error stats). */
double orc_t_f64(const rtq_ray* ray, const float* v9);

/* Brute force over all triangles in index order with strict t < best
 * (raytracer.glsl:149-157): closest hit, lowest index on ties; in ANY mode the
 * lowest-index hitting triangle.  hits: (id|-1, t, u, v). */
void orc_trace_brute(const float* verts9, uint32_t n_tris, const rtq_ray* rays, uint64_t n_rays,
                     int leaf, int mode, rtq_hit* hits, int n_threads);
/* The reference traversal, raytracer.glsl:182-295: stack DFS over the pre-order
 * node array, children pushed left then right (right visited first), the
 * reference slab test recomputing 1/d per node, no interval pruning; a leaf
 * result replaces the closest hit only when it is a hit and (t, id) is
 * smaller (the shader's line 279 omits both checks).  In ANY mode the closest
 * hit is still computed and reported (the reference has no any-hit path). */
void orc_trace_ref(const rtq_bvh_node* nodes, const float* verts9, const rtq_ray* rays,
                   uint64_t n_rays, int leaf, rtq_hit* hits, int n_threads);
/* The checker's own CPU traversal on the same node array: nearer child first,
 * interval pruning, conservative slab test on boxes padded by `pad`.
 * Node-visit / triangle-test counts are accumulated into *visits / *tests when
 * those are non-NULL. */
void orc_trace_bvh(const rtq_bvh_node* nodes, const float* verts9, const rtq_ray* rays,
                   uint64_t n_rays, int leaf, int mode, float pad, rtq_hit* hits, int n_threads,
                   uint64_t* visits, uint64_t* tests);

/* The GPU kernels' box tests restated (trace.hip visit(): FMA plane distances
 * from the approximate reciprocal, boxes padded by `pad`, far distance widened
 * by |tf| * rel; the 8-wide form quantised outward in the parent box's frame).
 * For each ray with a hit, every ancestor box of the hit triangle's leaf in the
 * pre-order array is tested with tbest = the hit's t and the reciprocal rounded
 * to nearest and one ulp either way; *lost_binary / *lost_wide count the rays
 * with at least one ancestor box that would be skipped. */
void orc_gpu_box_check(const rtq_bvh_node* nodes, uint32_t n_nodes, const rtq_ray* rays, const rtq_hit* hits,
                       uint64_t n_rays, float pad, float rel, uint64_t* lost_binary, uint64_t* lost_wide);
/* Counter-based RNG: SplitMix64 of (seed, stream, index); u01 = top 24 bits / 2^24. */
uint64_t orc_rng_hash(uint64_t seed, uint64_t stream, uint64_t index);
float orc_u01f(uint64_t seed, uint64_t stream, uint64_t index);

void orc_raygen_pinhole(const float eye[3], const float right[3], const float up[3],
                        const float fwd[3], uint32_t width, uint32_t height, rtq_ray* rays);
void orc_raygen_diffuse(const float* verts9, const rtq_ray* prim, const rtq_hit* hits, uint64_t n,
                        uint64_t seed, float eps, rtq_ray* out);
void orc_raygen_shadow(const float* verts9, const rtq_ray* prim, const rtq_hit* hits, uint64_t n,
                       const float* lights, uint32_t n_lights, float eps, rtq_ray* out);

/* Exclusive scan (wrapping u32). */
void orc_scan_exclusive(const uint32_t* in, uint32_t* out, uint64_t n);

/* ---- learned child ordering (order.c; no reference counterpart) ---------- */
/* Parameters: W1 [32][64] (input-major), b1 [64], W2 [64][64], b2, W3 [64], b3.
 * Samples: 32 features, label, weight (34 floats). */
uint32_t orc_mlp_param_count(void);
float orc_bf16_round(float f);
/* Logits of n rows of 32 features `stride` floats apart, in FP64; bf16 != 0
 * rounds W1, W2, the inputs and the layer-1 activations to BF16 first (the
 * rounding points of the GPU's MFMA inference). */
void orc_mlp_logits(const float* params, const float* x, uint64_t n, uint32_t stride, int bf16, double* out);
/* Mean weighted binary cross-entropy over n samples, FP64. */
double orc_mlp_loss(const float* params, const float* samples, uint64_t n);
/* Its gradient (orc_mlp_param_count() doubles); returns the loss. */
double orc_mlp_grad(const float* params, const float* samples, uint64_t n, double* grad);
/* Features of one 64-B traversal node for (octant, mode); see order.c. */
void orc_order_features(const float* node16, const uint32_t* aux4, uint32_t octant, uint32_t any, float* f32);

/* ---- shading (shade.c): raytracer.glsl:159-179, 182-276, 299-331 -------- */
/* intersectBVH: 0 miss, 1 hit, 2 hit near two slab planes of the box. */
int orc_ref_box_hit(const rtq_ray* ray, const rtq_aabb* b, float threshold);
/* bvhColor per ray for uDepthDisplayBVH = depth (4 floats per ray); cut != 0
 * stops descending below the display depth. */
void orc_bvh_overlay(const rtq_bvh_node* nodes, uint32_t n_nodes, const rtq_ray* rays, uint64_t n,
                     uint32_t depth, int cut, float* rgba);
/* Pixel colour per hit record (flags bit 0 = wireframe); overlay may be NULL. */
void orc_shade(const rtq_triangle* tris, uint32_t n_tris, const rtq_mesh_model* models, const rtq_material* materials,
               uint32_t n_materials, const rtq_hit* hits, uint64_t n, const float* overlay, uint32_t flags,
               float* rgba);

#ifdef __cplusplus
}
#endif

#endif
