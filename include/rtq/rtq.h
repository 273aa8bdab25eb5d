/*
 * rtq — C ABI of the MI355X (gfx950) ray-query engine.
 *
 * The reference has no FFI: its seam is the srcCommon C++ scene API
 * (cr::BVH built in its constructor, bvh.hpp:84-124 / bvh.cpp:11-24), the
 * pre-order flatten done by the OpenGL scene (scene.cpp:189-208) and the GLSL
 * traversal dispatched per frame (application.cpp:220-246, raytracer.glsl:246-331).
 * Every entry point below replaces one of those, cited on its declaration.
 *
 * Conventions
 *   - Every call returns an rtq_status; nothing exits the process (the
 *     reference calls exit() through ErrorHandler::handle, errorHandler.cpp:13-45).
 *     rtq_last_error() gives the message of the last failing call on a context.
 *   - "_dev" pointers are device (HBM) pointers on the context's device;
 *     everything else is host memory.  Streams are hipStream_t passed as void*;
 *     NULL means the context's own stream.
 *   - A context is bound to one device and is not thread-safe.
 *   - Plain pointers and sizes only: no C++ or torch types cross this boundary.
 */
#ifndef RTQ_RTQ_H
#define RTQ_RTQ_H

#include "scene_types.h"

#ifdef __cplusplus
extern "C" {
#endif

/* The library is built with hidden visibility; only these declarations export. */
#if defined(__GNUC__) || defined(__clang__)
#pragma GCC visibility push(default)
#endif

#define RTQ_ABI_VERSION 2

typedef struct rtq_ctx rtq_ctx;

typedef enum rtq_status {
    RTQ_OK = 0,
    RTQ_EINVAL = 1,      /* bad argument (null pointer, N = 0, non-finite vertex, ...) */
    RTQ_ENOMEM = 2,      /* device allocation failed */
    RTQ_EHIP = 3,        /* a HIP runtime call failed */
    RTQ_ECAPACITY = 4,   /* caller buffer too small / size beyond a documented limit */
    RTQ_EINTERNAL = 5,   /* an internal invariant failed (e.g. a PLOC pass made no merge) */
    RTQ_ENOTBUILT = 6    /* scene has no BVH yet */
} rtq_status;

/* ---- context ------------------------------------------------------------ */

/* Replaces the GL context + SSBO set created by glr::Scene (scene.cpp:69-108). */
rtq_status rtq_create(rtq_ctx** out, int hip_device);
void rtq_destroy(rtq_ctx* ctx);
const char* rtq_last_error(const rtq_ctx* ctx);
const char* rtq_status_string(rtq_status s);
int rtq_abi_version(void);
/* Synchronise the context's stream (all work issued with stream == NULL). */
rtq_status rtq_synchronize(rtq_ctx* ctx);

/* ---- scene -------------------------------------------------------------- */

/* Upload a scene given in the reference layouts (triangles: 64-B TriangleGPU,
 * models: 68-B MeshModelGPU).  Replaces Scene::bindSSBO's triangle/model
 * uploads (scene.cpp:130-156) and the input copy of BVH::BVH (bvh.cpp:11-17).
 * Vertices are transformed by their model matrix once, here, in the order
 * glm's mat4 * vec4 uses (bvh.cpp:238-240), instead of per intersection
 * (raytracer.glsl:107-109).  n_tris has no 65,536 cap (triangle.hpp:21).
 * Rejects n_tris == 0, model ids >= n_models and vertex coordinates that are
 * not finite or exceed 2^28 in magnitude (checked on the device after the
 * copy; the message names the lowest failing triangle).
 * Host arrays; the library copies them. */
rtq_status rtq_scene_upload(rtq_ctx* ctx, const rtq_triangle* tris, uint32_t n_tris,
                            const rtq_mesh_model* models, uint32_t n_models,
                            uint32_t* scene_id);
rtq_status rtq_scene_release(rtq_ctx* ctx, uint32_t scene_id);

/* Per-frame scene edits (no counterpart in the reference, which builds its BVH
 * once when it binds the scene, scene.cpp:110-187): new triangle vertices
 * and/or model matrices, same counts as uploaded (NULL keeps an array), with
 * the upload's checks.  Waits for work queued on the device.  New triangles
 * are copied to a staging array (N x 64 B, allocated on the first such update
 * and kept with the scene), checked there and only then copied over the
 * scene's, so a rejected update leaves the scene unchanged and its arrays keep
 * their addresses.  A built scene becomes stale: trace, export and shading
 * calls return RTQ_ENOTBUILT until rtq_bvh_refit or rtq_bvh_build. */
rtq_status rtq_scene_update(rtq_ctx* ctx, uint32_t scene_id, const rtq_triangle* tris, uint32_t n_tris,
                            const rtq_mesh_model* models, uint32_t n_models);

/* rtq_scene_update with the new triangles already in device memory (n_tris
 * 64-B records at tris_dev on the context's device, e.g. written by an
 * animation kernel; no counterpart in the reference either): checked where
 * they lie and copied over the scene's array on the device, with no host
 * copy and no staging array.  Waits for work queued on the device; the
 * caller's array may be reused once the call returns.  Rejects an array that
 * is not device memory. */
rtq_status rtq_scene_update_device(rtq_ctx* ctx, uint32_t scene_id, const rtq_triangle* tris_dev, uint32_t n_tris);

/* ---- OBJ ingest (host only; no context or device needed) ------------------ */

/* Mesh::load (mesh.cpp:186-263): the triangles of a Wavefront OBJ file in
 * file order as TriangleGPU records (w = 1, model_id as given), the way the
 * reference reads it through tinyobjloader with triangulation on.  Reads `v`
 * and `f` statements (v, v/vt, v//vn, v/vt/vn references, 1-based or
 * negative); faces with fewer than 3 vertices are skipped; a quad is split
 * along its shorter diagonal (tinyobjloader 2.x), larger polygons into a fan
 * (0, k, k+1).  Call with tris == NULL to get *n_tris, then with a buffer of
 * `capacity` records (RTQ_ECAPACITY when too small).  Malformed input gives
 * RTQ_EINVAL with a message in msg (may be NULL; msg_capacity bytes). */
rtq_status rtq_obj_load(const char* path, uint32_t model_id, rtq_triangle* tris, uint64_t capacity,
                        uint64_t* n_tris, char* msg, uint64_t msg_capacity);
/* The same for OBJ text in memory (len bytes). */
rtq_status rtq_obj_parse(const char* text, uint64_t len, uint32_t model_id, rtq_triangle* tris,
                         uint64_t capacity, uint64_t* n_tris, char* msg, uint64_t msg_capacity);

/* ---- BVH build ---------------------------------------------------------- */

/* Build flags. */
#define RTQ_BUILD_REF_PADDED_BOUNDS 0x1u  /* scene bounds also cover the origin when
                                             n_tris < 65,536, as the reference does by
                                             scanning its zero-padded triangle vector
                                             (bvh.cpp:237, scene.cpp:26-40) */
#define RTQ_BUILD_PLOC_SPLIT 0x2u         /* PLOC passes as separate neighbour-search, flag,
                                             scan and merge launches instead of one fused
                                             launch per pass; the same clusters either way */

typedef struct rtq_build_params {
    uint32_t search_radius;   /* PLOC neighbour window radius (bvh.hpp:66: 16); 1..32 */
    uint32_t max_leaf_size;   /* triangles per traversal leaf, 1..8 (the reference
                                 node array keeps one triangle per leaf regardless) */
    uint32_t flags;           /* RTQ_BUILD_* */
    uint32_t node_width;      /* traversal layout: 2 = binary 64-B nodes (default, 0 means 2),
                                 8 = compressed 8-wide 128-B nodes with 8-bit outward-rounded
                                 child boxes; the binary layout is built in both cases */
} rtq_build_params;

void rtq_build_params_default(rtq_build_params* p);

/* GPU build: scene bounds (bvh.cpp:235-251), circumscribed cube (253-303),
 * centroids and Morton codes (305-372), stable (code, index) sort (214-232),
 * PLOC (26-210) and the pre-order flatten (scene.cpp:189-208), then the
 * traversal layout.  Replaces BVH::BVH (bvh.cpp:11-24).  Waits for work queued
 * on the device first: a rebuild rewrites the scene's arrays in place, and
 * its node arrays reuse the previous build's memory where it is large enough
 * (a scene keeps its largest node arrays until rtq_scene_release).  The
 * 8-wide layout's cost table runs on a second stream of the context beside
 * the binary layout.  Temporaries come from the context's scratch, which keeps
 * its largest size. */
rtq_status rtq_bvh_build(rtq_ctx* ctx, uint32_t scene_id, const rtq_build_params* params);

/* Refit: the built tree's topology kept (PLOC merges, leaf order, binary
 * traversal node indices), every box recomputed bottom-up for the scene's
 * current vertices (leaf = triangle box, bvh.cpp:383-399; parent = merge of its
 * children, bvh.cpp:401-413), then the traversal nodes and triangles
 * rewritten in their buffers, so batch graphs captured before stay valid (the
 * reference node array is rewritten on the next export).  The 8-wide nodes keep their children (node count, depth,
 * child refs) and get boxes quantised from the new cluster boxes.  Learned
 * order words and tables are cleared.  Hit records stay exact (they never
 * depend on the tree); traversal cost grows as the vertices move away from
 * the ones the tree was built for. */
rtq_status rtq_bvh_refit(rtq_ctx* ctx, uint32_t scene_id);

typedef struct rtq_bvh_info {
    uint32_t n_tris;
    uint32_t n_ref_nodes;        /* 2N - 1 */
    uint32_t n_trav_nodes;       /* traversal (two-child) nodes */
    uint32_t depth;              /* max leaf depth of the binary tree (root = 0) */
    uint32_t ploc_iterations;
    uint32_t max_leaf_size;
    float box_pad;               /* outward padding of traversal boxes (world units) */
    float build_ms;              /* device time of the last build */
    uint32_t node_width;         /* layout traced by default: 2 or 8 */
    uint32_t n_wide_nodes;       /* compressed 8-wide nodes (0 unless node_width == 8) */
    uint32_t wide_depth;         /* levels of the 8-wide tree */
    uint32_t order_modes;        /* learned child order in the nodes (rtq_order_apply):
                                    bit 0 closest-hit decisions, bit 1 any-hit decisions */
    uint64_t scratch_spill_bytes;/* temporaries of the last build that did not fit the
                                    context scratch it reserved up front and were allocated
                                    on their own (0 normally; a later build reserves them) */
} rtq_bvh_info;

rtq_status rtq_bvh_info_get(rtq_ctx* ctx, uint32_t scene_id, rtq_bvh_info* out);

/* Pre-order BVH_NodeGPU array, root at 0, left child at i+1, exactly what
 * Scene::getBVH_NodesToGPUData returns (scene.cpp:189-208).  Host buffer of
 * `capacity` nodes; *n_nodes receives 2N-1. */
rtq_status rtq_bvh_export_ref(rtq_ctx* ctx, uint32_t scene_id, rtq_bvh_node* nodes,
                              uint64_t capacity, uint64_t* n_nodes);

/* The binary traversal layout (no counterpart in the reference, whose shader
 * reads the BVH_NodeGPU array directly, raytracer.glsl:246-295): 64-B nodes,
 *   [c0.lo.x c0.hi.x c0.lo.y c0.hi.y] [c0.lo.z c0.hi.z c1.lo.x c1.hi.x]
 *   [c1.lo.y c1.hi.y c1.lo.z c1.hi.z] [ref0 ref1 order 0]
 * (boxes widened by the build's pad; ref bit 31 set = leaf with first
 * leaf-order triangle in bits 30..3 and count - 1 in bits 2..0, else a node
 * index; order = learned order word), and per node 4 u32: the leaf-order
 * ranges of its children, [r0, r1) and [r1, r2), and its depth.  Either
 * buffer may be NULL; *n_nodes receives the node count. */
rtq_status rtq_bvh_export_trav(rtq_ctx* ctx, uint32_t scene_id, void* nodes, uint32_t* ranges,
                               uint64_t capacity, uint64_t* n_nodes);

/* The compressed 8-wide nodes (no counterpart in the reference), 128 B each
 * in level order from the root: [ox oy oz meta] with meta = ex | ey << 8 |
 * ez << 16 | valid << 24, then the 8-bit child planes [qlo_x[0..7] qhi_x[0..7]]
 * for x, y and z, then the 8 child refs (leaf: bit 31 set, first leaf-order
 * triangle in bits 30..3, count - 1 in bits 2..0; node: its index) and 32 B of
 * zeros.  A child box is o + q * 2^(e - 127) per axis.  Host buffer of
 * `capacity` nodes (NULL: only *n_nodes); RTQ_ENOTBUILT without 8-wide nodes. */
rtq_status rtq_bvh_export_wide(rtq_ctx* ctx, uint32_t scene_id, void* nodes, uint64_t capacity, uint64_t* n_nodes);

/* The build's cluster arrays in the form of BVH_Params (bvh.hpp:44-63):
 * cluster c < N is the leaf for sorted position c, c >= N are merges in the
 * order they were made (ids assigned in increasing position order within a
 * PLOC pass, where the reference numbers them in OpenMP critical-section order,
 * bvh.cpp:175-179).  Missing values (leaf children, root parent) are
 * 0xFFFFFFFF.  Every pointer may be NULL; each non-NULL array holds
 * `capacity` >= 2N-1 entries, except sorted_indices / sorted_codes (N). */
rtq_status rtq_bvh_export_clusters(rtq_ctx* ctx, uint32_t scene_id, uint64_t capacity,
                                   rtq_aabb* boxes, uint32_t* triangle_id,
                                   uint32_t* left, uint32_t* right, uint32_t* parent,
                                   uint32_t* sorted_indices, uint32_t* sorted_codes);

/* Unsorted 30-bit Morton codes of the uploaded triangles (bvh.cpp:330-348),
 * host output of n_tris entries. */
rtq_status rtq_morton_codes(rtq_ctx* ctx, uint32_t scene_id, uint32_t flags, uint32_t* codes);

/* Stable LSD radix sort of (key, value) pairs by key on device memory, in
 * place; with iota values it equals std::sort on (code, index) pairs
 * (bvh.cpp:223-231).  key_bits <= 32 selects how many low key bits are sorted;
 * n < 2^30.  One launch counts the digits of every pass, then one launch per
 * 8-bit digit finds its offsets by decoupled look-back and scatters
 * (the single-pass digit binning chainedScanDigitBinning.glsl:31-32 leaves
 * empty). */
rtq_status rtq_sort_pairs(rtq_ctx* ctx, uint32_t* keys_dev, uint32_t* vals_dev, uint64_t n,
                          uint32_t key_bits, void* stream);

/* Exclusive prefix sum of u32 (wrapping) on device memory; the scan behind
 * plocPrefixScan (bvh.cpp:125-148) and the reference's digit-count scan
 * (prefixSumOfGlobalDigitCounts.glsl:24-69) when applied per group. */
rtq_status rtq_scan_exclusive_u32(rtq_ctx* ctx, const uint32_t* in_dev, uint32_t* out_dev,
                                  uint64_t n, void* stream);

/* ---- trace -------------------------------------------------------------- */

#define RTQ_TRACE_CLOSEST     0x0u  /* closest hit, ties -> lowest triangle id */
#define RTQ_TRACE_ANY         0x1u  /* any hit (occlusion), ends at the first hit found */
#define RTQ_TRACE_LEAF_REFMT  0x2u  /* leaf test = the reference's Moller-Trumbore
                                       (raytracer.glsl:102-147: back-face culled,
                                       |det| < 1e-4 rejected, vertices 1/2 swapped);
                                       default leaf test is watertight FP32 */
#define RTQ_TRACE_BINARY      0x4u  /* trace the binary layout even if the scene was
                                       built with node_width 8 */
#define RTQ_TRACE_LEARNED     0x8u  /* binary layout, child order at nodes where both
                                       children are hit taken from the order bits that
                                       rtq_order_apply stored (hits are unchanged); a
                                       mode without decisions traces as plain rays do */

/* Trace n_rays rays against the scene's BVH.  Replaces the traversal of
 * getClosestHitBVH (raytracer.glsl:246-295) dispatched by drawOneFrame
 * (application.cpp:220-246).  Device buffers, asynchronous on `stream`; one
 * kernel launch.  Trace calls of one context share its work counter, so calls
 * issued on different streams must not run concurrently. */
rtq_status rtq_trace(rtq_ctx* ctx, uint32_t scene_id, const rtq_ray* rays_dev, uint64_t n_rays,
                     uint32_t flags, rtq_hit* hits_dev, void* stream);

/* Same trace, also writing per-ray traversal counters (node fetches, triangle
 * tests) for the algorithmic-byte count.  Counter arrays may be NULL. */
rtq_status rtq_trace_stats(rtq_ctx* ctx, uint32_t scene_id, const rtq_ray* rays_dev,
                           uint64_t n_rays, uint32_t flags, rtq_hit* hits_dev,
                           uint32_t* node_visits_dev, uint32_t* tri_tests_dev, void* stream);

/* rtq_trace without any device allocation.  No counterpart in the reference:
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same rays,
 * flags, layouts and hit records as rtq_trace.  The work counters and the
 * traversal stacks' spill live in a fixed device array of
 * RTQ_NOALLOC_SCRATCH_BYTES that the library holds on each GPU until the
 * process ends (rtq_create looks up its address once per device).  The
 * persistent grid takes as many workgroups as that array holds stacks for, so
 * very deep trees trace with fewer workgroups; RTQ_ECAPACITY when not even one
 * workgroup's stacks fit.  Every call clears the counters first.  All calls on
 * one GPU share the array: calls that overlap on one GPU (any context, any
 * stream) race, and the caller must serialise them.  Asynchronous on `stream`;
 * device buffers. */
#define RTQ_NOALLOC_SCRATCH_BYTES (64u << 20)
rtq_status rtq_trace_noalloc(rtq_ctx* ctx, uint32_t scene_id, const rtq_ray* rays_dev, uint64_t n_rays,
                             uint32_t flags, rtq_hit* hits_dev, void* stream);

/* Synchronous convenience form on host arrays (copies in, traces, copies out). */
rtq_status rtq_trace_host(rtq_ctx* ctx, uint32_t scene_id, const rtq_ray* rays, uint64_t n_rays,
                          uint32_t flags, rtq_hit* hits);

/* ---- batched traversal (no counterpart in the reference: one GL dispatch
 * per frame and scene, application.cpp:220-246) ------------------------- */

/* One ray set of a batch: its scene, its device buffers and its mode
 * (RTQ_TRACE_ANY or closest hit).  The leaf test (RTQ_TRACE_LEAF_REFMT) and
 * the node layout traced must be the same for every item of a launch. */
typedef struct rtq_batch_item {
    uint32_t scene_id;
    uint32_t flags;
    const rtq_ray* rays_dev;
    rtq_hit* hits_dev;
    uint64_t n_rays;
} rtq_batch_item;

#define RTQ_MAX_BATCH_ITEMS 64

/* Up to RTQ_MAX_BATCH_ITEMS ray sets of uneven sizes over any scenes in ONE
 * persistent launch: waves take chunks of consecutive rays from a single work
 * counter over all sets, so a long set never leaves the chip idle behind a
 * short one.  Asynchronous on `stream`. */
rtq_status rtq_trace_batch(rtq_ctx* ctx, const rtq_batch_item* items, uint32_t n_items, void* stream);

/* The same launch captured once in a hipGraph and replayed: create sizes the
 * batch's own scratch (so replays allocate nothing and share nothing with
 * other calls), captures, and instantiates; launch replays it on `stream`.
 * Item buffers must stay valid for the batch's lifetime.  Replays of one batch
 * on different streams must not overlap (they share the batch's scratch).
 * The graph holds the node and triangle buffers of the builds it was captured
 * on: a launch returns RTQ_ENOTBUILT once one of its scenes was rebuilt or
 * released (create a new batch), or edited and not yet refitted (a refit keeps
 * the buffers, so the batch replays again afterwards). */
typedef struct rtq_batch rtq_batch;
rtq_status rtq_batch_create(rtq_ctx* ctx, const rtq_batch_item* items, uint32_t n_items, rtq_batch** out);
rtq_status rtq_batch_launch(rtq_batch* batch, void* stream);
void rtq_batch_destroy(rtq_batch* batch);

/* ---- launch settings and autotuner ---------------------------------------- */

typedef struct rtq_launch_params {
    uint32_t refill;       /* refill a wave's idle lanes once this many (1..64) are idle */
    uint32_t chunk;        /* consecutive rays a wave takes per atomic (1..4096) */
    uint32_t prefer_wide;  /* trace 8-wide nodes when a scene has them */
    uint32_t spec_slack;   /* speculative traversal: a wave switches from node visits to
                              leaf tests once at most this many (0..63) of its lanes
                              are still looking for their first leaf (0: none) */
    uint32_t leaf_slack;   /* ... and back to node visits once at most this many (0..63)
                              lanes still have triangles to test; those lanes resume
                              their leaf in the next leaf phase (0: none) */
    uint32_t stack_lds;    /* traversal stack entries per ray held in LDS by batched
                              traces (rtq_trace_batch, rtq_batch_*), deeper entries in
                              HBM: 0 = the layout's default; binary layouts take 8, 12
                              (default) or 14, 8-wide layouts 8, 16 (default) or 20, and
                              a value of the other family means the default */
} rtq_launch_params;

/* A new context starts with refill 32, chunk 128, prefer_wide 1, spec_slack 16,
 * leaf_slack 8, stack_lds 0. */
rtq_status rtq_launch_params_get(rtq_ctx* ctx, rtq_launch_params* out);
rtq_status rtq_launch_params_set(rtq_ctx* ctx, const rtq_launch_params* p);

typedef struct rtq_tune_result {
    rtq_launch_params best;
    float best_ms;         /* batch time with the chosen settings */
    float default_ms;      /* batch time with the settings in force before tuning */
    uint32_t n_tried;      /* candidates timed (0: taken from the tuning cache) */
    uint32_t from_cache;
} rtq_tune_result;

/* Times candidate launch settings (refill, chunk, node layout, then the
 * speculation and leaf slacks, then the LDS stack depth) on the given batch, `reps` launches each, times
 * the four fastest again in three rounds, applies the fastest to the context and records
 * it under a key of (device, scene sizes, ray counts, modes); a later call
 * with the same key reuses it without timing.  Writes the items' hit buffers.
 * The applied settings stay in force for later traces on any scene. */
rtq_status rtq_autotune(rtq_ctx* ctx, const rtq_batch_item* items, uint32_t n_items, uint32_t reps,
                        rtq_tune_result* out);
/* Versioned, checksummed file of the tuning cache. */
rtq_status rtq_tune_save(rtq_ctx* ctx, const char* path);
rtq_status rtq_tune_load(rtq_ctx* ctx, const char* path);

/* ---- ray generation (caller side of the path) --------------------------- */

/* Pinhole camera given by precomputed basis vectors: for pixel (i, j) of a
 * width x height image, ray index j*width + i,
 *   s = i/width - 0.5, t = j/height - 0.5   (no half-pixel offset, as
 *                                            raytracer.glsl:303-307)
 *   d = normalize(right*s + up*t + forward), o = eye, interval [0, +inf).
 * right/up are scaled to the image plane at unit distance along forward
 * (the reference's plane width/height over near, camera.cpp:52-66). */
typedef struct rtq_pinhole {
    float eye[3];
    float right[3];
    float up[3];
    float forward[3];
    uint32_t width;
    uint32_t height;
} rtq_pinhole;

rtq_status rtq_raygen_pinhole(rtq_ctx* ctx, const rtq_pinhole* cam, rtq_ray* rays_dev, void* stream);

/* Secondary rays from primary hits (hit point p = o + t*d, geometric normal
 * facing the incoming ray, origin p + eps*n).  A primary miss yields an empty
 * ray (tmin = 1, tmax = 0) that traces as a miss.
 *  - diffuse: one cosine-weighted direction per hit, seeded counter-based RNG;
 *  - shadow:  n_lights rays per hit towards point lights (host array xyz),
 *             d = light - origin, interval [0, 1]; ray index = light*n + i. */
rtq_status rtq_raygen_diffuse(rtq_ctx* ctx, uint32_t scene_id, const rtq_ray* prim_dev,
                              const rtq_hit* hits_dev, uint64_t n, uint64_t seed, float eps,
                              rtq_ray* out_dev, void* stream);
rtq_status rtq_raygen_shadow(rtq_ctx* ctx, uint32_t scene_id, const rtq_ray* prim_dev,
                             const rtq_hit* hits_dev, uint64_t n, const float* lights,
                             uint32_t n_lights, float eps, rtq_ray* out_dev, void* stream);

/* ---- shading into an image (caller side of the path) -------------------- */

/* Shading flags. */
#define RTQ_SHADE_WIREFRAME 0x1u   /* uIsWireframeModeOn (raytracer.glsl:170-178) */

/* getColor + main's pixel value (raytracer.glsl:159-179, 299-331), one RGBA32F
 * pixel per hit record:
 *   c = (0, 0, 0, 1), or overlay[i] when overlay_dev is not NULL (the BVH view,
 *       uIsBVHDisplayed: c = bvhColor);
 *   a hit adds the colour of material models[tris[id].model_id].material_id
 *       (nothing for an id >= n_materials, as a read of the zero-filled
 *       material storage past the used entries);
 *   with RTQ_SHADE_WIREFRAME a hit with a barycentric weight below 0.02 is
 *       (0, 0, 0, 1).  The weights are (u, v, (1 - u) - v) of the hit record,
 *       the shader's (b0, b1, b2) for RefMT hits.
 * The reference's pixel keeps alpha 1 + material alpha (its += on the whole
 * vec4); so does this.  Device buffers of n records / pixels. */
rtq_status rtq_shade(rtq_ctx* ctx, uint32_t scene_id, const rtq_hit* hits_dev, uint64_t n,
                     const rtq_material* materials_dev, uint32_t n_materials, const float* overlay_dev,
                     uint32_t flags, float* rgba_dev, void* stream);

/* bvhColor of getClosestHitBVH (raytracer.glsl:246-276) for each ray: the
 * reference's DFS over the pre-order node array (left child pushed first, so
 * the right child is visited first) with its slab test (intersectBVH,
 * raytracer.glsl:182-237, GLSL min/max, 1/d per node, no ray interval); every
 * node at `depth` that the ray reaches sets the colour, the last one wins:
 * (0.7, 0, 0.7, 0.1) when the entry point lies within 0.05 / (depth + 1) of
 * two of the box's slab planes, else (0.5, 0, 0.5, 0.1); (0, 0, 0, 0) when no
 * node at `depth` is reached.  Nodes below `depth` cannot change the colour and
 * are not visited.  rgba_dev: n float4.  A depth above 62 on a tree that deep
 * returns RTQ_ECAPACITY. */
rtq_status rtq_bvh_overlay(rtq_ctx* ctx, uint32_t scene_id, const rtq_ray* rays_dev, uint64_t n, uint32_t depth,
                           float* rgba_dev, void* stream);

/* ---- learned child ordering (no counterpart in the reference: its
 * traversal pushes left then right, so it always enters the right child
 * first, raytracer.glsl:283-289) ----
 *
 * A 32 -> 64 -> 64 -> 1 MLP (ReLU, ReLU, logistic) scores a binary node for a
 * ray octant (bit k set <=> direction component k < 0) and a mode (closest /
 * any hit): the probability that entering child 0 first is the better order.
 * Parameters: RTQ_ORDER_PARAMS floats, W1 [32][64] (input-major), b1 [64],
 * W2 [64][64], b2 [64], W3 [64], b3.  A sample is RTQ_ORDER_SAMPLE floats:
 * 32 features, the label (1 = child 0 first) and a loss weight. */
#define RTQ_ORDER_INPUTS 32
#define RTQ_ORDER_HIDDEN 64
#define RTQ_ORDER_PARAMS 6337
#define RTQ_ORDER_SAMPLE 34

/* Traversal logs from every `stride`-th ray (closest hit, or any hit with
 * RTQ_TRACE_ANY).  From the root, each node where both children are hit is
 * priced in both visit orders by sub-traversals (nearer child first inside
 * them; node visits + triangle tests until the closest hit is certain, or
 * until an occluder is found); the sample's label is the cheaper child first
 * and its weight log2(1 + cost difference) (ties give no sample).  The walk
 * continues into the cheaper child and stops where no hit lies below it.
 * At most `capacity` samples are written (in no particular order); *n_samples
 * receives the count written.  Synchronous. */
rtq_status rtq_order_log(rtq_ctx* ctx, uint32_t scene_id, const rtq_ray* rays_dev, uint64_t n_rays,
                         uint32_t flags, uint32_t stride, float* samples_dev, uint64_t capacity,
                         uint64_t* n_samples);

/* He-uniform initial parameters from `seed` (Adam moments are the caller's
 * zeroed buffers). */
rtq_status rtq_order_init(rtq_ctx* ctx, uint64_t seed, float* params_dev);

/* Gradient of the mean weighted binary cross-entropy over samples [0, n)
 * (FP32, on the GPU); *loss (may be NULL) receives the mean loss.  Synchronous. */
rtq_status rtq_order_grad(rtq_ctx* ctx, const float* params_dev, const float* samples_dev, uint32_t n,
                          float* grad_dev, float* loss);

typedef struct rtq_train_params {
    uint32_t epochs;
    uint32_t batch;        /* samples per Adam step: a multiple of 64, at most 65536 */
    float lr;
    float beta1;
    float beta2;
    uint32_t reserved;
    uint64_t seed;         /* per-epoch sample permutation */
} rtq_train_params;

/* Adam on the GPU over `epochs` passes of the samples; *step counts Adam steps
 * (bias correction) across calls; epoch_loss (may be NULL) receives the mean
 * loss of each epoch, measured during its steps.  Synchronous. */
rtq_status rtq_order_train(rtq_ctx* ctx, const float* samples_dev, uint64_t n_samples,
                           const rtq_train_params* tp, float* params_dev, float* adam_m_dev,
                           float* adam_v_dev, uint32_t* step, float* epoch_loss);

/* Inference with BF16 MFMA tiles (FP32 accumulation): probabilities of n rows
 * of 32 features, `stride` floats apart.  Synchronous. */
rtq_status rtq_order_predict(rtq_ctx* ctx, const float* params_dev, const float* features_dev, uint64_t n,
                             uint32_t stride, float* prob_dev);

/* Scores every binary node of the scene for the 8 octants x 2 modes with the
 * MFMA inference and stores the decisions in the node; RTQ_TRACE_LEARNED then
 * takes the predicted child first where max(p, 1 - p) reaches the mode's
 * confidence (>= 0.5; >= 1 turns the mode's decisions off) and the nearer
 * child elsewhere (also where p is NaN).  params_dev NULL clears the
 * decisions (RTQ_TRACE_LEARNED is then refused).  Rebuilding the BVH clears
 * them too.  Synchronous. */
rtq_status rtq_order_apply(rtq_ctx* ctx, uint32_t scene_id, const float* params_dev, float confidence_closest,
                           float confidence_any);

/* Checkpoint: parameters, Adam moments and step in a versioned, checksummed
 * file (magic "RTOM", version 1). */
rtq_status rtq_order_save(rtq_ctx* ctx, const char* path, const float* params_dev, const float* adam_m_dev,
                          const float* adam_v_dev, uint32_t step);
rtq_status rtq_order_load(rtq_ctx* ctx, const char* path, float* params_dev, float* adam_m_dev,
                          float* adam_v_dev, uint32_t* step);

/* The same as one object that owns its device memory (parameters, Adam
 * moments and step, and a store of up to sample_capacity samples), so a caller
 * needs no device allocation of its own.  Destroy a model before its context;
 * calls on one model must not overlap. */
typedef struct rtq_order_model rtq_order_model;
rtq_status rtq_order_model_create(rtq_ctx* ctx, uint64_t seed, uint64_t sample_capacity, rtq_order_model** out);
void rtq_order_model_destroy(rtq_order_model* model);
/* rtq_order_log into the model's store, after the samples already there; a
 * full store keeps what fits.  *n_added (may be NULL) receives the count. */
rtq_status rtq_order_model_log(rtq_order_model* model, uint32_t scene_id, const rtq_ray* rays_dev,
                               uint64_t n_rays, uint32_t flags, uint32_t stride, uint64_t* n_added);
rtq_status rtq_order_model_clear_samples(rtq_order_model* model);
/* rtq_order_train on the stored samples (epoch_loss may be NULL). */
rtq_status rtq_order_model_train(rtq_order_model* model, const rtq_train_params* tp, float* epoch_loss);
rtq_status rtq_order_model_apply(rtq_order_model* model, uint32_t scene_id, float confidence_closest,
                                 float confidence_any);
rtq_status rtq_order_model_save(rtq_order_model* model, const char* path);
rtq_status rtq_order_model_load(rtq_order_model* model, const char* path);
/* Stored samples and Adam steps taken (either pointer may be NULL). */
rtq_status rtq_order_model_info(const rtq_order_model* model, uint64_t* n_samples, uint32_t* step);

#if defined(__GNUC__) || defined(__clang__)
#pragma GCC visibility pop
#endif

#ifdef __cplusplus
}  /* extern "C" */
#endif

#endif /* RTQ_RTQ_H */
