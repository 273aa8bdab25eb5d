/*
 * ORACLE — test infrastructure only (see oracle.h).  CPU restatement of the
 * reference hot path; compiled with -ffp-contract=off so every float
 * expression rounds exactly as written.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define NONE 0xFFFFFFFFu

/* std::min / std::max semantics (bvh.cpp uses them throughout). */
static inline float smin(float a, float b) { return (b < a) ? b : a; }
static inline float smax(float a, float b) { return (a < b) ? b : a; }
/* GLSL min/max as specified: min(x,y) = y < x ? y : x, max(x,y) = x < y ? y : x. */
static inline float gmin(float a, float b) { return (b < a) ? b : a; }
static inline float gmax(float a, float b) { return (a < b) ? b : a; }

/* ------------------------------------------------------------------------ */
/* glm arithmetic used at the call sites of bvh.cpp / triangle.cpp           */

void orc_mat4_mul_vec4(const float m[16], const float v[4], float out[4]) {
    /* glm's mat4*vec4: (m0*v.x + m1*v.y) + (m2*v.z + m3*v.w), column-wise. */
    for (int r = 0; r < 4; ++r) {
        float mul0 = m[0 * 4 + r] * v[0];
        float mul1 = m[1 * 4 + r] * v[1];
        float add0 = mul0 + mul1;
        float mul2 = m[2 * 4 + r] * v[2];
        float mul3 = m[3 * 4 + r] * v[3];
        float add1 = mul2 + mul3;
        out[r] = add0 + add1;
    }
}

static void xform(const rtq_mesh_model* model, const rtq_vec4* p, float out[4]) {
    float v[4] = {p->x, p->y, p->z, p->w};
    orc_mat4_mul_vec4(model->m, v, out);
}

static inline float canon0(float x) { return x == 0.0f ? 0.0f : x; }

rtq_aabb orc_scene_aabb(const rtq_triangle* tris, uint32_t n, const rtq_mesh_model* models,
                        int pad_origin) {
    /* bvh.cpp:235-251 */
    rtq_aabb b;
    memset(&b, 0, sizeof b);
    for (int k = 0; k < 3; ++k) { b.min[k] = INFINITY; b.max[k] = -INFINITY; }
    for (uint32_t i = 0; i < n; ++i) {
        const rtq_triangle* t = &tris[i];
        float p0[4], p1[4], p2[4];
        xform(&models[t->model_id], &t->p0, p0);
        xform(&models[t->model_id], &t->p1, p1);
        xform(&models[t->model_id], &t->p2, p2);
        for (int k = 0; k < 3; ++k) {
            b.max[k] = smax(b.max[k], smax(p0[k], smax(p1[k], p2[k])));
            b.min[k] = smin(b.min[k], smin(p0[k], smin(p1[k], p2[k])));
        }
    }
    if (pad_origin && n < 65536u) {
        /* a zero TriangleGPU transforms to the origin whatever the matrix */
        for (int k = 0; k < 3; ++k) {
            b.max[k] = smax(b.max[k], 0.0f);
            b.min[k] = smin(b.min[k], 0.0f);
        }
    }
    for (int k = 0; k < 3; ++k) { b.min[k] = canon0(b.min[k]); b.max[k] = canon0(b.max[k]); }
    return b;
}

rtq_aabb orc_circumscribed_cube(rtq_aabb s) {
    /* bvh.cpp:253-303, axis choice as written at 263-272 */
    rtq_aabb c = s;
    float dist_x = s.max[0] - s.min[0];
    float dist_y = s.max[1] - s.min[1];
    float dist_z = s.max[2] - s.min[2];
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
    return c;
}

void orc_normalized_centroid(const rtq_triangle* t, const rtq_mesh_model* model,
                             const rtq_aabb* cube, float out[3]) {
    /* triangle.cpp:31-33: (1.f/3.f) * model * (P0 + P1 + P2) parses as
     * ((1/3 * model) * ((P0 + P1) + P2)). */
    const float third = 1.0f / 3.0f;
    float ms[16];
    for (int i = 0; i < 16; ++i) ms[i] = third * model->m[i];
    float s[4] = {(t->p0.x + t->p1.x) + t->p2.x, (t->p0.y + t->p1.y) + t->p2.y,
                  (t->p0.z + t->p1.z) + t->p2.z, (t->p0.w + t->p1.w) + t->p2.w};
    float c[4];
    orc_mat4_mul_vec4(ms, s, c);
    /* bvh.cpp:314-328 */
    for (int k = 0; k < 3; ++k) {
        float len = cube->max[k] - cube->min[k];
        out[k] = (c[k] - cube->min[k]) / len;
    }
}

uint32_t orc_expand_bits(uint32_t v) {
    /* bvh.cpp:350-356 */
    v = (v * 0x00010001u) & 0xFF0000FFu;
    v = (v * 0x00000101u) & 0x0F00F00Fu;
    v = (v * 0x00000011u) & 0xC30C30C3u;
    v = (v * 0x00000005u) & 0x49249249u;
    return v;
}

static uint32_t quantize10(float x) {
    /* bvh.cpp:363-367; a NaN (zero-extent axis) maps to 0 where the reference
     * cast is undefined. */
    x = smin(smax(x * 1024.0f, 0.0f), 1023.0f);
    if (x != x) return 0u;
    return (uint32_t)x;
}

uint32_t orc_morton3d(float x, float y, float z) {
    /* bvh.cpp:358-372 */
    uint32_t xx = orc_expand_bits(quantize10(x));
    uint32_t yy = orc_expand_bits(quantize10(y));
    uint32_t zz = orc_expand_bits(quantize10(z));
    return (xx << 2) | (yy << 1) | zz;
}

void orc_morton_codes(const rtq_triangle* tris, uint32_t n, const rtq_mesh_model* models,
                      int pad_origin, uint32_t* codes) {
    /* bvh.cpp:330-348 */
    rtq_aabb scene = orc_scene_aabb(tris, n, models, pad_origin);
    rtq_aabb cube = orc_circumscribed_cube(scene);
    for (uint32_t i = 0; i < n; ++i) {
        float c[3];
        orc_normalized_centroid(&tris[i], &models[tris[i].model_id], &cube, c);
        codes[i] = orc_morton3d(c[0], c[1], c[2]);
    }
}

static int cmp_pair(const void* a, const void* b) {
    const uint64_t x = *(const uint64_t*)a, y = *(const uint64_t*)b;
    return (x > y) - (x < y);
}

void orc_sort_pairs(uint32_t* codes, uint32_t* idx, uint32_t n) {
    /* bvh.cpp:223-231: std::sort of (code, index) pairs, lexicographic */
    uint64_t* p = (uint64_t*)malloc(sizeof(uint64_t) * (n ? n : 1));
    for (uint32_t i = 0; i < n; ++i) p[i] = ((uint64_t)codes[i] << 32) | idx[i];
    qsort(p, n, sizeof(uint64_t), cmp_pair);
    for (uint32_t i = 0; i < n; ++i) { codes[i] = (uint32_t)(p[i] >> 32); idx[i] = (uint32_t)p[i]; }
    free(p);
}

rtq_aabb orc_triangle_box(const rtq_triangle* t, const rtq_mesh_model* model) {
    /* bvh.cpp:383-399 */
    float p0[4], p1[4], p2[4];
    xform(model, &t->p0, p0);
    xform(model, &t->p1, p1);
    xform(model, &t->p2, p2);
    rtq_aabb b;
    memset(&b, 0, sizeof b);
    for (int k = 0; k < 3; ++k) {
        b.min[k] = smin(p0[k], smin(p1[k], p2[k]));
        b.max[k] = smax(p0[k], smax(p1[k], p2[k]));
    }
    return b;
}

float orc_surface_area(const rtq_aabb* b) {
    /* bvh.cpp:378-381 */
    float dx = b->max[0] - b->min[0];
    float dy = b->max[1] - b->min[1];
    float dz = b->max[2] - b->min[2];
    return 2.0f * ((dx * dy + dy * dz) + dz * dx);
}

rtq_aabb orc_merge(const rtq_aabb* a, const rtq_aabb* b) {
    /* bvh.cpp:401-413 */
    rtq_aabb m;
    memset(&m, 0, sizeof m);
    for (int k = 0; k < 3; ++k) {
        m.min[k] = smin(a->min[k], b->min[k]);
        m.max[k] = smax(a->max[k], b->max[k]);
    }
    return m;
}

int orc_ploc_build(const rtq_triangle* tris, uint32_t n, const rtq_mesh_model* models,
                   int pad_origin, uint32_t radius,
                   rtq_aabb* boxes, uint32_t* tri_id, uint32_t* left, uint32_t* right,
                   uint32_t* parent, uint32_t* leaf_count,
                   uint32_t* sorted_idx, uint32_t* sorted_codes) {
    if (n == 0) return -1;
    const uint32_t total = 2 * n - 1;
    for (uint32_t c = 0; c < total; ++c) {
        memset(&boxes[c], 0, sizeof(rtq_aabb));
        tri_id[c] = 0; left[c] = NONE; right[c] = NONE; parent[c] = NONE; leaf_count[c] = 0;
    }
    /* plocPreprocessing, bvh.cpp:26-46 */
    orc_morton_codes(tris, n, models, pad_origin, sorted_codes);
    for (uint32_t i = 0; i < n; ++i) sorted_idx[i] = i;
    orc_sort_pairs(sorted_codes, sorted_idx, n);
    uint32_t* c_in = (uint32_t*)malloc(sizeof(uint32_t) * n);
    uint32_t* c_out = (uint32_t*)malloc(sizeof(uint32_t) * n);
    uint32_t* nn = (uint32_t*)calloc(n, sizeof(uint32_t));
    uint32_t* scan = (uint32_t*)malloc(sizeof(uint32_t) * n);
    for (uint32_t i = 0; i < n; ++i) {
        uint32_t ti = sorted_idx[i];
        boxes[i] = orc_triangle_box(&tris[ti], &models[tris[ti].model_id]);
        tri_id[i] = ti;
        leaf_count[i] = 1;
        c_in[i] = i;
    }
    uint32_t iter = n, n_total = n;
    int passes = 0;
    while (iter > 1) {
        /* plocNearestNeighborSearch, bvh.cpp:193-210, positions in parallel as
         * the reference's OpenMP loop runs them (each result is independent) */
#pragma omp parallel for schedule(static)
        for (uint32_t i = 0; i < iter; ++i) {
            float min_dist = INFINITY;
            const rtq_aabb* bi = &boxes[c_in[i]];
            uint32_t start = (i > radius) ? i - radius : 0;
            uint32_t end = (i + radius + 1 < iter) ? i + radius + 1 : iter;
            for (uint32_t j = start; j < end; ++j) {
                if (j == i) continue;
                rtq_aabb m = orc_merge(bi, &boxes[c_in[j]]);
                float d = orc_surface_area(&m);
                if (d < min_dist) { min_dist = d; nn[i] = j; }
            }
        }
        /* plocMerging, bvh.cpp:159-191; new ids in increasing i */
        uint32_t merges = 0;
        for (uint32_t i = 0; i < iter; ++i) {
            uint32_t j = nn[i];
            if (j < iter && nn[j] == i && i < j) {
                uint32_t ci = c_in[i], cj = c_in[j];
                uint32_t id = n_total++;
                boxes[id] = orc_merge(&boxes[ci], &boxes[cj]);
                tri_id[id] = 0;
                left[id] = ci; right[id] = cj;
                parent[ci] = id; parent[cj] = id;
                leaf_count[id] = leaf_count[ci] + leaf_count[cj];
                c_in[j] = NONE;
                c_in[i] = id;
                ++merges;
            }
        }
        ++passes;
        if (merges == 0) { passes = -1; break; }
        /* plocPrefixScan + plocCompaction, bvh.cpp:125-156, iteration update 105-113 */
        uint32_t acc = 0;
        for (uint32_t i = 0; i < iter; ++i) { scan[i] = acc; acc += (c_in[i] != NONE); }
        for (uint32_t i = 0; i < iter; ++i)
            if (c_in[i] != NONE) c_out[scan[i]] = c_in[i];
        iter = acc;
        uint32_t* tmp = c_in; c_in = c_out; c_out = tmp;
    }
    free(c_in); free(c_out); free(nn); free(scan);
    return passes;
}

int orc_flatten_preorder(uint32_t n, const rtq_aabb* boxes, const uint32_t* tri_id,
                         const uint32_t* left, const uint32_t* right, rtq_bvh_node* out) {
    /* scene.cpp:189-208: recursiveTopDownTraversalBVH from root 2N-2, written
     * with an explicit stack; a node is appended, then its left subtree, then
     * its right subtree. */
    if (n == 0) return -1;
    const uint32_t total = 2 * n - 1;
    uint32_t root = 2 * n - 2;
    /* stack of (cluster, parent position, is_right) */
    uint32_t* st_c = (uint32_t*)malloc(sizeof(uint32_t) * (total + 1));
    uint32_t* st_p = (uint32_t*)malloc(sizeof(uint32_t) * (total + 1));
    uint8_t* st_r = (uint8_t*)malloc(total + 1);
    uint32_t sp = 0, pos = 0;
    st_c[sp] = root; st_p[sp] = NONE; st_r[sp] = 0; ++sp;
    while (sp) {
        --sp;
        uint32_t c = st_c[sp], pp = st_p[sp];
        uint8_t isr = st_r[sp];
        uint32_t here = pos++;
        memset(&out[here], 0, sizeof(rtq_bvh_node));
        out[here].box = boxes[c];
        out[here].box.pad0_ = 0.0f; out[here].box.pad1_ = 0.0f;
        out[here].triangle_id = tri_id[c];
        if (pp != NONE) {
            if (isr) out[pp].right_child = here; else out[pp].left_child = here;
        }
        if (left[c] != NONE) {
            /* right pushed first so the left subtree is emitted first */
            st_c[sp] = right[c]; st_p[sp] = here; st_r[sp] = 1; ++sp;
            st_c[sp] = left[c]; st_p[sp] = here; st_r[sp] = 0; ++sp;
        }
    }
    free(st_c); free(st_p); free(st_r);
    return pos == total ? 0 : -1;
}

int orc_build_ref_nodes(const rtq_triangle* tris, uint32_t n, const rtq_mesh_model* models,
                        int pad_origin, uint32_t radius, rtq_bvh_node* out) {
    if (n == 0) return -1;
    uint32_t total = 2 * n - 1;
    rtq_aabb* boxes = (rtq_aabb*)malloc(sizeof(rtq_aabb) * total);
    uint32_t* a = (uint32_t*)malloc(sizeof(uint32_t) * total * 5 + sizeof(uint32_t) * 2 * n);
    uint32_t *tri_id = a, *left = a + total, *right = a + 2 * total, *parent = a + 3 * total,
             *lc = a + 4 * total, *sidx = a + 5 * total, *scode = a + 5 * total + n;
    int passes = orc_ploc_build(tris, n, models, pad_origin, radius, boxes, tri_id, left, right,
                                parent, lc, sidx, scode);
    if (passes >= 0 && orc_flatten_preorder(n, boxes, tri_id, left, right, out) != 0) passes = -1;
    free(boxes); free(a);
    return passes;
}

uint32_t orc_tree_depth(const rtq_bvh_node* nodes, uint32_t n_nodes) {
    if (n_nodes == 0) return 0;
    uint32_t* depth = (uint32_t*)calloc(n_nodes, sizeof(uint32_t));
    uint32_t best = 0;
    for (uint32_t i = 0; i < n_nodes; ++i) {   /* pre-order: parents come first */
        const rtq_bvh_node* nd = &nodes[i];
        if (nd->left_child == 0 && nd->right_child == 0) {
            if (depth[i] > best) best = depth[i];
        } else {
            depth[nd->left_child] = depth[i] + 1;
            depth[nd->right_child] = depth[i] + 1;
        }
    }
    free(depth);
    return best;
}

void orc_world_vertices(const rtq_triangle* tris, uint32_t n, const rtq_mesh_model* models,
                        float* v9) {
    for (uint32_t i = 0; i < n; ++i) {
        float p[4];
        const rtq_mesh_model* m = &models[tris[i].model_id];
        xform(m, &tris[i].p0, p); v9[9 * i + 0] = p[0]; v9[9 * i + 1] = p[1]; v9[9 * i + 2] = p[2];
        xform(m, &tris[i].p1, p); v9[9 * i + 3] = p[0]; v9[9 * i + 4] = p[1]; v9[9 * i + 5] = p[2];
        xform(m, &tris[i].p2, p); v9[9 * i + 6] = p[0]; v9[9 * i + 7] = p[1]; v9[9 * i + 8] = p[2];
    }
}

/* ------------------------------------------------------------------------ */
/* Leaf tests                                                               */

static inline void cross3(const float a[3], const float b[3], float o[3]) {
    /* glm::cross: (a.y*b.z - b.y*a.z, a.z*b.x - b.z*a.x, a.x*b.y - b.x*a.y) */
    o[0] = a[1] * b[2] - b[1] * a[2];
    o[1] = a[2] * b[0] - b[2] * a[0];
    o[2] = a[0] * b[1] - b[0] * a[1];
}
static inline float dot3(const float a[3], const float b[3]) {
    return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2];
}

int orc_refmt(const rtq_ray* ray, const float* v9, float tmax, float* t_out, float* u_out,
              float* v_out) {
    /* raytracer.glsl:102-147.  The shader's struct order _P0,_P2,_P1
     * (raytracer.glsl:20-25) makes its p1 the C++ P2 and its p2 the C++ P1. */
    const float* p0 = v9;
    const float* p1 = v9 + 6;  /* shader p1 = C++ P2 */
    const float* p2 = v9 + 3;  /* shader p2 = C++ P1 */
    float e0[3] = {p1[0] - p0[0], p1[1] - p0[1], p1[2] - p0[2]};
    float e1[3] = {p2[0] - p0[0], p2[1] - p0[1], p2[2] - p0[2]};
    float nrm[3];
    cross3(e1, e0, nrm);
    float q[3];
    cross3(ray->d, e1, q);
    float a = dot3(e0, q);
    if (dot3(nrm, ray->d) >= 0.0f || fabsf(a) < 1e-4f) return 0;
    float s[3] = {(ray->o[0] - p0[0]) / a, (ray->o[1] - p0[1]) / a, (ray->o[2] - p0[2]) / a};
    float r[3];
    cross3(s, e0, r);
    float b0 = dot3(s, q);
    float b1 = dot3(r, ray->d);
    float b2 = (1.0f - b0) - b1;
    if (b0 < 0.0f || b1 < 0.0f || b2 < 0.0f) return 0;
    float t = dot3(e1, r);
    if (t < 0.0f) return 0;
    if (!(t >= ray->tmin && t <= tmax)) return 0;
    *t_out = t; *u_out = b0; *v_out = b1;
    return 1;
}

/* Ray set-up of the watertight test: dominant axis kz (first maximum of
 * |d|), kx/ky following it, swapped when d[kz] < 0, and the shear. */
typedef struct { int kx, ky, kz; float sx, sy, sz; int ok; } wt_setup;

static wt_setup wt_prepare(const float d[3]) {
    wt_setup w;
    float ax = fabsf(d[0]), ay = fabsf(d[1]), az = fabsf(d[2]);
    int kz = 0;
    float m = ax;
    if (ay > m) { kz = 1; m = ay; }
    if (az > m) { kz = 2; m = az; }
    int kx = kz + 1; if (kx == 3) kx = 0;
    int ky = kx + 1; if (ky == 3) ky = 0;
    if (d[kz] < 0.0f) { int tmp = kx; kx = ky; ky = tmp; }
    w.kx = kx; w.ky = ky; w.kz = kz;
    w.ok = (d[kz] != 0.0f);
    w.sx = d[kx] / d[kz];
    w.sy = d[ky] / d[kz];
    w.sz = 1.0f / d[kz];
    return w;
}

static int wt_test(const wt_setup* w, const rtq_ray* ray, const float* v9, float tmax,
                   float* t_out, float* u_out, float* v_out) {
    if (!w->ok) return 0;
    const int kx = w->kx, ky = w->ky, kz = w->kz;
    float A[3] = {v9[0] - ray->o[0], v9[1] - ray->o[1], v9[2] - ray->o[2]};
    float B[3] = {v9[3] - ray->o[0], v9[4] - ray->o[1], v9[5] - ray->o[2]};
    float C[3] = {v9[6] - ray->o[0], v9[7] - ray->o[1], v9[8] - ray->o[2]};
    float Ax = A[kx] - w->sx * A[kz];
    float Ay = A[ky] - w->sy * A[kz];
    float Bx = B[kx] - w->sx * B[kz];
    float By = B[ky] - w->sy * B[kz];
    float Cx = C[kx] - w->sx * C[kz];
    float Cy = C[ky] - w->sy * C[kz];
    float U = Cx * By - Cy * Bx;
    float V = Ax * Cy - Ay * Cx;
    float W = Bx * Ay - By * Ax;
    if (U == 0.0f || V == 0.0f || W == 0.0f) {
        double CxBy = (double)Cx * (double)By, CyBx = (double)Cy * (double)Bx;
        U = (float)(CxBy - CyBx);
        double AxCy = (double)Ax * (double)Cy, AyCx = (double)Ay * (double)Cx;
        V = (float)(AxCy - AyCx);
        double BxAy = (double)Bx * (double)Ay, ByAx = (double)By * (double)Ax;
        W = (float)(BxAy - ByAx);
    }
    if ((U < 0.0f || V < 0.0f || W < 0.0f) && (U > 0.0f || V > 0.0f || W > 0.0f)) return 0;
    float det = (U + V) + W;
    if (det == 0.0f) return 0;
    float Az = w->sz * A[kz];
    float Bz = w->sz * B[kz];
    float Cz = w->sz * C[kz];
    float T = (U * Az + V * Bz) + W * Cz;
    float t = T / det;
    if (!(t >= ray->tmin && t <= tmax)) return 0;
    *t_out = t;
    *u_out = V / det;
    *v_out = W / det;
    return 1;
}

int orc_watertight(const rtq_ray* ray, const float* v9, float tmax, float* t, float* u, float* v) {
    wt_setup w = wt_prepare(ray->d);
    return wt_test(&w, ray, v9, tmax, t, u, v);
}

double orc_t_f64(const rtq_ray* ray, const float* v9) {
    double e1[3], e2[3], nrm[3], w[3];
    for (int k = 0; k < 3; ++k) {
        e1[k] = (double)v9[3 + k] - (double)v9[k];
        e2[k] = (double)v9[6 + k] - (double)v9[k];
        w[k] = (double)v9[k] - (double)ray->o[k];
    }
    nrm[0] = e1[1] * e2[2] - e1[2] * e2[1];
    nrm[1] = e1[2] * e2[0] - e1[0] * e2[2];
    nrm[2] = e1[0] * e2[1] - e1[1] * e2[0];
    double num = nrm[0] * w[0] + nrm[1] * w[1] + nrm[2] * w[2];
    double den = nrm[0] * ray->d[0] + nrm[1] * ray->d[1] + nrm[2] * ray->d[2];
    return num / den;
}

/* ------------------------------------------------------------------------ */
/* Traversals                                                                */

static inline int better(float t, int32_t id, const rtq_hit* best) {
    return best->id < 0 || t < best->t || (t == best->t && id < best->id);
}

static void miss(rtq_hit* h) { h->id = -1; h->t = INFINITY; h->u = 0.0f; h->v = 0.0f; }

static int set_threads(int n_threads) {
#ifdef _OPENMP
    return n_threads > 0 ? n_threads : omp_get_max_threads();
#else
    (void)n_threads;
    return 1;
#endif
}

void orc_trace_brute(const float* v9, uint32_t n_tris, const rtq_ray* rays, uint64_t n_rays,
                     int leaf, int mode, rtq_hit* hits, int n_threads) {
    /* raytracer.glsl:149-157 (getAllHits): index order, strict t < best */
    int nt = set_threads(n_threads);
    (void)nt;
#pragma omp parallel for schedule(dynamic, 16) num_threads(nt)
    for (int64_t r = 0; r < (int64_t)n_rays; ++r) {
        const rtq_ray* ray = &rays[r];
        rtq_hit best;
        miss(&best);
        wt_setup w = wt_prepare(ray->d);
        int empty = !(ray->tmin <= ray->tmax);
        for (uint32_t i = 0; i < n_tris && !empty; ++i) {
            float t, u, v;
            int h = (leaf == ORC_LEAF_REFMT) ? orc_refmt(ray, v9 + 9 * (size_t)i, ray->tmax, &t, &u, &v)
                                             : wt_test(&w, ray, v9 + 9 * (size_t)i, ray->tmax, &t, &u, &v);
            if (!h) continue;
            if (best.id < 0 || t < best.t) {
                best.id = (int32_t)i; best.t = t; best.u = u; best.v = v;
                if (mode == ORC_MODE_ANY) break;
            }
        }
        hits[r] = best;
    }
}

/* raytracer.glsl:182-237 with GLSL min/max; returns nonzero on a hit. */
static int ref_slab(const rtq_ray* ray, const rtq_aabb* b) {
    float inv_x = 1.0f / ray->d[0];
    float tx1 = (b->min[0] - ray->o[0]) * inv_x;
    float tx2 = (b->max[0] - ray->o[0]) * inv_x;
    float tmin = gmin(tx1, tx2);
    float tmax = gmax(tx1, tx2);
    if (tmax < 0.0f || tmin > tmax) return 0;
    float inv_y = 1.0f / ray->d[1];
    float ty1 = (b->min[1] - ray->o[1]) * inv_y;
    float ty2 = (b->max[1] - ray->o[1]) * inv_y;
    tmin = gmax(tmin, gmin(ty1, ty2));
    tmax = gmin(tmax, gmax(ty1, ty2));
    if (tmax < 0.0f || tmin > tmax) return 0;
    float inv_z = 1.0f / ray->d[2];
    float tz1 = (b->min[2] - ray->o[2]) * inv_z;
    float tz2 = (b->max[2] - ray->o[2]) * inv_z;
    tmin = gmax(tmin, gmin(tz1, tz2));
    tmax = gmin(tmax, gmax(tz1, tz2));
    return (tmax >= 0.0f && tmin <= tmax);
}

void orc_trace_ref(const rtq_bvh_node* nodes, const float* v9, const rtq_ray* rays,
                   uint64_t n_rays, int leaf, rtq_hit* hits, int n_threads) {
    /* raytracer.glsl:246-295 */
    int nt = set_threads(n_threads);
    (void)nt;
#pragma omp parallel num_threads(nt)
    {
        uint32_t cap = 4096;
        uint32_t* stack = (uint32_t*)malloc(sizeof(uint32_t) * cap);
#pragma omp for schedule(dynamic, 64)
        for (int64_t r = 0; r < (int64_t)n_rays; ++r) {
            const rtq_ray* ray = &rays[r];
            rtq_hit best;
            miss(&best);
            wt_setup w = wt_prepare(ray->d);
            uint32_t sp = 0;
            if (ray->tmin <= ray->tmax) stack[sp++] = 0;
            while (sp > 0) {
                uint32_t cur = stack[--sp];
                const rtq_bvh_node* nd = &nodes[cur];
                if (!ref_slab(ray, &nd->box)) continue;
                if (nd->left_child == 0 && nd->right_child == 0) {
                    float t, u, v;
                    const float* tri = v9 + 9 * (size_t)nd->triangle_id;
                    int h = (leaf == ORC_LEAF_REFMT) ? orc_refmt(ray, tri, ray->tmax, &t, &u, &v)
                                                     : wt_test(&w, ray, tri, ray->tmax, &t, &u, &v);
                    if (h && better(t, (int32_t)nd->triangle_id, &best)) {
                        best.id = (int32_t)nd->triangle_id; best.t = t; best.u = u; best.v = v;
                    }
                } else {
                    if (sp + 2 > cap) {
                        cap *= 2;
                        stack = (uint32_t*)realloc(stack, sizeof(uint32_t) * cap);
                    }
                    stack[sp++] = nd->left_child;
                    stack[sp++] = nd->right_child;
                }
            }
            hits[r] = best;
        }
        free(stack);
    }
}

/* Relative widening of a far distance: a box is entered while
 * tn <= tf + |tf| * ORC_SLAB_REL.  Errors proportional to the distance (the
 * reciprocal, the subtraction of the origin and the product here; the
 * watertight test's translation of far vertices and its t = T / det) stay
 * below a sixteenth of it, so with boxes padded by `pad` (errors proportional
 * to the scene's coordinates) the test keeps every watertight hit for any ray
 * origin (Ize, "Robust BVH Ray Traversal", JCGT 2013, widens the far distance
 * the same way).  The same constant widens the GPU kernels' far distances. */
#define ORC_SLAB_REL 0x1p-19f

static int slab_keep(float tn, float tf) { return tn <= fmaf(fabsf(tf), ORC_SLAB_REL, tf); }

/* Conservative slab test: boxes widened by `pad`, far distance widened
 * relatively (slab_keep), inclusive comparison. */
static int cons_slab(const float o[3], const float inv[3], const rtq_aabb* b, float pad,
                     float t0, float t1, float* tnear) {
    float lo, hi, tn = t0, tf = t1;
    for (int k = 0; k < 3; ++k) {
        float a = ((b->min[k] - pad) - o[k]) * inv[k];
        float c = ((b->max[k] + pad) - o[k]) * inv[k];
        lo = smin(a, c);
        hi = smax(a, c);
        tn = smax(tn, lo);
        tf = smin(tf, hi);
    }
    *tnear = tn;
    return slab_keep(tn, tf);
}

void orc_trace_bvh(const rtq_bvh_node* nodes, const float* v9, const rtq_ray* rays,
                   uint64_t n_rays, int leaf, int mode, float pad, rtq_hit* hits, int n_threads,
                   uint64_t* visits_out, uint64_t* tests_out) {
    uint32_t n_nodes_guess = 0;
    (void)n_nodes_guess;
    int nt = set_threads(n_threads);
    (void)nt;
    uint64_t visits = 0, tests = 0;
#pragma omp parallel num_threads(nt) reduction(+ : visits, tests)
    {
        uint32_t cap = 256;
        uint32_t* st_n = (uint32_t*)malloc(sizeof(uint32_t) * cap);
        float* st_t = (float*)malloc(sizeof(float) * cap);
#pragma omp for schedule(dynamic, 64)
        for (int64_t r = 0; r < (int64_t)n_rays; ++r) {
            const rtq_ray* ray = &rays[r];
            rtq_hit best;
            miss(&best);
            float tbest = ray->tmax;
            if (!(ray->tmin <= ray->tmax)) { hits[r] = best; continue; }
            wt_setup w = wt_prepare(ray->d);
            float inv[3];
            for (int k = 0; k < 3; ++k) {
                float dk = ray->d[k];
                if (fabsf(dk) < 1e-30f) dk = copysignf(1e-30f, dk);
                inv[k] = 1.0f / dk;
            }
            uint32_t sp = 0;
            uint32_t cur = 0;
            float tn;
            int done = 0;
            if (!cons_slab(ray->o, inv, &nodes[0].box, pad, ray->tmin, tbest, &tn)) done = 1;
            while (!done) {
                const rtq_bvh_node* nd = &nodes[cur];
                ++visits;
                if (nd->left_child == 0 && nd->right_child == 0) {
                    float t, u, v;
                    const float* tri = v9 + 9 * (size_t)nd->triangle_id;
                    ++tests;
                    int h = (leaf == ORC_LEAF_REFMT) ? orc_refmt(ray, tri, tbest, &t, &u, &v)
                                                     : wt_test(&w, ray, tri, tbest, &t, &u, &v);
                    if (h && better(t, (int32_t)nd->triangle_id, &best)) {
                        best.id = (int32_t)nd->triangle_id; best.t = t; best.u = u; best.v = v;
                        tbest = t;
                        if (mode == ORC_MODE_ANY) break;
                    }
                } else {
                    float ta, tb;
                    int ha = cons_slab(ray->o, inv, &nodes[nd->left_child].box, pad, ray->tmin, tbest, &ta);
                    int hb = cons_slab(ray->o, inv, &nodes[nd->right_child].box, pad, ray->tmin, tbest, &tb);
                    if (ha && hb) {
                        uint32_t nearc = nd->left_child, farc = nd->right_child;
                        float tfar = tb;
                        if (tb < ta) { nearc = nd->right_child; farc = nd->left_child; tfar = ta; }
                        if (sp == cap) {
                            cap *= 2;
                            st_n = (uint32_t*)realloc(st_n, sizeof(uint32_t) * cap);
                            st_t = (float*)realloc(st_t, sizeof(float) * cap);
                        }
                        st_n[sp] = farc; st_t[sp] = tfar; ++sp;
                        cur = nearc;
                        continue;
                    }
                    if (ha) { cur = nd->left_child; continue; }
                    if (hb) { cur = nd->right_child; continue; }
                }
                /* pop, skipping entries whose entry distance is beyond the best hit */
                done = 1;
                while (sp > 0) {
                    --sp;
                    if (slab_keep(st_t[sp], tbest)) { cur = st_n[sp]; done = 0; break; }
                }
            }
            hits[r] = best;
        }
        free(st_n); free(st_t);
    }
    if (visits_out) *visits_out = visits;
    if (tests_out) *tests_out = tests;
}

/* ------------------------------------------------------------------------ */
/* The GPU kernels' box tests, restated to check them against brute force    */

static float gpu_rcp(float d, int var) {
    /* v_rcp_f32 is within 1 ulp of 1/d: nearest, or one ulp up or down */
    float r = 1.0f / d;
    if (var == 1) r = nextafterf(r, INFINITY);
    if (var == 2) r = nextafterf(r, -INFINITY);
    return r;
}

/* trace.hip, binary visit: plane distance fma(b, inv, -(o * inv)) on boxes
 * padded by `pad`, entered while tn <= tf + |tf| * rel. */
static int gpu_box_binary(const rtq_ray* r, int var, const rtq_aabb* b, float pad, float tbest, float rel) {
    float tn = r->tmin, tf = tbest;
    for (int k = 0; k < 3; ++k) {
        float e = fabsf(r->d[k]) < 1e-30f ? copysignf(1e-30f, r->d[k]) : r->d[k];
        float inv = gpu_rcp(e, var);
        float oi = r->o[k] * inv;
        float a = fmaf(b->min[k] - pad, inv, -oi), c = fmaf(b->max[k] + pad, inv, -oi);
        tn = fmaxf(tn, fminf(a, c));
        tf = fminf(tf, fmaxf(a, c));
    }
    return tn <= fmaf(fabsf(tf), rel, tf);
}

static int wide_exponent(double extent) {
    /* wide.hip exponent_for: smallest e with 255 * 2^e >= extent */
    int e = extent > 0.0 ? ilogb(extent) - 7 : -126;
    if (e < -126) e = -126;
    while (e < 127 && ldexp(255.0, e) < extent) ++e;
    while (e > -126 && ldexp(255.0, e - 1) >= extent) --e;
    return e;
}

/* trace.hip / wide.hip, 8-wide visit: the child box quantised outward to 8 bits
 * in a node frame (here the parent box's padded frame stands in for the 8-wide
 * node's), planes fma(q, 2^e * inv, fma(origin, inv, -(o * inv))), near and far
 * plane picked by the sign of inv. */
static int gpu_box_wide(const rtq_ray* r, int var, const rtq_aabb* b, const rtq_aabb* frame, float pad, float tbest,
                        float rel) {
    float tn = r->tmin, tf = tbest;
    for (int k = 0; k < 3; ++k) {
        float org = frame->min[k] - pad;
        int ex = wide_exponent((double)(frame->max[k] + pad) - (double)org);
        double scale = ldexp(1.0, ex);
        double l = floor(((double)(b->min[k] - pad) - (double)org) / scale);
        double h = ceil(((double)(b->max[k] + pad) - (double)org) / scale);
        l = fmin(fmax(l, 0.0), 255.0);
        h = fmin(fmax(h, 0.0), 255.0);
        float e = fabsf(r->d[k]) < 1e-30f ? copysignf(1e-30f, r->d[k]) : r->d[k];
        float inv = gpu_rcp(e, var);
        float oi = r->o[k] * inv;
        float sc = ldexpf(1.0f, ex) * inv;
        float base = fmaf(org, inv, -oi);
        float tl = fmaf((float)l, sc, base), th = fmaf((float)h, sc, base);
        tn = fmaxf(tn, inv >= 0.0f ? tl : th);
        tf = fminf(tf, inv >= 0.0f ? th : tl);
    }
    return tn <= fmaf(fabsf(tf), rel, tf);
}

void orc_gpu_box_check(const rtq_bvh_node* nodes, uint32_t n_nodes, const rtq_ray* rays, const rtq_hit* hits,
                       uint64_t n_rays, float pad, float rel, uint64_t* lost_binary, uint64_t* lost_wide) {
    uint32_t* parent = (uint32_t*)malloc(sizeof(uint32_t) * (n_nodes ? n_nodes : 1));
    uint32_t n_tris = (n_nodes + 1) / 2;
    uint32_t* leaf_of = (uint32_t*)malloc(sizeof(uint32_t) * (n_tris ? n_tris : 1));
    for (uint32_t i = 0; i < n_nodes; ++i) parent[i] = 0xFFFFFFFFu;
    for (uint32_t i = 0; i < n_nodes; ++i) {
        const rtq_bvh_node* nd = &nodes[i];
        if (nd->left_child == 0 && nd->right_child == 0) {
            if (nd->triangle_id < n_tris) leaf_of[nd->triangle_id] = i;
        } else {
            parent[nd->left_child] = i;
            parent[nd->right_child] = i;
        }
    }
    uint64_t lb = 0, lw = 0;
    for (uint64_t r = 0; r < n_rays; ++r) {
        if (hits[r].id < 0 || (uint32_t)hits[r].id >= n_tris) continue;
        int bad_b = 0, bad_w = 0;
        for (int var = 0; var < 3; ++var) {
            for (uint32_t x = leaf_of[hits[r].id]; x != 0xFFFFFFFFu; x = parent[x]) {
                if (!gpu_box_binary(&rays[r], var, &nodes[x].box, pad, hits[r].t, rel)) bad_b = 1;
                if (parent[x] != 0xFFFFFFFFu &&
                    !gpu_box_wide(&rays[r], var, &nodes[x].box, &nodes[parent[x]].box, pad, hits[r].t, rel))
                    bad_w = 1;
            }
        }
        lb += (uint64_t)bad_b;
        lw += (uint64_t)bad_w;
    }
    free(parent);
    free(leaf_of);
    *lost_binary = lb;
    *lost_wide = lw;
}

/* ------------------------------------------------------------------------ */
/* RNG and ray generation                                                    */

static inline uint64_t splitmix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

uint64_t orc_rng_hash(uint64_t seed, uint64_t stream, uint64_t index) {
    return splitmix64(splitmix64(splitmix64(seed) ^ stream) ^ index);
}

float orc_u01f(uint64_t seed, uint64_t stream, uint64_t index) {
    return (float)(uint32_t)(orc_rng_hash(seed, stream, index) >> 40) * (1.0f / 16777216.0f);
}

void orc_raygen_pinhole(const float eye[3], const float right[3], const float up[3],
                        const float fwd[3], uint32_t width, uint32_t height, rtq_ray* rays) {
    for (uint32_t j = 0; j < height; ++j) {
        for (uint32_t i = 0; i < width; ++i) {
            float s = (float)i / (float)width - 0.5f;
            float t = (float)j / (float)height - 0.5f;
            float d[3];
            for (int k = 0; k < 3; ++k) d[k] = (right[k] * s + up[k] * t) + fwd[k];
            float len = sqrtf((d[0] * d[0] + d[1] * d[1]) + d[2] * d[2]);
            rtq_ray* r = &rays[(size_t)j * width + i];
            for (int k = 0; k < 3; ++k) { r->o[k] = eye[k]; r->d[k] = d[k] / len; }
            r->tmin = 0.0f;
            r->tmax = INFINITY;
        }
    }
}

/* Hit point and unit geometric normal facing the incoming ray. */
static void surface(const float* v9, const rtq_ray* ray, const rtq_hit* h, float p[3], float nrm[3]) {
    const float* tri = v9 + 9 * (size_t)h->id;
    float e1[3] = {tri[3] - tri[0], tri[4] - tri[1], tri[5] - tri[2]};
    float e2[3] = {tri[6] - tri[0], tri[7] - tri[1], tri[8] - tri[2]};
    float c[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2],
                  e1[0] * e2[1] - e1[1] * e2[0]};
    if (dot3(c, ray->d) > 0.0f) { c[0] = -c[0]; c[1] = -c[1]; c[2] = -c[2]; }
    float len = sqrtf(dot3(c, c));
    for (int k = 0; k < 3; ++k) {
        nrm[k] = c[k] / len;
        p[k] = ray->o[k] + h->t * ray->d[k];
    }
}

static void empty_ray(rtq_ray* r) {
    memset(r, 0, sizeof *r);
    r->tmin = 1.0f;
    r->tmax = 0.0f;
}

void orc_raygen_diffuse(const float* v9, const rtq_ray* prim, const rtq_hit* hits, uint64_t n,
                        uint64_t seed, float eps, rtq_ray* out) {
    for (uint64_t i = 0; i < n; ++i) {
        if (hits[i].id < 0) { empty_ray(&out[i]); continue; }
        float p[3], nrm[3];
        surface(v9, &prim[i], &hits[i], p, nrm);
        /* orthonormal basis from the normal: Duff et al., JCGT 2017 */
        float sign = copysignf(1.0f, nrm[2]);
        float a = -1.0f / (sign + nrm[2]);
        float b = nrm[0] * nrm[1] * a;
        float tx[3] = {1.0f + sign * nrm[0] * nrm[0] * a, sign * b, -sign * nrm[0]};
        float bx[3] = {b, sign + nrm[1] * nrm[1] * a, -nrm[1]};
        /* cosine-weighted: uniform point in the unit disk by rejection, lifted */
        float sx = 0.0f, sy = 0.0f;
        for (uint64_t k = 0; k < 8; ++k) {
            float ux = 2.0f * orc_u01f(seed, 2 * k, i) - 1.0f;
            float uy = 2.0f * orc_u01f(seed, 2 * k + 1, i) - 1.0f;
            if (ux * ux + uy * uy < 1.0f) { sx = ux; sy = uy; break; }
        }
        float sz = sqrtf(smax(0.0f, (1.0f - sx * sx) - sy * sy));
        rtq_ray* r = &out[i];
        for (int k = 0; k < 3; ++k) {
            r->o[k] = p[k] + eps * nrm[k];
            r->d[k] = (tx[k] * sx + bx[k] * sy) + nrm[k] * sz;
        }
        r->tmin = 0.0f;
        r->tmax = INFINITY;
    }
}

void orc_raygen_shadow(const float* v9, const rtq_ray* prim, const rtq_hit* hits, uint64_t n,
                       const float* lights, uint32_t n_lights, float eps, rtq_ray* out) {
    for (uint32_t l = 0; l < n_lights; ++l) {
        for (uint64_t i = 0; i < n; ++i) {
            rtq_ray* r = &out[(uint64_t)l * n + i];
            if (hits[i].id < 0) { empty_ray(r); continue; }
            float p[3], nrm[3];
            surface(v9, &prim[i], &hits[i], p, nrm);
            for (int k = 0; k < 3; ++k) {
                r->o[k] = p[k] + eps * nrm[k];
                r->d[k] = lights[3 * l + k] - r->o[k];
            }
            r->tmin = 0.0f;
            r->tmax = 1.0f;
        }
    }
}

void orc_scan_exclusive(const uint32_t* in, uint32_t* out, uint64_t n) {
    uint32_t acc = 0;
    for (uint64_t i = 0; i < n; ++i) { uint32_t x = in[i]; out[i] = acc; acc += x; }
}

/* rtq_bvh_refit: boxes of a kept cluster tree for new vertices.  Leaves
 * (c < n) are triangle boxes (bvh.cpp:383-399); merged clusters (c >= n) the
 * merge of their children (bvh.cpp:401-413), in increasing id order: a merge
 * always gets a higher id than both children. */
void orc_refit_boxes(const rtq_triangle* tris, const rtq_mesh_model* models, uint32_t n, const uint32_t* tri_id,
                     const uint32_t* left, const uint32_t* right, rtq_aabb* boxes) {
    for (uint32_t c = 0; c < n; ++c) {
        const rtq_triangle* t = &tris[tri_id[c]];
        boxes[c] = orc_triangle_box(t, &models[t->model_id]);
    }
    for (uint32_t c = n; c + 1 < 2 * n; ++c) boxes[c] = orc_merge(&boxes[left[c]], &boxes[right[c]]);
}
