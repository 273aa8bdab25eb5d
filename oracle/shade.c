/*
 * ORACLE — test infrastructure only (see oracle.h).
 *
 * The shader's pixel colour (getColor, raytracer.glsl:159-179, and main's
 * value, :299-331) and its BVH box colouring (intersectBVH :182-237 and
 * getClosestHitBVH :246-276), restated literally: the overlay walks the whole
 * tree like the shader (no cut below the display depth) unless `cut` is set.
 */
#include <math.h>
#include <stdlib.h>

#include "oracle.h"

/* GLSL 4.60 section 8.3 */
static float smin_glsl(float x, float y) { return y < x ? y : x; }
static float smax_glsl(float x, float y) { return x < y ? y : x; }

int orc_ref_box_hit(const rtq_ray* ray, const rtq_aabb* b, float threshold) {
    /* raytracer.glsl:182-237 */
    float inv_x = 1.0f / ray->d[0];
    float tx1 = (b->min[0] - ray->o[0]) * inv_x;
    float tx2 = (b->max[0] - ray->o[0]) * inv_x;
    float tmin = smin_glsl(tx1, tx2);
    float tmax = smax_glsl(tx1, tx2);
    if (tmax < 0.0f || tmin > tmax) return 0;
    float inv_y = 1.0f / ray->d[1];
    float ty1 = (b->min[1] - ray->o[1]) * inv_y;
    float ty2 = (b->max[1] - ray->o[1]) * inv_y;
    tmin = smax_glsl(tmin, smin_glsl(ty1, ty2));
    tmax = smin_glsl(tmax, smax_glsl(ty1, ty2));
    if (tmax < 0.0f || tmin > tmax) return 0;
    float inv_z = 1.0f / ray->d[2];
    float tz1 = (b->min[2] - ray->o[2]) * inv_z;
    float tz2 = (b->max[2] - ray->o[2]) * inv_z;
    tmin = smax_glsl(tmin, smin_glsl(tz1, tz2));
    tmax = smin_glsl(tmax, smax_glsl(tz1, tz2));
    if (tmax >= 0.0f && tmin <= tmax) {
        float e[3];
        for (int k = 0; k < 3; ++k) e[k] = ray->o[k] + ray->d[k] * tmin;
        int close[3];
        for (int k = 0; k < 3; ++k)
            close[k] = (fabsf(e[k] - b->min[k]) < threshold) || (fabsf(e[k] - b->max[k]) < threshold);
        if ((close[0] && close[1]) || (close[0] && close[2]) || (close[1] && close[2])) return 2;
        return 1;
    }
    return 0;
}

void orc_bvh_overlay(const rtq_bvh_node* nodes, uint32_t n_nodes, const rtq_ray* rays, uint64_t n,
                     uint32_t depth, int cut, float* rgba) {
    /* raytracer.glsl:246-276 with uDepthDisplayBVH = depth */
    const float threshold = 0.05f / ((float)depth + 1.0f);
    uint32_t* stack = (uint32_t*)malloc(sizeof(uint32_t) * 2 * (size_t)(n_nodes + 1));
    uint32_t* dstack = stack + (n_nodes + 1);
    for (uint64_t r = 0; r < n; ++r) {
        float c[4] = {0.0f, 0.0f, 0.0f, 0.0f};
        uint32_t sp = 0;
        stack[sp] = 0;
        dstack[sp] = 0;
        ++sp;
        while (sp > 0) {
            --sp;
            uint32_t cur = stack[sp], d = dstack[sp];
            const rtq_bvh_node* nd = &nodes[cur];
            int h = orc_ref_box_hit(&rays[r], &nd->box, threshold);
            if (h == 0) continue;
            if (d == depth) {
                const float line[4] = {0.7f, 0.0f, 0.7f, 0.1f}, box[4] = {0.5f, 0.0f, 0.5f, 0.1f};
                for (int k = 0; k < 4; ++k) c[k] = (h == 2) ? line[k] : box[k];
                if (cut) continue;
            }
            if (nd->left_child == 0 && nd->right_child == 0) continue;
            stack[sp] = nd->left_child;
            dstack[sp] = d + 1;
            ++sp;
            stack[sp] = nd->right_child;
            dstack[sp] = d + 1;
            ++sp;
        }
        for (int k = 0; k < 4; ++k) rgba[4 * r + k] = c[k];
    }
    free(stack);
}

void orc_shade(const rtq_triangle* tris, uint32_t n_tris, const rtq_mesh_model* models, const rtq_material* materials,
               uint32_t n_materials, const rtq_hit* hits, uint64_t n, const float* overlay, uint32_t flags,
               float* rgba) {
    /* raytracer.glsl:301 value = (0,0,0,1); getColor :159-179 */
    for (uint64_t i = 0; i < n; ++i) {
        float c[4] = {0.0f, 0.0f, 0.0f, 1.0f};
        if (overlay)
            for (int k = 0; k < 4; ++k) c[k] = overlay[4 * i + k];
        const rtq_hit* h = &hits[i];
        if (h->id >= 0) {
            if ((uint32_t)h->id < n_tris) {
                uint32_t m = models[tris[h->id].model_id].material_id;
                if (m < n_materials)
                    for (int k = 0; k < 4; ++k) c[k] = c[k] + materials[m].color[k];
            }
            if (flags & 1u) {
                float b2 = (1.0f - h->u) - h->v;
                if (h->u < 0.02f || h->v < 0.02f || b2 < 0.02f) {
                    c[0] = 0.0f; c[1] = 0.0f; c[2] = 0.0f; c[3] = 1.0f;
                }
            }
        }
        for (int k = 0; k < 4; ++k) rgba[4 * i + k] = c[k];
    }
}
