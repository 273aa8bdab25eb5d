// C ABI of rtq: contexts, scenes, build/export entry points, ray generation
// and trace dispatch.  No entry point exits the process; failures come back
// as rtq_status with a message in rtq_last_error().
#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstring>

#include "rtq_internal.hpp"

namespace rtq {

rtq_status build_scene(rtq_ctx* ctx, Scene* s, const rtq_build_params& prm);
rtq_status morton_codes(rtq_ctx* ctx, Scene* s, uint32_t flags, uint32_t* host_codes);
rtq_status trace(rtq_ctx* ctx, const Scene* s, const rtq_ray* rays, uint64_t n, uint32_t flags, rtq_hit* hits,
                 uint32_t* stat_nodes, uint32_t* stat_tris, bool stats, hipStream_t st);

rtq_status fail(rtq_ctx* ctx, rtq_status s, const char* fmt, ...) {
    if (ctx) {
        char buf[512];
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(buf, sizeof buf, fmt, ap);
        va_end(ap);
        ctx->err = buf;
    }
    return s;
}

rtq_status hip_fail(rtq_ctx* ctx, hipError_t e, const char* what) {
    const rtq_status s = (e == hipErrorOutOfMemory) ? RTQ_ENOMEM : RTQ_EHIP;
    return fail(ctx, s, "%s: %s", what, hipGetErrorString(e));
}

Scene* find_scene(rtq_ctx* ctx, uint32_t id) {
    auto it = ctx->scenes.find(id);
    return it == ctx->scenes.end() ? nullptr : it->second;
}

hipStream_t pick_stream(rtq_ctx* ctx, void* stream) {
    return stream ? reinterpret_cast<hipStream_t>(stream) : ctx->stream;
}

template <typename T>
static void release(T*& p) {
    if (p) (void)hipFree(p);
    p = nullptr;
}

// Before a rebuild: the buffers sized by the tree are let go, their memory
// kept in s->kept_* for the build to take back when it is large enough; those
// sized by the triangle count (cluster arrays, sorted codes, reference nodes,
// traversal triangles) and the fixed-size top table stay for the build to
// overwrite.
void reset_build(Scene* s) {
    s->nodes = nullptr;
    s->node_aux = nullptr;
    s->wide = nullptr;
    s->wide_trav = nullptr;
    s->wide_kids = nullptr;
    s->wide_ordered = false;
    s->d_passes = nullptr;
    s->lds_from = 0;
    s->n_top = 0;
    s->built = false;
    s->stale = false;
    s->generation = 0;
    s->passes.clear();
    s->ordered = false;
    s->order_closest = false;
    s->order_any = false;
    s->n_trav = 0;
    s->n_wide = 0;
    s->wide_depth = 0;
    s->ref_ready = false;
}

static void release(Kept& k) {
    if (k.p) (void)hipFree(k.p);
    k.p = nullptr;
    k.bytes = 0;
}

void free_build(Scene* s) {
    reset_build(s);
    release(s->kept_nodes);
    release(s->kept_aux);
    release(s->kept_wide);
    release(s->kept_wide_trav);
    release(s->kept_wide_kids);
    release(s->kept_passes);
    release(s->cbox);
    release(s->ctri);
    release(s->cleft);
    release(s->cright);
    release(s->cparent);
    release(s->ccount);
    release(s->sorted_idx);
    release(s->sorted_codes);
    release(s->ref_nodes);
    release(s->preorder);
    release(s->first_leaf);
    release(s->node_depth);
    release(s->trav_index);
    release(s->top);
    release(s->tris);
}

rtq_status arena_reserve(rtq_ctx* ctx) {
    Arena& a = ctx->arena;
    if (a.top != 0 || a.extra != 0) return fail(ctx, RTQ_EINTERNAL, "arena in use at the start of an operation");
    if (a.peak <= a.cap) return RTQ_OK;
    if (a.base) {
        (void)hipFree(a.base);
        a.base = nullptr;
        a.cap = 0;
    }
    const size_t want = (a.peak + ((size_t)1 << 20) - 1) & ~(((size_t)1 << 20) - 1);
    const hipError_t e = hipMalloc(reinterpret_cast<void**>(&a.base), want);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        a.base = nullptr;  // temporaries fall back to their own allocations
        return RTQ_OK;
    }
    a.cap = want;
    return RTQ_OK;
}

static void free_scene(Scene* s) {
    free_build(s);
    release(s->tris_ref);
    release(s->models);
    release(s->tri_world);
    release(s->tris_stage);
    release(s->check_words);
    delete s;
}

namespace {

constexpr int kThreads = 256;
inline uint32_t blocks_for(uint64_t n) { return (uint32_t)((n + kThreads - 1) / kThreads); }

// World-space vertices, model matrix applied in glm's order (bvh.cpp:238-240):
// 3 float4 per triangle, the original id in the w of the first.
__global__ __launch_bounds__(kThreads) void k_world(const rtq_triangle* __restrict__ tris,
                                                    const DevModel* __restrict__ models, uint32_t n,
                                                    float4* __restrict__ out) {
    const uint32_t i = blockIdx.x * kThreads + threadIdx.x;
    if (i >= n) return;
    const rtq_triangle t = tris[i];
    const DevModel& md = models[t.model_id];
    const rtq_vec4 ps[3] = {t.p0, t.p1, t.p2};
#pragma unroll
    for (int v = 0; v < 3; ++v) {
        float o[3];
#pragma unroll
        for (int r = 0; r < 3; ++r) {
            float a0 = md.m[0 + r] * ps[v].x + md.m[4 + r] * ps[v].y;
            float a1 = md.m[8 + r] * ps[v].z + md.m[12 + r] * ps[v].w;
            o[r] = a0 + a1;
        }
        out[3ull * i + v] = make_float4(o[0], o[1], o[2], v == 0 ? __uint_as_float(i) : 0.0f);
    }
}

// The upload checks of every triangle: bad[0] becomes the lowest index whose
// model id is out of range, bad[1] the lowest whose vertex coordinates (w
// included) are not all finite and at most 2^28 in magnitude; both start at
// kNone.  A float is finite with |x| <= 2^28 exactly when its bits without the
// sign are at most those of 2^28 (0x4D800000): infinities and NaNs lie above.
__global__ __launch_bounds__(kThreads) void k_check_tris(const uint4* __restrict__ tris, uint32_t n,
                                                         uint32_t n_models, uint32_t* __restrict__ bad) {
    const uint32_t i = blockIdx.x * kThreads + threadIdx.x;
    if (i >= n) return;
    const uint4 p0 = tris[4ull * i], p1 = tris[4ull * i + 1], p2 = tris[4ull * i + 2], rest = tris[4ull * i + 3];
    auto mag = [](uint4 v) {   // largest magnitude bits of the four words
        return max(max(v.x & 0x7FFFFFFFu, v.y & 0x7FFFFFFFu), max(v.z & 0x7FFFFFFFu, v.w & 0x7FFFFFFFu));
    };
    const uint32_t m = max(mag(p0), max(mag(p1), mag(p2)));
    if (rest.x >= n_models) atomicMin(bad, i);
    if (m > 0x4D800000u) atomicMin(bad + 1, i);
}

__device__ __forceinline__ uint64_t splitmix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// Counter-based uniform in [0,1): top 24 bits of SplitMix64(seed, stream, index).
__device__ __forceinline__ float u01f(uint64_t seed, uint64_t stream, uint64_t index) {
    const uint64_t h = splitmix64(splitmix64(splitmix64(seed) ^ stream) ^ index);
    return (float)(uint32_t)(h >> 40) * (1.0f / 16777216.0f);
}

__global__ __launch_bounds__(kThreads) void k_pinhole(rtq_pinhole cam, rtq_ray* __restrict__ rays) {
    const uint64_t idx = (uint64_t)blockIdx.x * kThreads + threadIdx.x;
    const uint64_t total = (uint64_t)cam.width * cam.height;
    if (idx >= total) return;
    const uint32_t i = (uint32_t)(idx % cam.width), j = (uint32_t)(idx / cam.width);
    const float s = (float)i / (float)cam.width - 0.5f;
    const float t = (float)j / (float)cam.height - 0.5f;
    float d[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) d[k] = (cam.right[k] * s + cam.up[k] * t) + cam.forward[k];
    const float len = sqrtf((d[0] * d[0] + d[1] * d[1]) + d[2] * d[2]);
    rtq_ray r;
    for (int k = 0; k < 3; ++k) { r.o[k] = cam.eye[k]; r.d[k] = d[k] / len; }
    r.tmin = 0.0f;
    r.tmax = INFINITY;
    rays[idx] = r;
}

__device__ __forceinline__ void empty_ray(rtq_ray& r) {
    r.o[0] = r.o[1] = r.o[2] = 0.0f;
    r.d[0] = r.d[1] = r.d[2] = 0.0f;
    r.tmin = 1.0f;
    r.tmax = 0.0f;
}

// Hit point and unit geometric normal facing the incoming ray.
__device__ __forceinline__ void surface(const float4* __restrict__ tw, const rtq_ray& ray, const rtq_hit& h,
                                        float* p, float* nrm) {
    const float4 v0 = tw[3ull * h.id], v1 = tw[3ull * h.id + 1], v2 = tw[3ull * h.id + 2];
    const float e1[3] = {v1.x - v0.x, v1.y - v0.y, v1.z - v0.z};
    const float e2[3] = {v2.x - v0.x, v2.y - v0.y, v2.z - v0.z};
    float c[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
    if ((c[0] * ray.d[0] + c[1] * ray.d[1]) + c[2] * ray.d[2] > 0.0f) { c[0] = -c[0]; c[1] = -c[1]; c[2] = -c[2]; }
    const float len = sqrtf((c[0] * c[0] + c[1] * c[1]) + c[2] * c[2]);
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        nrm[k] = c[k] / len;
        p[k] = ray.o[k] + h.t * ray.d[k];
    }
}

__global__ __launch_bounds__(kThreads) void k_diffuse(const float4* __restrict__ tw, const rtq_ray* __restrict__ prim,
                                                      const rtq_hit* __restrict__ hits, uint64_t n, uint64_t seed,
                                                      float eps, rtq_ray* __restrict__ out) {
    const uint64_t i = (uint64_t)blockIdx.x * kThreads + threadIdx.x;
    if (i >= n) return;
    const rtq_hit h = hits[i];
    rtq_ray r;
    if (h.id < 0) { empty_ray(r); out[i] = r; return; }
    const rtq_ray pr = prim[i];
    float p[3], nrm[3];
    surface(tw, pr, h, p, nrm);
    // orthonormal basis (Duff et al., JCGT 2017)
    const float sign = copysignf(1.0f, nrm[2]);
    const float a = -1.0f / (sign + nrm[2]);
    const float b = nrm[0] * nrm[1] * a;
    const float tx[3] = {1.0f + sign * nrm[0] * nrm[0] * a, sign * b, -sign * nrm[0]};
    const float bx[3] = {b, sign + nrm[1] * nrm[1] * a, -nrm[1]};
    // cosine-weighted: rejection-sampled unit disk point lifted to the hemisphere
    float sx = 0.0f, sy = 0.0f;
    for (uint64_t k = 0; k < 8; ++k) {
        const float ux = 2.0f * u01f(seed, 2 * k, i) - 1.0f;
        const float uy = 2.0f * u01f(seed, 2 * k + 1, i) - 1.0f;
        if (ux * ux + uy * uy < 1.0f) { sx = ux; sy = uy; break; }
    }
    const float q = (1.0f - sx * sx) - sy * sy;
    const float sz = sqrtf(q < 0.0f ? 0.0f : q);
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        r.o[k] = p[k] + eps * nrm[k];
        r.d[k] = (tx[k] * sx + bx[k] * sy) + nrm[k] * sz;
    }
    r.tmin = 0.0f;
    r.tmax = INFINITY;
    out[i] = r;
}

struct Lights {
    float xyz[3 * 16];
};

__global__ __launch_bounds__(kThreads) void k_shadow(const float4* __restrict__ tw, const rtq_ray* __restrict__ prim,
                                                     const rtq_hit* __restrict__ hits, uint64_t n, Lights lights,
                                                     uint32_t n_lights, float eps, rtq_ray* __restrict__ out) {
    const uint64_t i = (uint64_t)blockIdx.x * kThreads + threadIdx.x;
    if (i >= n) return;
    const rtq_hit h = hits[i];
    if (h.id < 0) {
        rtq_ray r;
        empty_ray(r);
        for (uint32_t l = 0; l < n_lights; ++l) out[(uint64_t)l * n + i] = r;
        return;
    }
    const rtq_ray pr = prim[i];
    float p[3], nrm[3];
    surface(tw, pr, h, p, nrm);
    float o[3];
    for (int k = 0; k < 3; ++k) o[k] = p[k] + eps * nrm[k];
    for (uint32_t l = 0; l < n_lights; ++l) {
        rtq_ray r;
        for (int k = 0; k < 3; ++k) {
            r.o[k] = o[k];
            r.d[k] = lights.xyz[3 * l + k] - o[k];
        }
        r.tmin = 0.0f;
        r.tmax = 1.0f;
        out[(uint64_t)l * n + i] = r;
    }
}

bool finite_and_bounded(float x, float lim) { return std::isfinite(x) && std::fabs(x) <= lim; }

}  // namespace
}  // namespace rtq

using namespace rtq;

extern "C" {

int rtq_abi_version(void) { return RTQ_ABI_VERSION; }

const char* rtq_status_string(rtq_status s) {
    switch (s) {
        case RTQ_OK: return "ok";
        case RTQ_EINVAL: return "invalid argument";
        case RTQ_ENOMEM: return "out of device memory";
        case RTQ_EHIP: return "HIP runtime error";
        case RTQ_ECAPACITY: return "capacity exceeded";
        case RTQ_EINTERNAL: return "internal error";
        case RTQ_ENOTBUILT: return "scene has no BVH";
    }
    return "unknown status";
}

rtq_status rtq_create(rtq_ctx** out, int hip_device) {
    if (!out) return RTQ_EINVAL;
    *out = nullptr;
    int n_dev = 0;
    if (hipGetDeviceCount(&n_dev) != hipSuccess || n_dev <= 0) return RTQ_EHIP;
    if (hip_device < 0 || hip_device >= n_dev) return RTQ_EINVAL;
    if (hipSetDevice(hip_device) != hipSuccess) return RTQ_EHIP;
    rtq_ctx* ctx = new rtq_ctx();
    ctx->device = hip_device;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, hip_device) == hipSuccess) ctx->n_cu = prop.multiProcessorCount;
    if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&ctx->side, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&ctx->fork, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&ctx->join, hipEventDisableTiming) != hipSuccess ||
        hipHostMalloc(&ctx->pinned, kPinnedBytes, hipHostMallocDefault) != hipSuccess ||
        scratch_alloc(ctx, &ctx->scratch) != RTQ_OK || noalloc_setup(ctx) != RTQ_OK) {
        scratch_free(&ctx->scratch);
        if (ctx->pinned) (void)hipHostFree(ctx->pinned);
        if (ctx->join) (void)hipEventDestroy(ctx->join);
        if (ctx->fork) (void)hipEventDestroy(ctx->fork);
        if (ctx->side) (void)hipStreamDestroy(ctx->side);
        if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
        delete ctx;
        return RTQ_EHIP;
    }
    if (const char* e = getenv("RTQ_REFILL")) {
        const int v = atoi(e);
        if (v >= 1 && v <= 64) ctx->refill = (uint32_t)v;
    }
    if (const char* e = getenv("RTQ_CHUNK")) {
        const int v = atoi(e);
        if (v >= 1 && v <= 4096) ctx->chunk = (uint32_t)v;
    }
    if (const char* e = getenv("RTQ_SPEC_SLACK")) {
        const int v = atoi(e);
        if (v >= 0 && v <= 63) ctx->spec_slack = (uint32_t)v;
    }
    if (const char* e = getenv("RTQ_LEAF_SLACK")) {
        const int v = atoi(e);
        if (v >= 0 && v <= 63) ctx->leaf_slack = (uint32_t)v;
    }
    *out = ctx;
    return RTQ_OK;
}

void rtq_destroy(rtq_ctx* ctx) {
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipStreamSynchronize(ctx->side);
    for (auto& kv : ctx->scenes) free_scene(kv.second);
    ctx->scenes.clear();
    scratch_free(&ctx->scratch);
    if (ctx->arena.base) (void)hipFree(ctx->arena.base);
    (void)hipHostFree(ctx->pinned);
    (void)hipEventDestroy(ctx->join);
    (void)hipEventDestroy(ctx->fork);
    (void)hipStreamDestroy(ctx->side);
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

const char* rtq_last_error(const rtq_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

rtq_status rtq_synchronize(rtq_ctx* ctx) {
    if (!ctx) return RTQ_EINVAL;
    RTQ_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return RTQ_OK;
}

// Upload checks (shared with rtq_scene_update): model ids in range, vertices
// and matrix entries finite and bounded by 2^28 so every box surface area is
// finite.  The model matrices are checked here on the host, before anything
// is copied.
static rtq_status check_models(rtq_ctx* ctx, const char* what, const rtq_mesh_model* models, uint32_t n_models) {
    const float lim = 268435456.0f;
    for (uint32_t m = 0; m < n_models; ++m)
        for (int k = 0; k < 16; ++k)
            if (!finite_and_bounded(models[m].m[k], lim))
                return fail(ctx, RTQ_EINVAL, "%s: model %u has a non-finite or huge entry", what, m);
    return RTQ_OK;
}

// The triangles are checked on the device after their copy (dev, n of them;
// host = the caller's array for the message, or null when the triangles came
// from device memory): one pass over HBM instead of a second pass over host
// memory.  Waits for the stream; the message names the lowest failing
// triangle and, as a host loop over the triangles in order would, its model
// id first.
static rtq_status check_tris_on_device(rtq_ctx* ctx, const char* what, const rtq_triangle* host,
                                       const rtq_triangle* dev, uint32_t n, uint32_t n_models, uint32_t* words,
                                       hipStream_t st) {
    RTQ_HIP(ctx, hipMemsetAsync(words, 0xFF, 2 * sizeof(uint32_t), st));
    hipLaunchKernelGGL(k_check_tris, dim3(blocks_for(n)), dim3(kThreads), 0, st, reinterpret_cast<const uint4*>(dev),
                       n, n_models, words);
    RTQ_HIP(ctx, hipGetLastError());
    uint32_t bad[2];
    HostReads rd(ctx, st);
    RTQ_HIP(ctx, rd.add(bad, words, sizeof bad));
    RTQ_HIP(ctx, rd.wait());
    const uint32_t i = std::min(bad[0], bad[1]);
    if (i == kNone) return RTQ_OK;
    if (bad[0] == i) {
        uint32_t model_id = 0;
        if (host) {
            model_id = host[i].model_id;
        } else {
            RTQ_HIP(ctx, rd.add(&model_id, &dev[i].model_id, sizeof model_id));
            RTQ_HIP(ctx, rd.wait());
        }
        return fail(ctx, RTQ_EINVAL, "%s: triangle %u model id %u out of range", what, i, model_id);
    }
    return fail(ctx, RTQ_EINVAL, "%s: triangle %u has a non-finite or huge vertex", what, i);
}

rtq_status rtq_scene_upload(rtq_ctx* ctx, const rtq_triangle* tris, uint32_t n_tris, const rtq_mesh_model* models,
                            uint32_t n_models, uint32_t* scene_id) {
    if (!ctx) return RTQ_EINVAL;
    if (!tris || !models || !scene_id) return fail(ctx, RTQ_EINVAL, "scene_upload: null pointer");
    if (n_tris == 0) return fail(ctx, RTQ_EINVAL, "scene_upload: no triangles (the reference root id 2N-2 underflows)");
    if (n_models == 0) return fail(ctx, RTQ_EINVAL, "scene_upload: no mesh models");
    if (n_tris > kMaxTris) return fail(ctx, RTQ_ECAPACITY, "scene_upload: %u triangles exceeds %u", n_tris, kMaxTris);
    rtq_status rs = check_models(ctx, "scene_upload", models, n_models);
    if (rs != RTQ_OK) return rs;
    RTQ_HIP(ctx, hipSetDevice(ctx->device));
    Scene* s = new Scene();
    s->n_tris = n_tris;
    s->n_models = n_models;
    hipError_t e = dev_alloc(&s->tris_ref, n_tris);
    if (e == hipSuccess) e = dev_alloc(&s->models, n_models);
    if (e == hipSuccess) e = dev_alloc(&s->tri_world, 3ull * n_tris);
    if (e == hipSuccess) e = dev_alloc(&s->check_words, 2);
    if (e == hipSuccess) e = hipMemcpyAsync(s->tris_ref, tris, sizeof(rtq_triangle) * n_tris, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(s->models, models, sizeof(DevModel) * n_models, hipMemcpyHostToDevice, ctx->stream);
    if (e != hipSuccess) {
        free_scene(s);
        return hip_fail(ctx, e, "scene_upload");
    }
    rs = check_tris_on_device(ctx, "scene_upload", tris, s->tris_ref, n_tris, n_models, s->check_words, ctx->stream);
    if (rs != RTQ_OK) {
        free_scene(s);
        return rs;
    }
    hipLaunchKernelGGL(k_world, dim3(blocks_for(n_tris)), dim3(kThreads), 0, ctx->stream, s->tris_ref, s->models,
                       n_tris, s->tri_world);
    e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) {
        free_scene(s);
        return hip_fail(ctx, e, "scene_upload");
    }
    const uint32_t id = ctx->next_scene++;
    ctx->scenes[id] = s;
    *scene_id = id;
    return RTQ_OK;
}

rtq_status rtq_scene_update(rtq_ctx* ctx, uint32_t scene_id, const rtq_triangle* tris, uint32_t n_tris,
                            const rtq_mesh_model* models, uint32_t n_models) {
    if (!ctx) return RTQ_EINVAL;
    Scene* s = find_scene(ctx, scene_id);
    if (!s) return fail(ctx, RTQ_EINVAL, "scene_update: unknown scene %u", scene_id);
    if (tris && n_tris != s->n_tris)
        return fail(ctx, RTQ_EINVAL, "scene_update: %u triangles, the scene has %u", n_tris, s->n_tris);
    if (models && n_models != s->n_models)
        return fail(ctx, RTQ_EINVAL, "scene_update: %u models, the scene has %u", n_models, s->n_models);
    if (!tris && !models) return RTQ_OK;
    if (models) {
        const rtq_status rs = check_models(ctx, "scene_update", models, s->n_models);
        if (rs != RTQ_OK) return rs;
    }
    RTQ_HIP(ctx, hipSetDevice(ctx->device));
    // work already queued on the scene may still read the arrays
    RTQ_HIP(ctx, hipDeviceSynchronize());
    if (tris) {
        // new triangles go to the staging array first and replace the
        // scene's only once they pass the checks (model ids against the kept
        // model count), so a rejected update leaves the scene as it was; the
        // scene's array keeps its address for work captured on it
        if (!s->tris_stage) RTQ_HIP(ctx, dev_alloc(&s->tris_stage, s->n_tris));
        const size_t bytes = sizeof(rtq_triangle) * s->n_tris;
        RTQ_HIP(ctx, hipMemcpyAsync(s->tris_stage, tris, bytes, hipMemcpyHostToDevice, ctx->stream));
        const rtq_status rs = check_tris_on_device(ctx, "scene_update", tris, s->tris_stage, s->n_tris, s->n_models,
                                                   s->check_words, ctx->stream);
        if (rs != RTQ_OK) return rs;
        RTQ_HIP(ctx, hipMemcpyAsync(s->tris_ref, s->tris_stage, bytes, hipMemcpyDeviceToDevice, ctx->stream));
    }
    if (models)
        RTQ_HIP(ctx, hipMemcpyAsync(s->models, models, sizeof(DevModel) * s->n_models, hipMemcpyHostToDevice,
                                    ctx->stream));
    hipLaunchKernelGGL(k_world, dim3(blocks_for(s->n_tris)), dim3(kThreads), 0, ctx->stream, s->tris_ref, s->models,
                       s->n_tris, s->tri_world);
    RTQ_HIP(ctx, hipGetLastError());
    RTQ_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (s->built) {
        s->built = false;
        s->stale = true;
    }
    return RTQ_OK;
}

rtq_status rtq_scene_update_device(rtq_ctx* ctx, uint32_t scene_id, const rtq_triangle* tris_dev, uint32_t n_tris) {
    if (!ctx) return RTQ_EINVAL;
    Scene* s = find_scene(ctx, scene_id);
    if (!s) return fail(ctx, RTQ_EINVAL, "scene_update_device: unknown scene %u", scene_id);
    if (!tris_dev) return fail(ctx, RTQ_EINVAL, "scene_update_device: null pointer");
    if (n_tris != s->n_tris)
        return fail(ctx, RTQ_EINVAL, "scene_update_device: %u triangles, the scene has %u", n_tris, s->n_tris);
    RTQ_HIP(ctx, hipSetDevice(ctx->device));
    hipPointerAttribute_t attr{};
    if (hipPointerGetAttributes(&attr, tris_dev) != hipSuccess || attr.type != hipMemoryTypeDevice ||
        attr.device != ctx->device) {
        (void)hipGetLastError();
        return fail(ctx, RTQ_EINVAL, "scene_update_device: the triangles are not in device memory of device %d",
                    ctx->device);
    }
    // work already queued on the scene may still read the arrays
    RTQ_HIP(ctx, hipDeviceSynchronize());
    // checked where they lie, so a rejected update leaves the scene as it was
    const rtq_status rs = check_tris_on_device(ctx, "scene_update_device", nullptr, tris_dev, s->n_tris, s->n_models,
                                               s->check_words, ctx->stream);
    if (rs != RTQ_OK) return rs;
    RTQ_HIP(ctx, hipMemcpyAsync(s->tris_ref, tris_dev, sizeof(rtq_triangle) * s->n_tris, hipMemcpyDeviceToDevice,
                                ctx->stream));
    hipLaunchKernelGGL(k_world, dim3(blocks_for(s->n_tris)), dim3(kThreads), 0, ctx->stream, s->tris_ref, s->models,
                       s->n_tris, s->tri_world);
    RTQ_HIP(ctx, hipGetLastError());
    RTQ_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (s->built) {
        s->built = false;
        s->stale = true;
    }
    return RTQ_OK;
}

rtq_status rtq_bvh_refit(rtq_ctx* ctx, uint32_t scene_id) {
    if (!ctx) return RTQ_EINVAL;
    Scene* s = find_scene(ctx, scene_id);
    if (!s) return fail(ctx, RTQ_EINVAL, "bvh_refit: unknown scene %u", scene_id);
    if (!s->built && !s->stale) return fail(ctx, RTQ_ENOTBUILT, "bvh_refit: scene %u has no BVH", scene_id);
    RTQ_HIP(ctx, hipSetDevice(ctx->device));
    RTQ_HIP(ctx, hipDeviceSynchronize());
    s->built = false;
    rtq_status rs = refit_scene(ctx, s);
    if (rs != RTQ_OK) return rs;
    // the traversal nodes were written with empty order words
    s->ordered = false;
    s->order_closest = false;
    s->order_any = false;
    s->stale = false;
    s->built = true;
    return RTQ_OK;
}

rtq_status rtq_scene_release(rtq_ctx* ctx, uint32_t scene_id) {
    if (!ctx) return RTQ_EINVAL;
    Scene* s = find_scene(ctx, scene_id);
    if (!s) return fail(ctx, RTQ_EINVAL, "scene_release: unknown scene %u", scene_id);
    (void)hipStreamSynchronize(ctx->stream);
    free_scene(s);
    ctx->scenes.erase(scene_id);
    return RTQ_OK;
}

void rtq_build_params_default(rtq_build_params* p) {
    if (!p) return;
    p->search_radius = 16;
    p->max_leaf_size = 4;
    p->flags = 0;
    p->node_width = 2;
}

rtq_status rtq_bvh_build(rtq_ctx* ctx, uint32_t scene_id, const rtq_build_params* params) {
    if (!ctx) return RTQ_EINVAL;
    Scene* s = find_scene(ctx, scene_id);
    if (!s) return fail(ctx, RTQ_EINVAL, "bvh_build: unknown scene %u", scene_id);
    rtq_build_params p;
    rtq_build_params_default(&p);
    if (params) p = *params;
    if (p.search_radius < 1 || p.search_radius > 32) return fail(ctx, RTQ_EINVAL, "bvh_build: search_radius must be 1..32");
    if (p.max_leaf_size < 1 || p.max_leaf_size > kMaxLeaf) return fail(ctx, RTQ_EINVAL, "bvh_build: max_leaf_size must be 1..8");
    if (p.flags & ~(RTQ_BUILD_REF_PADDED_BOUNDS | RTQ_BUILD_PLOC_SPLIT))
        return fail(ctx, RTQ_EINVAL, "bvh_build: unknown flags 0x%x", p.flags);
    if (p.node_width == 0) p.node_width = 2;
    if (p.node_width != 2 && p.node_width != 8) return fail(ctx, RTQ_EINVAL, "bvh_build: node_width must be 2 or 8");
    RTQ_HIP(ctx, hipSetDevice(ctx->device));
    rtq_status st = build_scene(ctx, s, p);
    if (st != RTQ_OK) free_build(s);
    return st;
}

rtq_status rtq_bvh_info_get(rtq_ctx* ctx, uint32_t scene_id, rtq_bvh_info* out) {
    if (!ctx || !out) return RTQ_EINVAL;
    Scene* s = find_scene(ctx, scene_id);
    if (!s) return fail(ctx, RTQ_EINVAL, "bvh_info: unknown scene %u", scene_id);
    if (!s->built) return fail(ctx, RTQ_ENOTBUILT, "bvh_info: scene %u has no BVH", scene_id);
    out->n_tris = s->n_tris;
    out->n_ref_nodes = s->n_clusters;
    out->n_trav_nodes = s->n_trav;
    out->depth = s->depth;
    out->ploc_iterations = s->ploc_passes;
    out->max_leaf_size = s->params.max_leaf_size;
    out->box_pad = s->box_pad;
    out->build_ms = s->build_ms;
    out->node_width = s->params.node_width;
    out->n_wide_nodes = s->n_wide;
    out->order_modes = s->ordered ? ((s->order_closest ? 1u : 0u) | (s->order_any ? 2u : 0u)) : 0u;
    out->wide_depth = s->wide_depth;
    out->scratch_spill_bytes = s->scratch_spill;
    return RTQ_OK;
}

rtq_status rtq_bvh_export_wide(rtq_ctx* ctx, uint32_t scene_id, void* nodes, uint64_t capacity, uint64_t* n_nodes) {
    if (!ctx) return RTQ_EINVAL;
    Scene* s = find_scene(ctx, scene_id);
    if (!s) return fail(ctx, RTQ_EINVAL, "bvh_export_wide: unknown scene %u", scene_id);
    if (!s->built) return fail(ctx, RTQ_ENOTBUILT, "bvh_export_wide: scene %u has no BVH", scene_id);
    if (!s->wide) return fail(ctx, RTQ_ENOTBUILT, "bvh_export_wide: scene %u has no 8-wide nodes", scene_id);
    if (n_nodes) *n_nodes = s->n_wide;
    if (!nodes) return RTQ_OK;
    if (capacity < s->n_wide) return fail(ctx, RTQ_ECAPACITY, "bvh_export_wide: need %u nodes", s->n_wide);
    RTQ_HIP(ctx, hipSetDevice(ctx->device));
    RTQ_HIP(ctx, hipMemcpyAsync(nodes, s->wide, sizeof(float4) * 8ull * s->n_wide, hipMemcpyDeviceToHost,
                                ctx->stream));
    RTQ_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return RTQ_OK;
}

rtq_status rtq_bvh_export_ref(rtq_ctx* ctx, uint32_t scene_id, rtq_bvh_node* nodes, uint64_t capacity,
                              uint64_t* n_nodes) {
    if (!ctx) return RTQ_EINVAL;
    Scene* s = find_scene(ctx, scene_id);
    if (!s) return fail(ctx, RTQ_EINVAL, "bvh_export_ref: unknown scene %u", scene_id);
    if (!s->built) return fail(ctx, RTQ_ENOTBUILT, "bvh_export_ref: scene %u has no BVH", scene_id);
    if (n_nodes) *n_nodes = s->n_clusters;
    if (!nodes) return RTQ_OK;
    if (capacity < s->n_clusters) return fail(ctx, RTQ_ECAPACITY, "bvh_export_ref: need %u nodes", s->n_clusters);
    RTQ_HIP(ctx, hipSetDevice(ctx->device));
    const rtq_status rs = ensure_ref_nodes(ctx, s, ctx->stream);
    if (rs != RTQ_OK) return rs;
    RTQ_HIP(ctx, hipMemcpyAsync(nodes, s->ref_nodes, sizeof(rtq_bvh_node) * s->n_clusters, hipMemcpyDeviceToHost, ctx->stream));
    RTQ_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return RTQ_OK;
}

rtq_status rtq_bvh_export_trav(rtq_ctx* ctx, uint32_t scene_id, void* nodes, uint32_t* ranges, uint64_t capacity,
                               uint64_t* n_nodes) {
    if (!ctx) return RTQ_EINVAL;
    Scene* s = find_scene(ctx, scene_id);
    if (!s) return fail(ctx, RTQ_EINVAL, "bvh_export_trav: unknown scene %u", scene_id);
    if (!s->built) return fail(ctx, RTQ_ENOTBUILT, "bvh_export_trav: scene %u has no BVH", scene_id);
    if (n_nodes) *n_nodes = s->n_trav;
    if (!nodes && !ranges) return RTQ_OK;
    if (capacity < s->n_trav) return fail(ctx, RTQ_ECAPACITY, "bvh_export_trav: need %u nodes", s->n_trav);
    RTQ_HIP(ctx, hipSetDevice(ctx->device));
    if (nodes)
        RTQ_HIP(ctx, hipMemcpyAsync(nodes, s->nodes, 4 * sizeof(float4) * s->n_trav, hipMemcpyDeviceToHost, ctx->stream));
    if (ranges)
        RTQ_HIP(ctx, hipMemcpyAsync(ranges, s->node_aux, sizeof(uint4) * s->n_trav, hipMemcpyDeviceToHost, ctx->stream));
    RTQ_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return RTQ_OK;
}

rtq_status rtq_bvh_export_clusters(rtq_ctx* ctx, uint32_t scene_id, uint64_t capacity, rtq_aabb* boxes,
                                   uint32_t* triangle_id, uint32_t* left, uint32_t* right, uint32_t* parent,
                                   uint32_t* sorted_indices, uint32_t* sorted_codes) {
    if (!ctx) return RTQ_EINVAL;
    Scene* s = find_scene(ctx, scene_id);
    if (!s) return fail(ctx, RTQ_EINVAL, "bvh_export_clusters: unknown scene %u", scene_id);
    if (!s->built) return fail(ctx, RTQ_ENOTBUILT, "bvh_export_clusters: scene %u has no BVH", scene_id);
    if (capacity < s->n_clusters) return fail(ctx, RTQ_ECAPACITY, "bvh_export_clusters: need %u entries", s->n_clusters);
    RTQ_HIP(ctx, hipSetDevice(ctx->device));
    const size_t nc = s->n_clusters, n = s->n_tris;
    hipStream_t st = ctx->stream;
    if (boxes) RTQ_HIP(ctx, hipMemcpyAsync(boxes, s->cbox, sizeof(rtq_aabb) * nc, hipMemcpyDeviceToHost, st));
    if (triangle_id) RTQ_HIP(ctx, hipMemcpyAsync(triangle_id, s->ctri, 4 * nc, hipMemcpyDeviceToHost, st));
    if (left) RTQ_HIP(ctx, hipMemcpyAsync(left, s->cleft, 4 * nc, hipMemcpyDeviceToHost, st));
    if (right) RTQ_HIP(ctx, hipMemcpyAsync(right, s->cright, 4 * nc, hipMemcpyDeviceToHost, st));
    if (parent) RTQ_HIP(ctx, hipMemcpyAsync(parent, s->cparent, 4 * nc, hipMemcpyDeviceToHost, st));
    if (sorted_indices) RTQ_HIP(ctx, hipMemcpyAsync(sorted_indices, s->sorted_idx, 4 * n, hipMemcpyDeviceToHost, st));
    if (sorted_codes) RTQ_HIP(ctx, hipMemcpyAsync(sorted_codes, s->sorted_codes, 4 * n, hipMemcpyDeviceToHost, st));
    RTQ_HIP(ctx, hipStreamSynchronize(st));
    return RTQ_OK;
}

rtq_status rtq_morton_codes(rtq_ctx* ctx, uint32_t scene_id, uint32_t flags, uint32_t* codes) {
    if (!ctx) return RTQ_EINVAL;
    Scene* s = find_scene(ctx, scene_id);
    if (!s) return fail(ctx, RTQ_EINVAL, "morton_codes: unknown scene %u", scene_id);
    if (!codes) return fail(ctx, RTQ_EINVAL, "morton_codes: null output");
    if (flags & ~RTQ_BUILD_REF_PADDED_BOUNDS) return fail(ctx, RTQ_EINVAL, "morton_codes: unknown flags");
    RTQ_HIP(ctx, hipSetDevice(ctx->device));
    return morton_codes(ctx, s, flags, codes);
}

rtq_status rtq_sort_pairs(rtq_ctx* ctx, uint32_t* keys_dev, uint32_t* vals_dev, uint64_t n, uint32_t key_bits,
                          void* stream) {
    if (!ctx) return RTQ_EINVAL;
    if (n == 0) return RTQ_OK;
    if (!keys_dev || !vals_dev) return fail(ctx, RTQ_EINVAL, "sort_pairs: null pointer");
    if (key_bits > 32) return fail(ctx, RTQ_EINVAL, "sort_pairs: key_bits must be <= 32");
    RTQ_HIP(ctx, hipSetDevice(ctx->device));
    RTQ_HIP(ctx, hipDeviceSynchronize());  // nothing queued still uses the arena
    rtq_status s = arena_reserve(ctx);
    if (s != RTQ_OK) return s;
    const hipStream_t st = pick_stream(ctx, stream);
    {
        Temps tmp(ctx);
        uint32_t *tk = nullptr, *tv = nullptr, *err = nullptr;
        RTQ_HIP(ctx, tmp.get(&tk, n));
        RTQ_HIP(ctx, tmp.get(&tv, n));
        RTQ_HIP(ctx, tmp.get(&err, 1));
        RTQ_HIP(ctx, hipMemsetAsync(err, 0, sizeof(uint32_t), st));
        s = radix_sort_pairs(ctx, keys_dev, vals_dev, tk, tv, n, key_bits, err, st);
        uint32_t h_err = 0;
        hipError_t e = hipMemcpyAsync(&h_err, err, sizeof h_err, hipMemcpyDeviceToHost, st);
        // the temporaries go back to the arena once the sort is done
        const hipError_t e2 = hipStreamSynchronize(st);
        if (e == hipSuccess) e = e2;
        if (s == RTQ_OK && e != hipSuccess) s = hip_fail(ctx, e, "hipStreamSynchronize");
        if (s == RTQ_OK && h_err) s = fail(ctx, RTQ_EINTERNAL, "sort_pairs: a digit pass's look-back gave up");
    }
    return s;
}

rtq_status rtq_scan_exclusive_u32(rtq_ctx* ctx, const uint32_t* in_dev, uint32_t* out_dev, uint64_t n, void* stream) {
    if (!ctx) return RTQ_EINVAL;
    if (n == 0) return RTQ_OK;
    if (!in_dev || !out_dev) return fail(ctx, RTQ_EINVAL, "scan: null pointer");
    RTQ_HIP(ctx, hipSetDevice(ctx->device));
    RTQ_HIP(ctx, hipDeviceSynchronize());  // nothing queued still uses the arena
    rtq_status s = arena_reserve(ctx);
    if (s != RTQ_OK) return s;
    const hipStream_t st = pick_stream(ctx, stream);
    s = scan_u32(ctx, in_dev, out_dev, n, nullptr, st);
    const hipError_t e = hipStreamSynchronize(st);  // its temporaries go back to the arena
    if (s == RTQ_OK && e != hipSuccess) s = hip_fail(ctx, e, "hipStreamSynchronize");
    return s;
}

static rtq_status trace_common(rtq_ctx* ctx, uint32_t scene_id, const rtq_ray* rays, uint64_t n, uint32_t flags,
                               rtq_hit* hits, uint32_t* sn, uint32_t* stt, bool stats, void* stream,
                               bool noalloc = false) {
    if (!ctx) return RTQ_EINVAL;
    Scene* s = find_scene(ctx, scene_id);
    if (!s) return fail(ctx, RTQ_EINVAL, "trace: unknown scene %u", scene_id);
    if (!s->built) return fail(ctx, RTQ_ENOTBUILT, "trace: scene %u has no BVH", scene_id);
    if (flags & ~(RTQ_TRACE_ANY | RTQ_TRACE_LEAF_REFMT | RTQ_TRACE_BINARY | RTQ_TRACE_LEARNED))
        return fail(ctx, RTQ_EINVAL, "trace: unknown flags 0x%x", flags);
    if (n == 0) return RTQ_OK;
    if (!rays || !hits) return fail(ctx, RTQ_EINVAL, "trace: null ray or hit buffer");
    RTQ_HIP(ctx, hipSetDevice(ctx->device));
    if (noalloc) return trace_noalloc(ctx, s, rays, n, flags, hits, pick_stream(ctx, stream));
    return trace(ctx, s, rays, n, flags, hits, sn, stt, stats, pick_stream(ctx, stream));
}

rtq_status rtq_trace(rtq_ctx* ctx, uint32_t scene_id, const rtq_ray* rays_dev, uint64_t n_rays, uint32_t flags,
                     rtq_hit* hits_dev, void* stream) {
    return trace_common(ctx, scene_id, rays_dev, n_rays, flags, hits_dev, nullptr, nullptr, false, stream);
}

rtq_status rtq_trace_noalloc(rtq_ctx* ctx, uint32_t scene_id, const rtq_ray* rays_dev, uint64_t n_rays,
                             uint32_t flags, rtq_hit* hits_dev, void* stream) {
    return trace_common(ctx, scene_id, rays_dev, n_rays, flags, hits_dev, nullptr, nullptr, false, stream, true);
}

rtq_status rtq_trace_stats(rtq_ctx* ctx, uint32_t scene_id, const rtq_ray* rays_dev, uint64_t n_rays, uint32_t flags,
                           rtq_hit* hits_dev, uint32_t* node_visits_dev, uint32_t* tri_tests_dev, void* stream) {
    return trace_common(ctx, scene_id, rays_dev, n_rays, flags, hits_dev, node_visits_dev, tri_tests_dev, true, stream);
}

rtq_status rtq_trace_host(rtq_ctx* ctx, uint32_t scene_id, const rtq_ray* rays, uint64_t n_rays, uint32_t flags,
                          rtq_hit* hits) {
    if (!ctx) return RTQ_EINVAL;
    if (n_rays == 0) return RTQ_OK;
    if (!rays || !hits) return fail(ctx, RTQ_EINVAL, "trace_host: null pointer");
    RTQ_HIP(ctx, hipSetDevice(ctx->device));
    rtq_ray* dr = nullptr;
    rtq_hit* dh = nullptr;
    RTQ_HIP(ctx, dev_alloc(&dr, n_rays));
    hipError_t e = dev_alloc(&dh, n_rays);
    if (e != hipSuccess) { (void)hipFree(dr); return hip_fail(ctx, e, "trace_host alloc"); }
    rtq_status s = RTQ_OK;
    e = hipMemcpyAsync(dr, rays, sizeof(rtq_ray) * n_rays, hipMemcpyHostToDevice, ctx->stream);
    if (e != hipSuccess) s = hip_fail(ctx, e, "trace_host copy-in");
    if (s == RTQ_OK) s = trace_common(ctx, scene_id, dr, n_rays, flags, dh, nullptr, nullptr, false, ctx->stream);
    if (s == RTQ_OK) {
        e = hipMemcpyAsync(hits, dh, sizeof(rtq_hit) * n_rays, hipMemcpyDeviceToHost, ctx->stream);
        if (e != hipSuccess) s = hip_fail(ctx, e, "trace_host copy-out");
    }
    e = hipStreamSynchronize(ctx->stream);
    if (s == RTQ_OK && e != hipSuccess) s = hip_fail(ctx, e, "trace_host sync");
    (void)hipFree(dr);
    (void)hipFree(dh);
    return s;
}

rtq_status rtq_raygen_pinhole(rtq_ctx* ctx, const rtq_pinhole* cam, rtq_ray* rays_dev, void* stream) {
    if (!ctx) return RTQ_EINVAL;
    if (!cam || !rays_dev) return fail(ctx, RTQ_EINVAL, "raygen_pinhole: null pointer");
    const uint64_t total = (uint64_t)cam->width * cam->height;
    if (total == 0) return RTQ_OK;
    RTQ_HIP(ctx, hipSetDevice(ctx->device));
    hipLaunchKernelGGL(k_pinhole, dim3(blocks_for(total)), dim3(kThreads), 0, pick_stream(ctx, stream), *cam, rays_dev);
    RTQ_HIP(ctx, hipGetLastError());
    return RTQ_OK;
}

rtq_status rtq_raygen_diffuse(rtq_ctx* ctx, uint32_t scene_id, const rtq_ray* prim_dev, const rtq_hit* hits_dev,
                              uint64_t n, uint64_t seed, float eps, rtq_ray* out_dev, void* stream) {
    if (!ctx) return RTQ_EINVAL;
    Scene* s = find_scene(ctx, scene_id);
    if (!s) return fail(ctx, RTQ_EINVAL, "raygen_diffuse: unknown scene %u", scene_id);
    if (n == 0) return RTQ_OK;
    if (!prim_dev || !hits_dev || !out_dev) return fail(ctx, RTQ_EINVAL, "raygen_diffuse: null pointer");
    RTQ_HIP(ctx, hipSetDevice(ctx->device));
    hipLaunchKernelGGL(k_diffuse, dim3(blocks_for(n)), dim3(kThreads), 0, pick_stream(ctx, stream), s->tri_world,
                       prim_dev, hits_dev, n, seed, eps, out_dev);
    RTQ_HIP(ctx, hipGetLastError());
    return RTQ_OK;
}

rtq_status rtq_raygen_shadow(rtq_ctx* ctx, uint32_t scene_id, const rtq_ray* prim_dev, const rtq_hit* hits_dev,
                             uint64_t n, const float* lights, uint32_t n_lights, float eps, rtq_ray* out_dev,
                             void* stream) {
    if (!ctx) return RTQ_EINVAL;
    Scene* s = find_scene(ctx, scene_id);
    if (!s) return fail(ctx, RTQ_EINVAL, "raygen_shadow: unknown scene %u", scene_id);
    if (n == 0 || n_lights == 0) return RTQ_OK;
    if (!prim_dev || !hits_dev || !out_dev || !lights) return fail(ctx, RTQ_EINVAL, "raygen_shadow: null pointer");
    if (n_lights > 16) return fail(ctx, RTQ_ECAPACITY, "raygen_shadow: at most 16 lights");
    Lights L{};
    for (uint32_t i = 0; i < 3 * n_lights; ++i) L.xyz[i] = lights[i];
    RTQ_HIP(ctx, hipSetDevice(ctx->device));
    hipLaunchKernelGGL(k_shadow, dim3(blocks_for(n)), dim3(kThreads), 0, pick_stream(ctx, stream), s->tri_world,
                       prim_dev, hits_dev, n, L, n_lights, eps, out_dev);
    RTQ_HIP(ctx, hipGetLastError());
    return RTQ_OK;
}

}  // extern "C"
