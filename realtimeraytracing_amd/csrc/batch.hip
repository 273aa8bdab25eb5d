// Batched traversal entry points, hipGraph capture of a batch, launch
// settings and the autotuner with its persistent cache.
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <vector>

#include "rtq_internal.hpp"

using namespace rtq;

namespace {

// FNV-1a, used for tuning keys and the cache file checksum.
uint64_t fnv1a(const void* data, size_t n, uint64_t h = 0xcbf29ce484222325ull) {
    const unsigned char* p = static_cast<const unsigned char*>(data);
    for (size_t i = 0; i < n; ++i) {
        h ^= p[i];
        h *= 0x100000001b3ull;
    }
    return h;
}

// Key of a batch for the tuning cache: device name, CU count, and per item
// the scene's triangle count and layouts, the ray count and the mode.
uint64_t tune_key(rtq_ctx* ctx, const rtq_batch_item* items, uint32_t n) {
    hipDeviceProp_t prop{};
    (void)hipGetDeviceProperties(&prop, ctx->device);
    uint64_t h = fnv1a(prop.gcnArchName, strnlen(prop.gcnArchName, sizeof prop.gcnArchName));
    h = fnv1a(&ctx->n_cu, sizeof ctx->n_cu, h);
    for (uint32_t i = 0; i < n; ++i) {
        const Scene* s = find_scene(ctx, items[i].scene_id);
        // learned decisions decide the layout of learned-order items, so they key the settings too
        const uint64_t order = s && s->ordered ? (s->order_closest ? 1u : 0u) | (s->order_any ? 2u : 0u) : 0u;
        const uint64_t rec[5] = {s ? s->n_tris : 0u, s ? (uint64_t)(s->wide != nullptr) : 0u, items[i].n_rays,
                                 items[i].flags, order};
        h = fnv1a(rec, sizeof rec, h);
    }
    return h;
}

bool valid_params(const rtq_launch_params& p) {
    return p.refill >= 1 && p.refill <= 64 && p.chunk >= 1 && p.chunk <= 4096 && p.prefer_wide <= 1 &&
           p.spec_slack <= 63 && p.leaf_slack <= 63 && valid_stack_lds(p.stack_lds);
}

void apply(rtq_ctx* ctx, const rtq_launch_params& p) {
    ctx->refill = p.refill;
    ctx->chunk = p.chunk;
    ctx->prefer_wide = p.prefer_wide != 0;
    ctx->spec_slack = p.spec_slack;
    ctx->leaf_slack = p.leaf_slack;
    ctx->stack_lds = p.stack_lds;
}

rtq_launch_params current(const rtq_ctx* ctx) {
    rtq_launch_params p{};
    p.refill = ctx->refill;
    p.chunk = ctx->chunk;
    p.prefer_wide = ctx->prefer_wide ? 1u : 0u;
    p.spec_slack = ctx->spec_slack;
    p.leaf_slack = ctx->leaf_slack;
    p.stack_lds = ctx->stack_lds;
    return p;
}

constexpr uint32_t kTuneMagic = 0x4e555452u;  // "RTUN"
constexpr uint32_t kTuneVersion = 3;

struct TuneRecord {
    uint64_t key;
    rtq_launch_params params;
};

}  // namespace

extern "C" {

rtq_status rtq_trace_batch(rtq_ctx* ctx, const rtq_batch_item* items, uint32_t n_items, void* stream) {
    if (!ctx) return RTQ_EINVAL;
    RTQ_HIP(ctx, hipSetDevice(ctx->device));
    return trace_batch(ctx, items, n_items, nullptr, false, pick_stream(ctx, stream));
}

rtq_status rtq_batch_create(rtq_ctx* ctx, const rtq_batch_item* items, uint32_t n_items, rtq_batch** out) {
    if (!ctx) return RTQ_EINVAL;
    if (!out || !items || n_items == 0) return fail(ctx, RTQ_EINVAL, "batch_create: no items");
    *out = nullptr;
    RTQ_HIP(ctx, hipSetDevice(ctx->device));
    rtq_batch* b = new rtq_batch();
    b->ctx = ctx;
    b->items.assign(items, items + n_items);
    auto bail = [&](rtq_status s) {
        rtq_batch_destroy(b);
        return s;
    };
    rtq_status s = scratch_alloc(ctx, &b->scratch);
    if (s != RTQ_OK) return bail(s);
    if (hipStreamCreateWithFlags(&b->capture, hipStreamNonBlocking) != hipSuccess)
        return bail(fail(ctx, RTQ_EHIP, "batch_create: stream"));
    // size the batch's scratch (validates the items too), then capture one launch
    s = trace_batch(ctx, items, n_items, &b->scratch, true, b->capture);
    if (s != RTQ_OK) return bail(s);
    hipError_t e = hipStreamBeginCapture(b->capture, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) return bail(hip_fail(ctx, e, "hipStreamBeginCapture"));
    s = trace_batch(ctx, items, n_items, &b->scratch, false, b->capture);
    hipGraph_t g = nullptr;
    e = hipStreamEndCapture(b->capture, &g);
    b->graph = g;
    if (s != RTQ_OK) return bail(s);
    if (e != hipSuccess) return bail(hip_fail(ctx, e, "hipStreamEndCapture"));
    e = hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0);
    if (e != hipSuccess) return bail(hip_fail(ctx, e, "hipGraphInstantiate"));
    for (uint32_t i = 0; i < n_items; ++i) b->generations.push_back(find_scene(ctx, items[i].scene_id)->generation);
    *out = b;
    return RTQ_OK;
}

rtq_status rtq_batch_launch(rtq_batch* b, void* stream) {
    if (!b || !b->exec) return RTQ_EINVAL;
    rtq_ctx* ctx = b->ctx;
    // the graph holds the node, triangle and spill buffers of the builds it
    // was captured on: a rebuild or release frees them, an edit without a
    // refit leaves the boxes behind the vertices
    for (size_t i = 0; i < b->items.size(); ++i) {
        const Scene* s = find_scene(ctx, b->items[i].scene_id);
        if (s && s->stale && s->generation == b->generations[i])
            return fail(ctx, RTQ_ENOTBUILT, "batch_launch: scene %u was edited and not refitted", b->items[i].scene_id);
        if (!s || !s->built || s->generation != b->generations[i])
            return fail(ctx, RTQ_ENOTBUILT, "batch_launch: scene %u was rebuilt or released after the batch was created",
                        b->items[i].scene_id);
    }
    RTQ_HIP(ctx, hipSetDevice(ctx->device));
    RTQ_HIP(ctx, hipGraphLaunch(b->exec, pick_stream(ctx, stream)));
    return RTQ_OK;
}

void rtq_batch_destroy(rtq_batch* b) {
    if (!b) return;
    (void)hipSetDevice(b->ctx->device);
    (void)hipDeviceSynchronize();
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    if (b->capture) (void)hipStreamDestroy(b->capture);
    scratch_free(&b->scratch);
    delete b;
}

rtq_status rtq_launch_params_get(rtq_ctx* ctx, rtq_launch_params* out) {
    if (!ctx || !out) return RTQ_EINVAL;
    *out = current(ctx);
    return RTQ_OK;
}

rtq_status rtq_launch_params_set(rtq_ctx* ctx, const rtq_launch_params* p) {
    if (!ctx || !p) return RTQ_EINVAL;
    if (!valid_params(*p))
        return fail(ctx, RTQ_EINVAL,
                    "launch_params: refill 1..64, chunk 1..4096, prefer_wide 0|1, slacks 0..63, stack_lds 0 or one of "
                    "8, 12, 14 (binary), 8, 16, 20 (8-wide)");
    apply(ctx, *p);
    return RTQ_OK;
}

rtq_status rtq_autotune(rtq_ctx* ctx, const rtq_batch_item* items, uint32_t n_items, uint32_t reps,
                        rtq_tune_result* out) {
    if (!ctx) return RTQ_EINVAL;
    if (!items || n_items == 0 || !out) return fail(ctx, RTQ_EINVAL, "autotune: no items");
    RTQ_HIP(ctx, hipSetDevice(ctx->device));
    memset(out, 0, sizeof *out);
    const uint64_t key = tune_key(ctx, items, n_items);
    auto hit = ctx->tuned.find(key);
    if (hit != ctx->tuned.end()) {
        apply(ctx, hit->second);
        out->best = hit->second;
        out->from_cache = 1;
        return RTQ_OK;
    }
    if (reps == 0) reps = 3;
    hipStream_t st = ctx->stream;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    RTQ_HIP(ctx, hipEventCreate(&e0));
    RTQ_HIP(ctx, hipEventCreate(&e1));
    const rtq_launch_params before = current(ctx);
    bool any_wide = false;
    for (uint32_t i = 0; i < n_items; ++i) {
        const Scene* s = find_scene(ctx, items[i].scene_id);
        if (s && s->wide) any_wide = true;
    }
    auto time_params = [&](const rtq_launch_params& p, float* ms) -> rtq_status {
        apply(ctx, p);
        rtq_status s = trace_batch(ctx, items, n_items, nullptr, false, st);  // warm-up (sizes scratch)
        if (s != RTQ_OK) return s;
        RTQ_HIP(ctx, hipEventRecord(e0, st));
        for (uint32_t r = 0; r < reps && s == RTQ_OK; ++r) s = trace_batch(ctx, items, n_items, nullptr, false, st);
        if (s != RTQ_OK) return s;
        RTQ_HIP(ctx, hipEventRecord(e1, st));
        RTQ_HIP(ctx, hipEventSynchronize(e1));
        RTQ_HIP(ctx, hipEventElapsedTime(ms, e0, e1));
        *ms /= (float)reps;
        return RTQ_OK;
    };
    rtq_status status = RTQ_OK;
    float base_ms = 0.0f;
    status = time_params(before, &base_ms);
    rtq_launch_params best = before;
    float best_ms = base_ms;
    uint32_t tried = 1;
    std::vector<std::pair<float, rtq_launch_params>> seen{{base_ms, before}};
    // stage 1: refill x chunk x node layout with the slacks in force; stage 2:
    // both slacks with the best stage-1 settings
    auto consider = [&](const rtq_launch_params& p) -> rtq_status {
        if (memcmp(&p, &before, sizeof p) == 0) return RTQ_OK;
        float ms = 0.0f;
        const rtq_status s = time_params(p, &ms);
        if (s != RTQ_OK) return s;
        ++tried;
        seen.emplace_back(ms, p);
        if (ms < best_ms) {
            best_ms = ms;
            best = p;
        }
        return RTQ_OK;
    };
    const uint32_t refills[] = {4, 8, 16, 32};
    const uint32_t chunks[] = {128, 256, 512};
    for (uint32_t wide = 0; wide <= (any_wide ? 1u : 0u) && status == RTQ_OK; ++wide)
        for (uint32_t rf : refills)
            for (uint32_t ch : chunks)
                if (status == RTQ_OK)
                    status = consider(rtq_launch_params{rf, ch, wide, before.spec_slack, before.leaf_slack,
                                                        before.stack_lds});
    const rtq_launch_params stage1 = best;
    const uint32_t spec[] = {0, 4, 8, 16, 24, 32};
    const uint32_t leaf[] = {0, 8, 16, 24};
    for (uint32_t sl : spec)
        for (uint32_t ll : leaf)
            if (status == RTQ_OK && !(sl == stage1.spec_slack && ll == stage1.leaf_slack))
                status = consider(rtq_launch_params{stage1.refill, stage1.chunk, stage1.prefer_wide, sl, ll,
                                                    stage1.stack_lds});
    // stage 3: the LDS stack depths of the layout the best settings trace
    // (stored explicitly, so the result names the depth)
    if (status == RTQ_OK) {
        const rtq_launch_params stage2 = best;
        apply(ctx, stage2);
        int layout = -1;
        status = batch_layout(ctx, items, n_items, &layout);
        const uint32_t* choices = (layout == 1 || layout == 3) ? kStackChoicesWide : kStackChoicesBinary;
        for (uint32_t k = 0; k < 3 && status == RTQ_OK; ++k) {
            rtq_launch_params p = stage2;
            p.stack_lds = choices[k];
            if (k == 1 && (stage2.stack_lds == 0 || stage2.stack_lds == choices[1])) {
                // the depth in force: relabel its timings instead of timing it again
                for (auto& sp : seen)
                    if (memcmp(&sp.second, &stage2, sizeof stage2) == 0) sp.second.stack_lds = choices[1];
                if (memcmp(&best, &stage2, sizeof best) == 0) best.stack_lds = choices[1];
                continue;
            }
            if (p.stack_lds != stage2.stack_lds) status = consider(p);
        }
    }
    // stage 4: one timing of a ~1 ms launch spreads by a few percent, enough
    // to misorder close settings, so the four fastest run again in three
    // interleaved rounds and each keeps its lowest time
    if (status == RTQ_OK) {
        std::stable_sort(seen.begin(), seen.end(),
                         [](const auto& a, const auto& b) { return a.first < b.first; });
        const size_t k = std::min<size_t>(4, seen.size());
        std::vector<float> low(k, INFINITY);
        for (int round = 0; round < 3 && status == RTQ_OK; ++round)
            for (size_t i = 0; i < k && status == RTQ_OK; ++i) {
                float ms = 0.0f;
                status = time_params(seen[i].second, &ms);
                low[i] = std::min(low[i], ms);
            }
        if (status == RTQ_OK) {
            const size_t w = (size_t)(std::min_element(low.begin(), low.end()) - low.begin());
            best = seen[w].second;
            best_ms = low[w];
        }
    }
    if (status == RTQ_OK && best.stack_lds == 0) {
        // name the default depth of the layout the chosen settings trace
        apply(ctx, best);
        int layout = -1;
        status = batch_layout(ctx, items, n_items, &layout);
        best.stack_lds = (layout == 1 || layout == 3) ? kStackChoicesWide[1] : kStackChoicesBinary[1];
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (status != RTQ_OK) {
        apply(ctx, before);
        return status;
    }
    apply(ctx, best);
    ctx->tuned[key] = best;
    out->best = best;
    out->best_ms = best_ms;
    out->default_ms = base_ms;
    out->n_tried = tried;
    return RTQ_OK;
}

rtq_status rtq_tune_save(rtq_ctx* ctx, const char* path) {
    if (!ctx || !path) return RTQ_EINVAL;
    std::vector<TuneRecord> recs;
    for (const auto& kv : ctx->tuned) recs.push_back({kv.first, kv.second});
    const uint32_t head[3] = {kTuneMagic, kTuneVersion, (uint32_t)recs.size()};
    const uint64_t sum = fnv1a(recs.data(), recs.size() * sizeof(TuneRecord), fnv1a(head, sizeof head));
    FILE* f = fopen(path, "wb");
    if (!f) return fail(ctx, RTQ_EINVAL, "tune_save: cannot open %s", path);
    bool ok = fwrite(head, sizeof head, 1, f) == 1;
    if (!recs.empty()) ok = ok && fwrite(recs.data(), sizeof(TuneRecord), recs.size(), f) == recs.size();
    ok = ok && fwrite(&sum, sizeof sum, 1, f) == 1;
    ok = (fclose(f) == 0) && ok;
    return ok ? RTQ_OK : fail(ctx, RTQ_EINVAL, "tune_save: write failed for %s", path);
}

rtq_status rtq_tune_load(rtq_ctx* ctx, const char* path) {
    if (!ctx || !path) return RTQ_EINVAL;
    FILE* f = fopen(path, "rb");
    if (!f) return fail(ctx, RTQ_EINVAL, "tune_load: cannot open %s", path);
    uint32_t head[3] = {0, 0, 0};
    std::vector<TuneRecord> recs;
    uint64_t sum = 0;
    bool ok = fread(head, sizeof head, 1, f) == 1 && head[0] == kTuneMagic && head[1] == kTuneVersion &&
              head[2] <= (1u << 20);
    if (ok) {
        recs.resize(head[2]);
        if (!recs.empty()) ok = fread(recs.data(), sizeof(TuneRecord), recs.size(), f) == recs.size();
        ok = ok && fread(&sum, sizeof sum, 1, f) == 1;
    }
    fclose(f);
    if (!ok) return fail(ctx, RTQ_EINVAL, "tune_load: %s is not a version-%u tuning file", path, kTuneVersion);
    if (sum != fnv1a(recs.data(), recs.size() * sizeof(TuneRecord), fnv1a(head, sizeof head)))
        return fail(ctx, RTQ_EINVAL, "tune_load: checksum mismatch in %s", path);
    for (const TuneRecord& r : recs)
        if (!valid_params(r.params)) return fail(ctx, RTQ_EINVAL, "tune_load: invalid record in %s", path);
    for (const TuneRecord& r : recs) ctx->tuned[r.key] = r.params;
    return RTQ_OK;
}

}  // extern "C"
