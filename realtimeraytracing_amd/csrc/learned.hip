// Learned child ordering for the binary traversal layout.
//
// A 32 -> 64 -> 64 -> 1 MLP (ReLU, ReLU, logistic) predicts, for a node, a ray
// direction octant and a mode (closest / any hit), the probability that
// entering child 0 first costs less work (node visits + triangle tests until
// the closest hit is certain, or until an occluder is found).
//  - Logs: k_order_log walks sampled rays with the plain traversal and emits,
//    at nodes where both children are hit, the node's features, the label and
//    a weight.
//  - Training: k_mlp_grad (forward + backward of 64 samples per block, FP32,
//    per-block partial gradients), k_mlp_reduce (deterministic reduction in
//    block order) and k_mlp_adam, all on the GPU.
//  - Inference: k_order_apply evaluates every (node, octant, mode) with BF16
//    MFMA tiles (v_mfma_f32_32x32x16_bf16, FP32 accumulation) and stores 16
//    decision bits and 16 confidence bits in the node's spare word; with
//    RTQ_TRACE_LEARNED the traversal takes the predicted child first where the
//    confidence bit is set and the nearer child elsewhere.  Only the visit
//    order changes, so hit records are exactly those of the plain traversal.
#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstring>
#include <vector>

#include "rtq_internal.hpp"

namespace rtq {
namespace {

constexpr int kIn = 32, kH = 64;
constexpr int kW1 = 0, kB1 = kW1 + kIn * kH, kW2 = kB1 + kH, kB2 = kW2 + kH * kH, kW3 = kB2 + kH, kB3 = kW3 + kH;
constexpr int kNP = kB3 + 1;      // 6337 parameters
constexpr int kSample = kIn + 2;  // features, label, weight
static_assert(kNP == RTQ_ORDER_PARAMS && kSample == RTQ_ORDER_SAMPLE, "model sizes");

__device__ __forceinline__ float area(float ex, float ey, float ez) { return 2.0f * (ex * ey + ey * ez + ez * ex); }

// Features of (node, octant, mode): octant bit k set <=> ray direction < 0 on axis k.
__device__ void node_features(float4 n0, float4 n1, float4 n2, float4 n3, uint4 aux, uint32_t octant, uint32_t any,
                              float* f) {
    const float l0[3] = {n0.x, n0.z, n1.x}, h0[3] = {n0.y, n0.w, n1.y};
    const float l1[3] = {n1.z, n2.x, n2.z}, h1[3] = {n1.w, n2.y, n2.w};
    float lo[3], ext[3], e = 1e-30f;
    for (int k = 0; k < 3; ++k) {
        lo[k] = fminf(l0[k], l1[k]);
        ext[k] = fmaxf(h0[k], h1[k]) - lo[k];
        e = fmaxf(e, ext[k]);
    }
    const float inv = 1.0f / e;
    float s[3];
    for (int k = 0; k < 3; ++k) s[k] = ((octant >> k) & 1u) ? -1.0f : 1.0f;
    f[0] = s[0];
    f[1] = s[1];
    f[2] = s[2];
    f[3] = any ? 1.0f : 0.0f;
    f[4] = (float)aux.w * (1.0f / 32.0f);
    float dot = 0.0f;
    for (int k = 0; k < 3; ++k) {
        f[5 + k] = (l0[k] - lo[k]) * inv;
        f[8 + k] = (h0[k] - lo[k]) * inv;
        f[11 + k] = (l1[k] - lo[k]) * inv;
        f[14 + k] = (h1[k] - lo[k]) * inv;
        const float dc = 0.5f * ((l1[k] + h1[k]) - (l0[k] + h0[k])) * inv;
        f[17 + k] = dc;
        dot += dc * s[k];
        f[27 + k] = ext[k] * inv;
    }
    f[20] = dot;
    const float sp = fmaxf(area(ext[0], ext[1], ext[2]), 1e-30f);
    f[21] = area(h0[0] - l0[0], h0[1] - l0[1], h0[2] - l0[2]) / sp;
    f[22] = area(h1[0] - l1[0], h1[1] - l1[1], h1[2] - l1[2]) / sp;
    f[23] = log2f((float)(aux.y - aux.x) + 1.0f) * (1.0f / 24.0f);
    f[24] = log2f((float)(aux.z - aux.y) + 1.0f) * (1.0f / 24.0f);
    f[25] = (__float_as_uint(n3.x) & kLeafBit) ? 1.0f : 0.0f;
    f[26] = (__float_as_uint(n3.y) & kLeafBit) ? 1.0f : 0.0f;
    f[30] = 0.0f;
    f[31] = 0.0f;
}

__device__ __forceinline__ float sel3(float4 v, int k) { return k == 0 ? v.x : (k == 1 ? v.y : v.z); }

// Watertight hit test returning t (the same test as the traversal kernel).
__device__ bool wt_hit(float4 p0, float4 p1, float4 p2, int kx, int ky, int kz, float opx, float opy, float opz,
                       float sx, float sy, float sz, float tmin, float tmax, float& t) {
    const float Akx = sel3(p0, kx) - opx, Aky = sel3(p0, ky) - opy, Akz = sel3(p0, kz) - opz;
    const float Bkx = sel3(p1, kx) - opx, Bky = sel3(p1, ky) - opy, Bkz = sel3(p1, kz) - opz;
    const float Ckx = sel3(p2, kx) - opx, Cky = sel3(p2, ky) - opy, Ckz = sel3(p2, kz) - opz;
    const float Ax = Akx - sx * Akz, Ay = Aky - sy * Akz;
    const float Bx = Bkx - sx * Bkz, By = Bky - sy * Bkz;
    const float Cx = Ckx - sx * Ckz, Cy = Cky - sy * Ckz;
    float U = Cx * By - Cy * Bx, V = Ax * Cy - Ay * Cx, W = Bx * Ay - By * Ax;
    if (U == 0.0f || V == 0.0f || W == 0.0f) {
        U = (float)((double)Cx * (double)By - (double)Cy * (double)Bx);
        V = (float)((double)Ax * (double)Cy - (double)Ay * (double)Cx);
        W = (float)((double)Bx * (double)Ay - (double)By * (double)Ax);
    }
    if ((U < 0.0f || V < 0.0f || W < 0.0f) && (U > 0.0f || V > 0.0f || W > 0.0f)) return false;
    const float det = (U + V) + W;
    if (det == 0.0f) return false;
    const float T = (U * (sz * Akz) + V * (sz * Bkz)) + W * (sz * Ckz);
    const float tt = T / det;
    if (!(tt >= tmin && tt <= tmax)) return false;
    t = tt;
    return true;
}

// Both children's box tests of a binary node (padded boxes, interval [t0, t1]).
__device__ __forceinline__ void child_hits(float4 n0, float4 n1, float4 n2, const float* o, const float* inv, float t0,
                                           float t1, bool& h0, bool& h1, float& tn0, float& tn1) {
    const float oi[3] = {o[0] * inv[0], o[1] * inv[1], o[2] * inv[2]};
    const float ax0 = fmaf(n0.x, inv[0], -oi[0]), bx0 = fmaf(n0.y, inv[0], -oi[0]);
    const float ay0 = fmaf(n0.z, inv[1], -oi[1]), by0 = fmaf(n0.w, inv[1], -oi[1]);
    const float az0 = fmaf(n1.x, inv[2], -oi[2]), bz0 = fmaf(n1.y, inv[2], -oi[2]);
    const float ax1 = fmaf(n1.z, inv[0], -oi[0]), bx1 = fmaf(n1.w, inv[0], -oi[0]);
    const float ay1 = fmaf(n2.x, inv[1], -oi[1]), by1 = fmaf(n2.y, inv[1], -oi[1]);
    const float az1 = fmaf(n2.z, inv[2], -oi[2]), bz1 = fmaf(n2.w, inv[2], -oi[2]);
    tn0 = fmaxf(fmaxf(fminf(ax0, bx0), fminf(ay0, by0)), fmaxf(fminf(az0, bz0), t0));
    const float tf0 = fminf(fminf(fmaxf(ax0, bx0), fmaxf(ay0, by0)), fminf(fmaxf(az0, bz0), t1));
    tn1 = fmaxf(fmaxf(fminf(ax1, bx1), fminf(ay1, by1)), fmaxf(fminf(az1, bz1), t0));
    const float tf1 = fminf(fminf(fmaxf(ax1, bx1), fmaxf(ay1, by1)), fminf(fmaxf(az1, bz1), t1));
    h0 = tn0 <= tf0;
    h1 = tn1 <= tf1;
}

constexpr int kLogStack = 96;
constexpr uint32_t kLogBudget = 1u << 14;  // node visits + triangle tests per sub-traversal

struct LogRay {
    float o[3], inv[3], tmin, tmax;
    int kx, ky, kz;
    float sx, sy, sz, opx, opy, opz;
};

// Traversal of the subtree at `ref` (nearer child first) over [R.tmin, tmax].
// Closest (any = false): returns the leaf-order index of the closest hit and
// shrinks tmax to its t.  Any hit: returns the first hit found.  kNone = no
// hit.  cost = node visits + triangle tests; ok = false when the walk ran past
// kLogBudget or its stack.
__device__ uint32_t log_walk(const float4* __restrict__ nodes, const float4* __restrict__ tris, const LogRay& R,
                             uint32_t ref, bool any, float& tmax, uint32_t& cost, bool& ok) {
    uint32_t stack[kLogStack];
    uint32_t sp = 0, cur = ref, target = kNone;
    cost = 0;
    ok = true;
    for (;;) {
        if (++cost > kLogBudget) {
            ok = false;
            return kNone;
        }
        if (!(cur & kLeafBit)) {
            const float4* nd = nodes + 4ull * cur;
            bool h0, h1;
            float tn0, tn1;
            child_hits(nd[0], nd[1], nd[2], R.o, R.inv, R.tmin, tmax, h0, h1, tn0, tn1);
            const float4 n3 = nd[3];
            const uint32_t r0 = __float_as_uint(n3.x), r1 = __float_as_uint(n3.y);
            if (h0 && h1) {
                if (sp >= kLogStack) {
                    ok = false;
                    return kNone;
                }
                const bool f0 = tn0 <= tn1;
                stack[sp++] = f0 ? r1 : r0;
                cur = f0 ? r0 : r1;
                continue;
            }
            if (h0 || h1) {
                cur = h0 ? r0 : r1;
                continue;
            }
        } else {
            const uint32_t first = (cur >> 3) & 0x0FFFFFFFu, cnt = (cur & 7u) + 1u;
            cost += cnt - 1;  // the leaf's own step counted one test
            for (uint32_t k = 0; k < cnt; ++k) {
                const float4* tp = tris + 3ull * (first + k);
                float t;
                if (wt_hit(tp[0], tp[1], tp[2], R.kx, R.ky, R.kz, R.opx, R.opy, R.opz, R.sx, R.sy, R.sz, R.tmin, tmax,
                           t) &&
                    (t < tmax || target == kNone)) {
                    tmax = t;
                    target = first + k;
                    if (any) return target;
                }
            }
        }
        if (sp == 0) return target;
        cur = stack[--sp];
    }
}

__device__ void emit(float* __restrict__ out, uint32_t cap, uint32_t* __restrict__ count, float4 n0, float4 n1,
                     float4 n2, float4 n3, uint4 ax4, uint32_t octant, uint32_t any, bool child0_first, float weight) {
    const uint32_t slot = atomicAdd(count, 1u);
    if (slot >= cap) return;
    float f[kIn];
    node_features(n0, n1, n2, n3, ax4, octant, any, f);
    float* o = out + (size_t)slot * kSample;
    for (int k = 0; k < kIn; ++k) o[k] = f[k];
    o[kIn] = child0_first ? 1.0f : 0.0f;
    o[kIn + 1] = weight;
}

// Samples of one ray every `stride`: from the root, each node whose two
// children are both hit is priced in both visit orders with sub-traversals
// (node visits + triangle tests until the closest hit is certain, or until an
// occluder is found for any-hit rays); label = the cheaper child first, weight
// log2(1 + cost difference), no sample on a tie.  The walk continues into the
// cheaper child and stops where no hit lies below it.
__global__ __launch_bounds__(256) void k_order_log(const float4* __restrict__ nodes, const uint4* __restrict__ aux,
                                                   const float4* __restrict__ tris, const rtq_ray* __restrict__ rays,
                                                   uint32_t n, uint32_t stride, uint32_t any, float* __restrict__ out,
                                                   uint32_t cap, uint32_t* __restrict__ count) {
    const uint64_t r = ((uint64_t)blockIdx.x * 256 + threadIdx.x) * stride;
    if (r >= n) return;
    const rtq_ray ray = rays[r];
    if (!(ray.tmin <= ray.tmax)) return;
    LogRay R;
    for (int k = 0; k < 3; ++k) {
        const float d = fabsf(ray.d[k]) < 1e-30f ? copysignf(1e-30f, ray.d[k]) : ray.d[k];
        R.o[k] = ray.o[k];
        R.inv[k] = 1.0f / d;
    }
    R.tmin = ray.tmin;
    R.tmax = ray.tmax;
    const float ax = fabsf(ray.d[0]), ay = fabsf(ray.d[1]), az = fabsf(ray.d[2]);
    int kz = 0;
    float m = ax;
    if (ay > m) { kz = 1; m = ay; }
    if (az > m) kz = 2;
    int kx = kz + 1 == 3 ? 0 : kz + 1;
    int ky = kx + 1 == 3 ? 0 : kx + 1;
    const float dkz = ray.d[kz];
    if (dkz == 0.0f) return;
    if (dkz < 0.0f) { const int tmp = kx; kx = ky; ky = tmp; }
    R.kx = kx;
    R.ky = ky;
    R.kz = kz;
    R.sx = ray.d[kx] / dkz;
    R.sy = ray.d[ky] / dkz;
    R.sz = 1.0f / dkz;
    R.opx = ray.o[kx];
    R.opy = ray.o[ky];
    R.opz = ray.o[kz];
    const uint32_t octant = (ray.d[0] < 0.0f ? 1u : 0u) | (ray.d[1] < 0.0f ? 2u : 0u) | (ray.d[2] < 0.0f ? 4u : 0u);
    // Walk from the root along the child each decision visits first: nothing
    // has been visited before such a node, so its interval is the ray's own.
    uint32_t cur = 0;
    while (!(cur & kLeafBit)) {
        const float4* nd = nodes + 4ull * cur;
        const float4 n0 = nd[0], n1 = nd[1], n2 = nd[2], n3 = nd[3];
        bool h0, h1;
        float tn0, tn1;
        child_hits(n0, n1, n2, R.o, R.inv, R.tmin, R.tmax, h0, h1, tn0, tn1);
        const uint32_t r0 = __float_as_uint(n3.x), r1 = __float_as_uint(n3.y);
        if (h0 && h1) {
            uint32_t cost01, cost10;
            bool f0, f1;
            if (any) {
                // an occluder below the first child ends the ray there
                uint32_t c0, c1;
                bool ok0, ok1;
                float t0 = R.tmax, t1 = R.tmax;
                f0 = log_walk(nodes, tris, R, r0, true, t0, c0, ok0) != kNone;
                f1 = log_walk(nodes, tris, R, r1, true, t1, c1, ok1) != kNone;
                if (!ok0 || !ok1) return;
                cost01 = f0 ? c0 : c0 + c1;
                cost10 = f1 ? c1 : c1 + c0;
            } else {
                // closest hit: the second child is walked with the interval the first left
                uint32_t a0, a1, b0, b1;
                bool ok[4];
                float ta = R.tmax, tb = R.tmax;
                f0 = log_walk(nodes, tris, R, r0, false, ta, a0, ok[0]) != kNone;
                log_walk(nodes, tris, R, r1, false, ta, a1, ok[1]);
                f1 = log_walk(nodes, tris, R, r1, false, tb, b1, ok[2]) != kNone;
                log_walk(nodes, tris, R, r0, false, tb, b0, ok[3]);
                if (!ok[0] || !ok[1] || !ok[2] || !ok[3]) return;
                cost01 = a0 + a1;
                cost10 = b1 + b0;
            }
            // with no hit below either child every order visits the same nodes
            if (!f0 && !f1) return;
            if (cost01 != cost10) {
                const float diff = (float)(cost01 > cost10 ? cost01 - cost10 : cost10 - cost01);
                emit(out, cap, count, n0, n1, n2, n3, aux[cur], octant, any, cost01 < cost10, log2f(1.0f + diff));
            }
            const bool go0 = cost01 <= cost10;
            if (!(go0 ? f0 : f1)) return;
            cur = go0 ? r0 : r1;
        } else if (h0 || h1) {
            cur = h0 ? r0 : r1;
        } else {
            return;
        }
    }
}

// ---- training ----------------------------------------------------------------

constexpr int kTB = 64;  // samples per block

// Forward + backward of up to 64 samples: sample i of the step is
// samples[(start + i) * mul % n]; writes the block's gradient sums and its sum
// of weighted losses.  Loss of a sample: weight * BCE(sigmoid(z), label).
__global__ __launch_bounds__(kTB) void k_mlp_grad(const float* __restrict__ P, const float* __restrict__ samples,
                                                  uint32_t n, uint64_t start, uint32_t count, uint64_t mul,
                                                  float* __restrict__ partial, float* __restrict__ loss_part) {
    __shared__ float sW2[kH * kH];
    // rows padded to kTB + 1 so a thread walking its own row hits its own bank
    __shared__ float sX[kIn][kTB + 1];
    __shared__ float sH1[kH][kTB + 1];
    __shared__ float sH2[kH][kTB + 1];
    __shared__ float sDz[kTB];
    const int t = threadIdx.x;
    for (int i = t; i < kH * kH; i += kTB) sW2[i] = P[kW2 + i];
    const uint32_t local = blockIdx.x * kTB + t;
    const bool valid = local < count;
    const uint64_t idx = ((start + local) * mul) % n;
    for (int k = 0; k < kIn; ++k) sX[k][t] = valid ? samples[idx * kSample + k] : 0.0f;
    const float y = valid ? samples[idx * kSample + kIn] : 0.0f;
    const float w = valid ? samples[idx * kSample + kIn + 1] : 0.0f;
    __syncthreads();
    // forward (one sample per thread)
    for (int j = 0; j < kH; ++j) {
        float acc = P[kB1 + j];
        for (int k = 0; k < kIn; ++k) acc = fmaf(sX[k][t], P[kW1 + k * kH + j], acc);
        sH1[j][t] = fmaxf(acc, 0.0f);
    }
    float z = P[kB3];
    for (int j = 0; j < kH; ++j) {
        float acc = P[kB2 + j];
        for (int k = 0; k < kH; ++k) acc = fmaf(sH1[k][t], sW2[k * kH + j], acc);
        const float h = fmaxf(acc, 0.0f);
        sH2[j][t] = h;
        z = fmaf(h, P[kW3 + j], z);
    }
    // BCE with logits: max(z, 0) - z y + log(1 + e^-|z|); dL/dz = w (sigmoid(z) - y)
    const float p = 1.0f / (1.0f + expf(-z));
    const float loss = w * ((fmaxf(z, 0.0f) - z * y) + log1pf(expf(-fabsf(z))));
    sDz[t] = valid ? w * (p - y) : 0.0f;
    __syncthreads();
    float* g = partial + (size_t)blockIdx.x * kNP;
    // output layer: dw3[j] = sum_s h2[s][j] dz[s]; db3 = sum dz (thread t owns j = t)
    {
        float acc = 0.0f;
        for (int s = 0; s < kTB; ++s) acc = fmaf(sH2[t][s], sDz[s], acc);
        g[kW3 + t] = acc;
        if (t == 0) {
            float a = 0.0f;
            for (int s = 0; s < kTB; ++s) a += sDz[s];
            g[kB3] = a;
        }
    }
    __syncthreads();
    // dH2 = dz * w3 * (h2 > 0), in place of H2 (thread t owns sample t)
    for (int j = 0; j < kH; ++j) sH2[j][t] = sH2[j][t] > 0.0f ? sDz[t] * P[kW3 + j] : 0.0f;
    __syncthreads();
    // dW2[k][j] = sum_s h1[s][k] dh2[s][j]; db2[j] = sum_s dh2[s][j]   (thread t owns column j = t)
    {
        float db = 0.0f;
        for (int s = 0; s < kTB; ++s) db += sH2[t][s];
        g[kB2 + t] = db;
        for (int k = 0; k < kH; ++k) {
            float acc = 0.0f;
            for (int s = 0; s < kTB; ++s) acc = fmaf(sH1[k][s], sH2[t][s], acc);
            g[kW2 + k * kH + t] = acc;
        }
    }
    __syncthreads();
    // dH1 = (dH2 W2^T) * (h1 > 0), in place of H1 (thread t owns sample t)
    for (int k = 0; k < kH; ++k) {
        float acc = 0.0f;
        for (int j = 0; j < kH; ++j) acc = fmaf(sH2[j][t], sW2[k * kH + j], acc);
        sH1[k][t] = sH1[k][t] > 0.0f ? acc : 0.0f;
    }
    __syncthreads();
    // dW1[k][j] = sum_s x[s][k] dh1[s][j]; db1[j] = sum_s dh1[s][j]   (thread t owns column j = t)
    {
        float db = 0.0f;
        for (int s = 0; s < kTB; ++s) db += sH1[t][s];
        g[kB1 + t] = db;
        for (int k = 0; k < kIn; ++k) {
            float acc = 0.0f;
            for (int s = 0; s < kTB; ++s) acc = fmaf(sX[k][s], sH1[t][s], acc);
            g[kW1 + k * kH + t] = acc;
        }
    }
    // loss sum of the block
    __shared__ float sL[kTB];
    sL[t] = valid ? loss : 0.0f;
    __syncthreads();
    if (t == 0) {
        float a = 0.0f;
        for (int s = 0; s < kTB; ++s) a += sL[s];
        loss_part[blockIdx.x] = a;
    }
}

// Sums the blocks' gradients in block order (deterministic), scales by 1/batch.
__global__ void k_mlp_reduce(const float* __restrict__ partial, uint32_t blocks, float scale, float* __restrict__ grad) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= kNP) return;
    float acc = 0.0f;
    for (uint32_t b = 0; b < blocks; ++b) acc += partial[(size_t)b * kNP + j];
    grad[j] = acc * scale;
}

__global__ void k_mlp_adam(float* __restrict__ P, float* __restrict__ M, float* __restrict__ V,
                           const float* __restrict__ grad, float lr, float b1, float b2, float eps, float c1, float c2) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= kNP) return;
    const float g = grad[j];
    const float m = b1 * M[j] + (1.0f - b1) * g;
    const float v = b2 * V[j] + (1.0f - b2) * g * g;
    M[j] = m;
    V[j] = v;
    P[j] -= lr * (m / c1) / (sqrtf(v / c2) + eps);
}

// ---- BF16 MFMA inference ---------------------------------------------------------

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(16 * sizeof(float)))) float f32x16;

__device__ __forceinline__ uint16_t to_bf16(float f) {
    uint32_t u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);  // round to nearest even (finite inputs)
    return (uint16_t)(u >> 16);
}

__device__ __forceinline__ bf16x8 load8(const uint16_t* p) {
    const uint4 v = *reinterpret_cast<const uint4*>(p);
    return __builtin_bit_cast(bf16x8, v);
}

constexpr int kIW = 4;  // waves per inference block

// The bf16 operand arrays come first and have sizes that are multiples of 16 B,
// so every 8-element (16-B) operand load is 16-B aligned.
struct alignas(16) InferLds {
    uint16_t w1t[kH][kIn];       // W1 transposed (hidden, feature), bf16
    uint16_t w2t[kH][kH];        // W2 transposed (out, in), bf16
    uint16_t x[kIW][32][kIn];    // per wave: 32 samples x 32 features
    uint16_t h1[kIW][32][kH];    // per wave: layer-1 activations
    float red[kIW][32][33];      // per wave: layer-3 partial products
    float b1[kH], b2[kH], w3[kH];
    float b3;
};
static_assert(offsetof(InferLds, x) % 16 == 0 && offsetof(InferLds, h1) % 16 == 0, "operand alignment");

__device__ void load_weights(InferLds& L, const float* P) {
    for (int i = threadIdx.x; i < kIn * kH; i += blockDim.x) {
        const int k = i / kH, j = i % kH;
        L.w1t[j][k] = to_bf16(P[kW1 + i]);
    }
    for (int i = threadIdx.x; i < kH * kH; i += blockDim.x) {
        const int k = i / kH, j = i % kH;
        L.w2t[j][k] = to_bf16(P[kW2 + i]);
    }
    for (int i = threadIdx.x; i < kH; i += blockDim.x) {
        L.b1[i] = P[kB1 + i];
        L.b2[i] = P[kB2 + i];
        L.w3[i] = P[kW3 + i];
    }
    if (threadIdx.x == 0) L.b3 = P[kB3];
}

// The wave's 32 samples (rows of L.x[w]) through the network; returns the
// logit of sample (lane & 31) in every lane.
// 32x32x16 operand layout: A row m = lane % 32 holds k = 8 * (lane / 32) + 0..7;
// B column n = lane % 32 likewise; D element (m, n) of register r sits in lane
// n + 32 * ((m / 4) % 2) with r = (m % 4) + 4 * (m / 8).
__device__ float mlp_tile(InferLds& L, int w, int lane) {
    const int row = lane & 31, half = lane >> 5;
    f32x16 acc[2] = {};
#pragma unroll
    for (int ks = 0; ks < kIn / 16; ++ks) {
        const bf16x8 a = load8(&L.x[w][row][16 * ks + 8 * half]);
#pragma unroll
        for (int tn = 0; tn < 2; ++tn) {
            const bf16x8 b = load8(&L.w1t[32 * tn + row][16 * ks + 8 * half]);
            acc[tn] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc[tn], 0, 0, 0);
        }
    }
#pragma unroll
    for (int tn = 0; tn < 2; ++tn) {
        const int ncol = 32 * tn + row;
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const int m = (r & 3) + 4 * half + 8 * (r >> 2);
            L.h1[w][m][ncol] = to_bf16(fmaxf(acc[tn][r] + L.b1[ncol], 0.0f));
        }
    }
    __syncthreads();
    f32x16 acc2[2] = {};
#pragma unroll
    for (int ks = 0; ks < kH / 16; ++ks) {
        const bf16x8 a = load8(&L.h1[w][row][16 * ks + 8 * half]);
#pragma unroll
        for (int tn = 0; tn < 2; ++tn) {
            const bf16x8 b = load8(&L.w2t[32 * tn + row][16 * ks + 8 * half]);
            acc2[tn] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc2[tn], 0, 0, 0);
        }
    }
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        const int m = (r & 3) + 4 * half + 8 * (r >> 2);
        float v = 0.0f;
#pragma unroll
        for (int tn = 0; tn < 2; ++tn) {
            const int ncol = 32 * tn + row;
            v = fmaf(fmaxf(acc2[tn][r] + L.b2[ncol], 0.0f), L.w3[ncol], v);
        }
        L.red[w][m][row] = v;
    }
    __syncthreads();
    float z = L.b3;
    for (int c = 0; c < 32; ++c) z += L.red[w][row][c];
    return z;
}

// Probabilities for M feature rows (fp32 in, fp32 out): the MFMA path on any samples.
__global__ __launch_bounds__(64 * kIW) void k_mlp_infer(const float* __restrict__ P, const float* __restrict__ X,
                                                        uint32_t M, uint32_t stride, float* __restrict__ prob) {
    __shared__ InferLds L;
    load_weights(L, P);
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const uint32_t tiles = (M + 31) / 32;
    const uint32_t per_iter = gridDim.x * kIW;
    // every wave of the block runs the same number of iterations (barriers inside)
    const uint32_t iters = (tiles + per_iter - 1) / per_iter;
    __syncthreads();
    for (uint32_t it = 0; it < iters; ++it) {
        const uint32_t tile = it * per_iter + blockIdx.x * kIW + w;
        if (lane < 32) {
            const uint32_t sidx = tile * 32 + lane;
            for (int k = 0; k < kIn; ++k)
                L.x[w][lane][k] = to_bf16((tile < tiles && sidx < M) ? X[(size_t)sidx * stride + k] : 0.0f);
        }
        __syncthreads();
        const float z = mlp_tile(L, w, lane);
        const uint32_t sidx = tile * 32 + lane;
        if (lane < 32 && tile < tiles && sidx < M) prob[sidx] = 1.0f / (1.0f + expf(-z));
        __syncthreads();
    }
}

// Order word of every binary node: 16 samples per node (8 octants x 2 modes);
// bit (any * 8 + octant) set when child 0 should go first, bit 16 + (any * 8 +
// octant) set when the logit's magnitude reaches zmin[any] (the decision is
// used; an infinite zmin never sets it).
__global__ __launch_bounds__(64 * kIW) void k_order_apply(const float* __restrict__ P, float4* __restrict__ nodes,
                                                          const uint4* __restrict__ aux, uint32_t n_nodes,
                                                          float zmin_closest, float zmin_any) {
    __shared__ InferLds L;
    load_weights(L, P);
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const uint32_t tiles = (n_nodes + 1) / 2;  // 2 nodes x 16 samples per tile
    const uint32_t per_iter = gridDim.x * kIW;
    const uint32_t iters = (tiles + per_iter - 1) / per_iter;
    __syncthreads();
    for (uint32_t it = 0; it < iters; ++it) {
        const uint32_t tile = it * per_iter + blockIdx.x * kIW + w;
        const uint32_t node = tile * 2 + ((lane >> 4) & 1);
        const bool live = lane < 32 && tile < tiles && node < n_nodes;
        if (lane < 32) {
            float f[kIn];
            if (live) {
                const float4* nd = nodes + 4ull * node;
                node_features(nd[0], nd[1], nd[2], nd[3], aux[node], lane & 7, (lane >> 3) & 1, f);
            } else {
                for (int k = 0; k < kIn; ++k) f[k] = 0.0f;
            }
            for (int k = 0; k < kIn; ++k) L.x[w][lane][k] = to_bf16(f[k]);
        }
        __syncthreads();
        const float z = mlp_tile(L, w, lane);
        // a NaN logit fails both tests: the node keeps the nearer-child order
        const uint64_t bits = __ballot(live && z > 0.0f);
        const uint64_t sure = __ballot(live && fabsf(z) >= (((lane >> 3) & 1) ? zmin_any : zmin_closest));
        if ((lane == 0 || lane == 16) && live) {
            const uint32_t word = ((uint32_t)(bits >> lane) & 0xFFFFu) | (((uint32_t)(sure >> lane) & 0xFFFFu) << 16);
            float4* n3 = nodes + 4ull * node + 3;
            n3->z = __uint_as_float(word);
        }
        __syncthreads();
    }
}

__global__ void k_order_clear(float4* __restrict__ nodes, uint32_t n_nodes) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i < n_nodes) nodes[4ull * i + 3].z = 0.0f;
}

// Learned first child of every 8-wide node, per (octant, mode): from the
// binary node of the wide node's root cluster, follow the binary decisions
// (the child order word's decided child) while they are confident, until the
// path reaches one of the wide node's children; that slot is entry
// octant + 8 * any_hit of the table (8 | slot; 0 where a decision is missing).
// The children of a wide node are the frontier of the binary subtree below its
// root, so every decided path reaches one within 7 steps.
__global__ __launch_bounds__(256) void k_wide_order(const float4* __restrict__ nodes, float4* __restrict__ wide,
                                                    const uint32_t* __restrict__ wide_trav, uint32_t n_wide,
                                                    uint32_t n_trav) {
    const uint32_t k = blockIdx.x * 256 + threadIdx.x;
    if (k >= n_wide) return;
    const float4* wn = wide + 8ull * k;
    const uint32_t valid = __float_as_uint(wn[0].w) >> 24;
    const float4 a = wn[4], b = wn[5];
    const uint32_t refs[8] = {__float_as_uint(a.x), __float_as_uint(a.y), __float_as_uint(a.z), __float_as_uint(a.w),
                              __float_as_uint(b.x), __float_as_uint(b.y), __float_as_uint(b.z), __float_as_uint(b.w)};
    // each child as the binary layout refers to it: a leaf ref is the same in
    // both layouts, an internal child is the binary node of its root cluster
    uint32_t as_binary[8];
    for (int j = 0; j < 8; ++j) {
        const uint32_t r = refs[j];
        as_binary[j] = !((valid >> j) & 1u) ? kNone : ((r & kLeafBit) ? r : wide_trav[r]);
    }
    uint32_t table[2] = {0u, 0u};
    const uint32_t root = wide_trav[k];
    for (uint32_t e = 0; e < 16 && root < n_trav; ++e) {
        uint32_t x = root, entry = 0;
        for (int step = 0; step < 8; ++step) {
            const float4 n3 = nodes[4ull * x + 3];
            const uint32_t ow = __float_as_uint(n3.z);
            if (!((ow >> (16 + e)) & 1u)) break;   // no confident decision
            const uint32_t child = ((ow >> e) & 1u) ? __float_as_uint(n3.x) : __float_as_uint(n3.y);
            int slot = -1;
            for (int j = 0; j < 8; ++j)
                if (as_binary[j] == child) slot = j;
            if (slot >= 0) {
                entry = 8u | (uint32_t)slot;
                break;
            }
            if ((child & kLeafBit) || child >= n_trav) break;   // not inside this node's subtree
            x = child;
        }
        table[e >> 3] |= entry << (4 * (e & 7));
    }
    wide[8ull * k + 6] = make_float4(__uint_as_float(table[0]), __uint_as_float(table[1]), 0.0f, 0.0f);
}

__global__ void k_wide_order_clear(float4* __restrict__ wide, uint32_t n_wide) {
    const uint32_t k = blockIdx.x * 256 + threadIdx.x;
    if (k < n_wide) wide[8ull * k + 6] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
}

}  // namespace

rtq_status write_wide_order(rtq_ctx* ctx, Scene* s, bool clear, hipStream_t st) {
    s->wide_ordered = false;
    if (!s->wide || s->n_wide == 0) return RTQ_OK;
    const uint32_t blocks = (s->n_wide + 255) / 256;
    if (clear) {
        hipLaunchKernelGGL(k_wide_order_clear, dim3(blocks), dim3(256), 0, st, s->wide, s->n_wide);
    } else {
        hipLaunchKernelGGL(k_wide_order, dim3(blocks), dim3(256), 0, st, s->nodes, s->wide, s->wide_trav, s->n_wide,
                           s->n_trav);
    }
    RTQ_HIP(ctx, hipGetLastError());
    s->wide_ordered = !clear;
    return RTQ_OK;
}

}  // namespace rtq

using namespace rtq;

extern "C" {

rtq_status rtq_order_log(rtq_ctx* ctx, uint32_t scene_id, const rtq_ray* rays_dev, uint64_t n_rays, uint32_t flags,
                         uint32_t stride, float* samples_dev, uint64_t capacity, uint64_t* n_samples) {
    if (!ctx) return RTQ_EINVAL;
    Scene* s = find_scene(ctx, scene_id);
    if (!s) return fail(ctx, RTQ_EINVAL, "order_log: unknown scene %u", scene_id);
    if (!s->built) return fail(ctx, RTQ_ENOTBUILT, "order_log: scene %u has no BVH", scene_id);
    if (!rays_dev || !samples_dev || !n_samples || stride == 0) return fail(ctx, RTQ_EINVAL, "order_log: bad argument");
    if (n_rays > 0xFFFFFFFFull || capacity > 0xFFFFFFFFull) return fail(ctx, RTQ_ECAPACITY, "order_log: too many rays");
    if (flags & ~RTQ_TRACE_ANY) return fail(ctx, RTQ_EINVAL, "order_log: only RTQ_TRACE_ANY may be set");
    RTQ_HIP(ctx, hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    uint32_t* cnt = nullptr;
    RTQ_HIP(ctx, dev_alloc(&cnt, 1));
    hipError_t e = hipMemsetAsync(cnt, 0, sizeof(uint32_t), st);
    const uint64_t threads = (n_rays + stride - 1) / stride;
    if (e == hipSuccess && threads > 0) {
        hipLaunchKernelGGL(k_order_log, dim3((uint32_t)((threads + 255) / 256)), dim3(256), 0, st, s->nodes,
                           s->node_aux, s->tris, rays_dev, (uint32_t)n_rays, stride,
                           (flags & RTQ_TRACE_ANY) ? 1u : 0u, samples_dev, (uint32_t)capacity, cnt);
        e = hipGetLastError();
    }
    uint32_t h = 0;
    if (e == hipSuccess) e = hipMemcpyAsync(&h, cnt, sizeof h, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    (void)hipFree(cnt);
    if (e != hipSuccess) return hip_fail(ctx, e, "order_log");
    *n_samples = std::min<uint64_t>(h, capacity);
    return RTQ_OK;
}

// One batch gradient: grad = mean over samples [0, n) of dLoss/dParams.
rtq_status rtq_order_grad(rtq_ctx* ctx, const float* params_dev, const float* samples_dev, uint32_t n,
                          float* grad_dev, float* loss) {
    if (!ctx) return RTQ_EINVAL;
    if (!params_dev || !samples_dev || !grad_dev || n == 0) return fail(ctx, RTQ_EINVAL, "order_grad: bad argument");
    RTQ_HIP(ctx, hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    const uint32_t blocks = (n + kTB - 1) / kTB;
    float *partial = nullptr, *lp = nullptr;
    RTQ_HIP(ctx, dev_alloc(&partial, (size_t)blocks * kNP));
    hipError_t e = dev_alloc(&lp, blocks);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_mlp_grad, dim3(blocks), dim3(kTB), 0, st, params_dev, samples_dev, n, 0ull, n, 1ull,
                           partial, lp);
        e = hipGetLastError();
    }
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_mlp_reduce, dim3((kNP + 255) / 256), dim3(256), 0, st, partial, blocks, 1.0f / (float)n,
                           grad_dev);
        e = hipGetLastError();
    }
    std::vector<float> hl(blocks);
    if (e == hipSuccess) e = hipMemcpyAsync(hl.data(), lp, sizeof(float) * blocks, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    (void)hipFree(partial);
    (void)hipFree(lp);
    if (e != hipSuccess) return hip_fail(ctx, e, "order_grad");
    if (loss) {
        double a = 0.0;
        for (float v : hl) a += v;
        *loss = (float)(a / n);
    }
    return RTQ_OK;
}

rtq_status rtq_order_train(rtq_ctx* ctx, const float* samples_dev, uint64_t n_samples, const rtq_train_params* tp,
                           float* params_dev, float* adam_m_dev, float* adam_v_dev, uint32_t* step,
                           float* epoch_loss) {
    if (!ctx) return RTQ_EINVAL;
    if (!samples_dev || !tp || !params_dev || !adam_m_dev || !adam_v_dev || !step || n_samples == 0)
        return fail(ctx, RTQ_EINVAL, "order_train: bad argument");
    if (n_samples > 0x7FFFFFFFull) return fail(ctx, RTQ_ECAPACITY, "order_train: too many samples");
    if (tp->batch == 0 || tp->batch % kTB || tp->batch > 65536)
        return fail(ctx, RTQ_EINVAL, "order_train: batch must be a multiple of 64, at most 65536");
    RTQ_HIP(ctx, hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    const uint32_t n = (uint32_t)n_samples;
    const uint32_t blocks = tp->batch / kTB;
    float *partial = nullptr, *lp = nullptr, *grad = nullptr;
    RTQ_HIP(ctx, dev_alloc(&partial, (size_t)blocks * kNP));
    hipError_t e = dev_alloc(&lp, (size_t)blocks * ((n + tp->batch - 1) / tp->batch));
    if (e == hipSuccess) e = dev_alloc(&grad, kNP);
    const uint32_t steps_per_epoch = (n + tp->batch - 1) / tp->batch;
    for (uint32_t ep = 0; ep < tp->epochs && e == hipSuccess; ++ep) {
        // a permutation of the samples per epoch: i -> i * mul mod n with mul coprime to n
        uint64_t mul = 1;
        if (n > 1) {
            uint64_t h = (tp->seed + 0x9E3779B97F4A7C15ull * (ep + 1)) ^ 0xD1B54A32D192ED03ull;
            mul = (h % (n - 1)) + 1;
            auto gcd = [](uint64_t a, uint64_t b) { while (b) { uint64_t t = a % b; a = b; b = t; } return a; };
            while (gcd(mul, n) != 1) mul = mul % (n - 1) + 1;
        }
        for (uint32_t sidx = 0; sidx < steps_per_epoch && e == hipSuccess; ++sidx) {
            const uint64_t start = (uint64_t)sidx * tp->batch;
            const uint32_t cnt = (uint32_t)std::min<uint64_t>(tp->batch, n - start);
            const uint32_t bl = (cnt + kTB - 1) / kTB;
            hipLaunchKernelGGL(k_mlp_grad, dim3(bl), dim3(kTB), 0, st, params_dev, samples_dev, n, start, cnt, mul,
                               partial, lp + (size_t)sidx * blocks);
            e = hipGetLastError();
            if (e != hipSuccess) break;
            hipLaunchKernelGGL(k_mlp_reduce, dim3((kNP + 255) / 256), dim3(256), 0, st, partial, bl,
                               1.0f / (float)cnt, grad);
            e = hipGetLastError();
            if (e != hipSuccess) break;
            const uint32_t tstep = ++*step;
            const float c1 = 1.0f - powf(tp->beta1, (float)tstep), c2 = 1.0f - powf(tp->beta2, (float)tstep);
            hipLaunchKernelGGL(k_mlp_adam, dim3((kNP + 255) / 256), dim3(256), 0, st, params_dev, adam_m_dev,
                               adam_v_dev, grad, tp->lr, tp->beta1, tp->beta2, 1e-8f, c1, c2);
            e = hipGetLastError();
        }
        if (e != hipSuccess) break;
        // loss of the epoch (sum of the per-block sums of all steps)
        std::vector<float> hl((size_t)steps_per_epoch * blocks, 0.0f);
        e = hipMemcpyAsync(hl.data(), lp, sizeof(float) * hl.size(), hipMemcpyDeviceToHost, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        if (e == hipSuccess && epoch_loss) {
            double a = 0.0;
            for (uint32_t si = 0; si < steps_per_epoch; ++si) {
                const uint64_t start = (uint64_t)si * tp->batch;
                const uint32_t cnt = (uint32_t)std::min<uint64_t>(tp->batch, n - start);
                const uint32_t bl = (cnt + kTB - 1) / kTB;
                for (uint32_t b = 0; b < bl; ++b) a += hl[(size_t)si * blocks + b];
            }
            epoch_loss[ep] = (float)(a / n);
        }
    }
    (void)hipFree(partial);
    (void)hipFree(lp);
    (void)hipFree(grad);
    if (e != hipSuccess) return hip_fail(ctx, e, "order_train");
    return RTQ_OK;
}

rtq_status rtq_order_predict(rtq_ctx* ctx, const float* params_dev, const float* samples_dev, uint64_t n,
                             uint32_t stride, float* prob_dev) {
    if (!ctx) return RTQ_EINVAL;
    if (!params_dev || !samples_dev || !prob_dev || stride < kIn) return fail(ctx, RTQ_EINVAL, "order_predict: bad argument");
    if (n == 0) return RTQ_OK;
    if (n > 0xFFFFFFFFull) return fail(ctx, RTQ_ECAPACITY, "order_predict: too many samples");
    RTQ_HIP(ctx, hipSetDevice(ctx->device));
    const uint32_t tiles = (uint32_t)((n + 31) / 32);
    const uint32_t blocks = std::min<uint32_t>((tiles + kIW - 1) / kIW, 2048u);
    hipLaunchKernelGGL(k_mlp_infer, dim3(blocks), dim3(64 * kIW), 0, ctx->stream, params_dev, samples_dev, (uint32_t)n,
                       stride, prob_dev);
    RTQ_HIP(ctx, hipGetLastError());
    RTQ_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return RTQ_OK;
}

rtq_status rtq_order_apply(rtq_ctx* ctx, uint32_t scene_id, const float* params_dev, float confidence_closest,
                           float confidence_any) {
    if (!ctx) return RTQ_EINVAL;
    Scene* s = find_scene(ctx, scene_id);
    if (!s) return fail(ctx, RTQ_EINVAL, "order_apply: unknown scene %u", scene_id);
    if (!s->built) return fail(ctx, RTQ_ENOTBUILT, "order_apply: scene %u has no BVH", scene_id);
    if (params_dev && !(confidence_closest >= 0.5f && confidence_any >= 0.5f))
        return fail(ctx, RTQ_EINVAL, "order_apply: confidences must be >= 0.5 (>= 1 disables a mode)");
    RTQ_HIP(ctx, hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    if (!params_dev) {
        hipLaunchKernelGGL(k_order_clear, dim3((s->n_trav + 255) / 256), dim3(256), 0, st, s->nodes, s->n_trav);
    } else {
        // max(p, 1 - p) >= c  <=>  |logit| >= log(c / (1 - c)); c >= 1: never
        auto zmin = [](float c) {
            return c >= 1.0f ? INFINITY : (float)std::log((double)c / (1.0 - (double)c));
        };
        const uint32_t tiles = (s->n_trav + 1) / 2;
        const uint32_t blocks = std::min<uint32_t>((tiles + kIW - 1) / kIW, 2048u);
        hipLaunchKernelGGL(k_order_apply, dim3(blocks), dim3(64 * kIW), 0, st, params_dev, s->nodes, s->node_aux,
                           s->n_trav, zmin(confidence_closest), zmin(confidence_any));
    }
    RTQ_HIP(ctx, hipGetLastError());
    // the LDS copy of the top of the tree carries the order words too
    rtq_status rs = write_top(ctx, s, st);
    if (rs != RTQ_OK) return rs;
    // the 8-wide nodes' first-child tables follow the binary decisions
    rs = write_wide_order(ctx, s, params_dev == nullptr, st);
    if (rs != RTQ_OK) return rs;
    RTQ_HIP(ctx, hipStreamSynchronize(st));
    s->ordered = params_dev != nullptr;
    s->order_closest = s->ordered && confidence_closest < 1.0f;
    s->order_any = s->ordered && confidence_any < 1.0f;
    return RTQ_OK;
}

rtq_status rtq_order_init(rtq_ctx* ctx, uint64_t seed, float* params_dev) {
    if (!ctx) return RTQ_EINVAL;
    if (!params_dev) return fail(ctx, RTQ_EINVAL, "order_init: null parameters");
    // He-uniform weights (limit sqrt(6 / fan_in)) from a SplitMix64 stream, zero biases
    std::vector<float> p(kNP, 0.0f);
    uint64_t st = seed;
    auto next = [&]() {
        uint64_t z = (st += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        return (double)(z >> 40) * (1.0 / 16777216.0) * 2.0 - 1.0;
    };
    for (int i = 0; i < kIn * kH; ++i) p[kW1 + i] = (float)(next() * std::sqrt(6.0 / kIn));
    for (int i = 0; i < kH * kH; ++i) p[kW2 + i] = (float)(next() * std::sqrt(6.0 / kH));
    for (int i = 0; i < kH; ++i) p[kW3 + i] = (float)(next() * std::sqrt(6.0 / kH));
    RTQ_HIP(ctx, hipSetDevice(ctx->device));
    RTQ_HIP(ctx, hipMemcpy(params_dev, p.data(), sizeof(float) * kNP, hipMemcpyHostToDevice));
    return RTQ_OK;
}

namespace {
constexpr uint32_t kCkptMagic = 0x4d4f5452u;  // "RTOM"
constexpr uint32_t kCkptVersion = 1;

uint64_t fnv1a(const void* data, size_t n, uint64_t h = 0xcbf29ce484222325ull) {
    const unsigned char* p = static_cast<const unsigned char*>(data);
    for (size_t i = 0; i < n; ++i) h = (h ^ p[i]) * 0x100000001b3ull;
    return h;
}
}  // namespace

// File: magic, version, parameter count, step (4 x u32), then params, m, v
// (3 x RTQ_ORDER_PARAMS f32), then FNV-1a 64 of everything before it.
rtq_status rtq_order_save(rtq_ctx* ctx, const char* path, const float* params_dev, const float* adam_m_dev,
                          const float* adam_v_dev, uint32_t step) {
    if (!ctx) return RTQ_EINVAL;
    if (!path || !params_dev || !adam_m_dev || !adam_v_dev) return fail(ctx, RTQ_EINVAL, "order_save: bad argument");
    std::vector<unsigned char> buf(16 + 3 * sizeof(float) * kNP);
    const uint32_t head[4] = {kCkptMagic, kCkptVersion, (uint32_t)kNP, step};
    std::memcpy(buf.data(), head, 16);
    RTQ_HIP(ctx, hipSetDevice(ctx->device));
    const float* src[3] = {params_dev, adam_m_dev, adam_v_dev};
    for (int k = 0; k < 3; ++k)
        RTQ_HIP(ctx, hipMemcpy(buf.data() + 16 + k * sizeof(float) * kNP, src[k], sizeof(float) * kNP,
                               hipMemcpyDeviceToHost));
    const uint64_t h = fnv1a(buf.data(), buf.size());
    FILE* f = std::fopen(path, "wb");
    if (!f) return fail(ctx, RTQ_EINVAL, "order_save: cannot open %s", path);
    const bool ok = std::fwrite(buf.data(), 1, buf.size(), f) == buf.size() && std::fwrite(&h, 8, 1, f) == 1;
    const bool closed = std::fclose(f) == 0;
    if (!ok || !closed) return fail(ctx, RTQ_EINVAL, "order_save: write to %s failed", path);
    return RTQ_OK;
}

rtq_status rtq_order_load(rtq_ctx* ctx, const char* path, float* params_dev, float* adam_m_dev, float* adam_v_dev,
                          uint32_t* step) {
    if (!ctx) return RTQ_EINVAL;
    if (!path || !params_dev) return fail(ctx, RTQ_EINVAL, "order_load: bad argument");
    FILE* f = std::fopen(path, "rb");
    if (!f) return fail(ctx, RTQ_EINVAL, "order_load: cannot open %s", path);
    std::vector<unsigned char> buf(16 + 3 * sizeof(float) * kNP);
    uint64_t h = 0;
    const bool ok = std::fread(buf.data(), 1, buf.size(), f) == buf.size() && std::fread(&h, 8, 1, f) == 1;
    const bool extra = ok && std::fgetc(f) != EOF;
    std::fclose(f);
    if (!ok || extra) return fail(ctx, RTQ_EINVAL, "order_load: %s has the wrong size", path);
    uint32_t head[4];
    std::memcpy(head, buf.data(), 16);
    if (head[0] != kCkptMagic) return fail(ctx, RTQ_EINVAL, "order_load: %s is not an order-model file", path);
    if (head[1] != kCkptVersion) return fail(ctx, RTQ_EINVAL, "order_load: %s: unsupported version %u", path, head[1]);
    if (head[2] != (uint32_t)kNP) return fail(ctx, RTQ_EINVAL, "order_load: %s: %u parameters", path, head[2]);
    if (fnv1a(buf.data(), buf.size()) != h) return fail(ctx, RTQ_EINVAL, "order_load: %s: checksum mismatch", path);
    RTQ_HIP(ctx, hipSetDevice(ctx->device));
    float* dst[3] = {params_dev, adam_m_dev, adam_v_dev};
    for (int k = 0; k < 3; ++k)
        if (dst[k])
            RTQ_HIP(ctx, hipMemcpy(dst[k], buf.data() + 16 + k * sizeof(float) * kNP, sizeof(float) * kNP,
                                   hipMemcpyHostToDevice));
    if (step) *step = head[3];
    return RTQ_OK;
}

}  // extern "C"

// ---- the order model as one object owning its device memory ----------------

struct rtq_order_model {
    rtq_ctx* ctx = nullptr;
    float* params = nullptr;   // kNP
    float* m = nullptr;        // kNP (Adam first moment)
    float* v = nullptr;        // kNP (Adam second moment)
    float* samples = nullptr;  // capacity x kSample
    uint64_t capacity = 0;
    uint64_t n_samples = 0;
    uint32_t step = 0;
};

namespace {
void free_model(rtq_order_model* m) {
    if (m->params) (void)hipFree(m->params);
    if (m->m) (void)hipFree(m->m);
    if (m->v) (void)hipFree(m->v);
    if (m->samples) (void)hipFree(m->samples);
    delete m;
}
}  // namespace

extern "C" {

rtq_status rtq_order_model_create(rtq_ctx* ctx, uint64_t seed, uint64_t sample_capacity, rtq_order_model** out) {
    if (!ctx) return RTQ_EINVAL;
    if (!out) return fail(ctx, RTQ_EINVAL, "order_model_create: null output");
    *out = nullptr;
    if (sample_capacity > 0xFFFFFFFFull)
        return fail(ctx, RTQ_ECAPACITY, "order_model_create: at most 2^32 - 1 samples");
    RTQ_HIP(ctx, hipSetDevice(ctx->device));
    rtq_order_model* m = new rtq_order_model;
    m->ctx = ctx;
    m->capacity = sample_capacity;
    hipError_t e = dev_alloc(&m->params, kNP);
    if (e == hipSuccess) e = dev_alloc(&m->m, kNP);
    if (e == hipSuccess) e = dev_alloc(&m->v, kNP);
    if (e == hipSuccess) e = dev_alloc(&m->samples, (size_t)sample_capacity * kSample);
    if (e == hipSuccess) e = hipMemset(m->m, 0, sizeof(float) * kNP);
    if (e == hipSuccess) e = hipMemset(m->v, 0, sizeof(float) * kNP);
    if (e != hipSuccess) {
        free_model(m);
        return hip_fail(ctx, e, "order_model_create");
    }
    const rtq_status s = rtq_order_init(ctx, seed, m->params);
    if (s != RTQ_OK) {
        free_model(m);
        return s;
    }
    *out = m;
    return RTQ_OK;
}

void rtq_order_model_destroy(rtq_order_model* model) {
    if (!model) return;
    (void)hipSetDevice(model->ctx->device);
    free_model(model);
}

rtq_status rtq_order_model_log(rtq_order_model* model, uint32_t scene_id, const rtq_ray* rays_dev, uint64_t n_rays,
                               uint32_t flags, uint32_t stride, uint64_t* n_added) {
    if (!model) return RTQ_EINVAL;
    uint64_t k = 0;
    const uint64_t room = model->capacity - model->n_samples;
    if (room > 0) {
        const rtq_status s = rtq_order_log(model->ctx, scene_id, rays_dev, n_rays, flags, stride,
                                           model->samples + model->n_samples * kSample, room, &k);
        if (s != RTQ_OK) return s;
    }
    model->n_samples += k;
    if (n_added) *n_added = k;
    return RTQ_OK;
}

rtq_status rtq_order_model_clear_samples(rtq_order_model* model) {
    if (!model) return RTQ_EINVAL;
    model->n_samples = 0;
    return RTQ_OK;
}

rtq_status rtq_order_model_train(rtq_order_model* model, const rtq_train_params* tp, float* epoch_loss) {
    if (!model) return RTQ_EINVAL;
    if (model->n_samples == 0) return fail(model->ctx, RTQ_EINVAL, "order_model_train: no samples logged");
    return rtq_order_train(model->ctx, model->samples, model->n_samples, tp, model->params, model->m, model->v,
                           &model->step, epoch_loss);
}

rtq_status rtq_order_model_apply(rtq_order_model* model, uint32_t scene_id, float confidence_closest,
                                 float confidence_any) {
    if (!model) return RTQ_EINVAL;
    return rtq_order_apply(model->ctx, scene_id, model->params, confidence_closest, confidence_any);
}

rtq_status rtq_order_model_save(rtq_order_model* model, const char* path) {
    if (!model) return RTQ_EINVAL;
    return rtq_order_save(model->ctx, path, model->params, model->m, model->v, model->step);
}

rtq_status rtq_order_model_load(rtq_order_model* model, const char* path) {
    if (!model) return RTQ_EINVAL;
    return rtq_order_load(model->ctx, path, model->params, model->m, model->v, &model->step);
}

rtq_status rtq_order_model_info(const rtq_order_model* model, uint64_t* n_samples, uint32_t* step) {
    if (!model) return RTQ_EINVAL;
    if (n_samples) *n_samples = model->n_samples;
    if (step) *step = model->step;
    return RTQ_OK;
}

}  // extern "C"
