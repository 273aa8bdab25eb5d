/*
 * ORACLE — test infrastructure only.
 *
 * CPU restatement of the learned child-ordering model: node features, the
 * 32 -> 64 -> 64 -> 1 MLP forward pass (FP64, optionally with the BF16
 * rounding points of the GPU inference) and the FP64 gradient of the weighted
 * binary cross-entropy.  The reference has no learned ordering (its shader
 * always enters the left child first, raytracer.glsl:264-290), so nothing here
 * follows a reference line; each function is pinned by the tests:
 *   orc_mlp_grad      central finite differences of orc_mlp_loss, and the
 *                     closed form of an all-zero network (only b3 moves)
 *   orc_mlp_logits    hand-built networks with known outputs; the BF16 mode
 *                     against round-to-nearest-even BF16 on hand values
 *   orc_order_features  hand-computed features of a hand-made node
 */
#include <math.h>
#include <stdint.h>
#include <string.h>

#include "oracle.h"

#define IN 32
#define H 64
#define OFF_W1 0
#define OFF_B1 (OFF_W1 + IN * H)
#define OFF_W2 (OFF_B1 + H)
#define OFF_B2 (OFF_W2 + H * H)
#define OFF_W3 (OFF_B2 + H)
#define OFF_B3 (OFF_W3 + H)
#define NP (OFF_B3 + 1)
#define SAMPLE (IN + 2)

uint32_t orc_mlp_param_count(void) { return NP; }

/* float -> bf16 -> float, round to nearest even (finite inputs). */
float orc_bf16_round(float f) {
    uint32_t u;
    memcpy(&u, &f, 4);
    u += 0x7FFFu + ((u >> 16) & 1u);
    u &= 0xFFFF0000u;
    float r;
    memcpy(&r, &u, 4);
    return r;
}

static double relu(double x) { return x > 0.0 ? x : 0.0; }

/* Forward pass of one row; h1/h2 receive the activations (may be NULL). */
static double forward(const float* P, const float* x, int bf16, double* h1, double* h2) {
    double a1[H], a2[H];
    for (int j = 0; j < H; ++j) {
        double acc = P[OFF_B1 + j];
        for (int k = 0; k < IN; ++k) {
            const double xv = bf16 ? orc_bf16_round(x[k]) : x[k];
            const double wv = bf16 ? orc_bf16_round(P[OFF_W1 + k * H + j]) : P[OFF_W1 + k * H + j];
            acc += xv * wv;
        }
        a1[j] = relu(acc);
        if (bf16) a1[j] = orc_bf16_round((float)a1[j]);
    }
    double z = P[OFF_B3];
    for (int j = 0; j < H; ++j) {
        double acc = P[OFF_B2 + j];
        for (int k = 0; k < H; ++k) {
            const double wv = bf16 ? orc_bf16_round(P[OFF_W2 + k * H + j]) : P[OFF_W2 + k * H + j];
            acc += a1[k] * wv;
        }
        a2[j] = relu(acc);
        z += a2[j] * (double)P[OFF_W3 + j];
    }
    if (h1) memcpy(h1, a1, sizeof a1);
    if (h2) memcpy(h2, a2, sizeof a2);
    return z;
}

void orc_mlp_logits(const float* P, const float* X, uint64_t n, uint32_t stride, int bf16, double* out) {
    for (uint64_t i = 0; i < n; ++i) out[i] = forward(P, X + i * stride, bf16, NULL, NULL);
}

static double bce(double z, double y) { return (z > 0.0 ? z : 0.0) - z * y + log1p(exp(-fabs(z))); }

double orc_mlp_loss(const float* P, const float* S, uint64_t n) {
    double acc = 0.0;
    for (uint64_t i = 0; i < n; ++i) {
        const float* s = S + i * SAMPLE;
        acc += (double)s[IN + 1] * bce(forward(P, s, 0, NULL, NULL), s[IN]);
    }
    return n ? acc / (double)n : 0.0;
}

/* Mean weighted BCE and its gradient (NP doubles), backpropagation in FP64.
 * ReLU'(x) = 1 for x > 0, else 0. */
double orc_mlp_grad(const float* P, const float* S, uint64_t n, double* grad) {
    memset(grad, 0, sizeof(double) * NP);
    double loss = 0.0;
    for (uint64_t i = 0; i < n; ++i) {
        const float* s = S + i * SAMPLE;
        double h1[H], h2[H];
        const double z = forward(P, s, 0, h1, h2);
        const double y = s[IN], w = s[IN + 1];
        loss += w * bce(z, y);
        const double dz = w * (1.0 / (1.0 + exp(-z)) - y);
        grad[OFF_B3] += dz;
        double d2[H];
        for (int j = 0; j < H; ++j) {
            grad[OFF_W3 + j] += h2[j] * dz;
            d2[j] = h2[j] > 0.0 ? dz * (double)P[OFF_W3 + j] : 0.0;
            grad[OFF_B2 + j] += d2[j];
        }
        double d1[H];
        for (int k = 0; k < H; ++k) {
            double acc = 0.0;
            for (int j = 0; j < H; ++j) {
                grad[OFF_W2 + k * H + j] += h1[k] * d2[j];
                acc += d2[j] * (double)P[OFF_W2 + k * H + j];
            }
            d1[k] = h1[k] > 0.0 ? acc : 0.0;
            grad[OFF_B1 + k] += d1[k];
        }
        for (int k = 0; k < IN; ++k)
            for (int j = 0; j < H; ++j) grad[OFF_W1 + k * H + j] += (double)s[k] * d1[j];
    }
    if (n) {
        for (int j = 0; j < NP; ++j) grad[j] /= (double)n;
        loss /= (double)n;
    }
    return loss;
}

static float area3(float ex, float ey, float ez) { return 2.0f * (ex * ey + ey * ez + ez * ex); }

/* Features of a 64-B traversal node (16 floats: child boxes interleaved as
 * [c0.lo.x c0.hi.x c0.lo.y c0.hi.y][c0.lo.z c0.hi.z c1.lo.x c1.hi.x]
 * [c1.lo.y c1.hi.y c1.lo.z c1.hi.z][ref0 ref1 order 0]) with its child
 * leaf ranges and depth (aux: r0, r1, r2, depth), for ray octant `octant`
 * (bit k set <=> direction component k < 0) and mode `any`:
 *   0-2   direction signs (+1 / -1)        3  mode (1 = any hit)
 *   4     depth / 32
 *   5-16  child 0 lo, child 0 hi, child 1 lo, child 1 hi, relative to the
 *         node box's low corner, over the node box's largest extent
 *   17-19 child-1 centre minus child-0 centre (same scale)
 *   20    that offset dotted with the direction signs
 *   21-22 child surface areas over the node box's
 *   23-24 log2(leaf count + 1) / 24 of each child
 *   25-26 child is a leaf (1 / 0)             27-29 node box extents (same scale)
 *   30-31 zero */
void orc_order_features(const float* nd, const uint32_t* aux, uint32_t octant, uint32_t any, float* f) {
    const float l0[3] = {nd[0], nd[2], nd[4]}, h0[3] = {nd[1], nd[3], nd[5]};
    const float l1[3] = {nd[6], nd[8], nd[10]}, h1[3] = {nd[7], nd[9], nd[11]};
    uint32_t r0, r1;
    memcpy(&r0, &nd[12], 4);
    memcpy(&r1, &nd[13], 4);
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
    f[4] = (float)aux[3] * (1.0f / 32.0f);
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
    const float sp = fmaxf(area3(ext[0], ext[1], ext[2]), 1e-30f);
    f[21] = area3(h0[0] - l0[0], h0[1] - l0[1], h0[2] - l0[2]) / sp;
    f[22] = area3(h1[0] - l1[0], h1[1] - l1[1], h1[2] - l1[2]) / sp;
    f[23] = log2f((float)(aux[1] - aux[0]) + 1.0f) * (1.0f / 24.0f);
    f[24] = log2f((float)(aux[2] - aux[1]) + 1.0f) * (1.0f / 24.0f);
    f[25] = (r0 & 0x80000000u) ? 1.0f : 0.0f;
    f[26] = (r1 & 0x80000000u) ? 1.0f : 0.0f;
    f[30] = 0.0f;
    f[31] = 0.0f;
}
