"""The oracle's learned-ordering functions, pinned without the GPU:
the FP64 gradient by central finite differences and an all-zero network's
closed form, the forward pass by hand-built networks, the BF16 rounding by
hand values, and the node features by a hand-computed node."""
import math

import numpy as np
import pytest

import oracle

NP = 6337
W1, B1, W2, B2, W3, B3 = 0, 2048, 2112, 6208, 6272, 6336


def rand_params(seed, scale=1.0):
    rng = np.random.default_rng(seed)
    p = np.zeros(NP, np.float32)
    p[W1:B1] = rng.uniform(-1, 1, 2048) * math.sqrt(6 / 32) * scale
    p[B1:W2] = rng.uniform(-0.1, 0.1, 64)
    p[W2:B2] = rng.uniform(-1, 1, 4096) * math.sqrt(6 / 64) * scale
    p[B2:W3] = rng.uniform(-0.1, 0.1, 64)
    p[W3:B3] = rng.uniform(-1, 1, 64) * math.sqrt(6 / 64) * scale
    p[B3] = 0.05
    return p


def rand_samples(n, seed):
    rng = np.random.default_rng(seed)
    s = np.zeros((n, 34), np.float32)
    s[:, :32] = rng.uniform(-1, 1, (n, 32))
    s[:, 32] = rng.integers(0, 2, n)
    s[:, 33] = rng.uniform(0.5, 2.0, n)
    return s


def test_param_count():
    assert oracle.mlp_param_count() == NP == 32 * 64 + 64 + 64 * 64 + 64 + 64 + 1


def test_grad_matches_finite_differences():
    P = rand_params(1)
    S = rand_samples(24, 2)
    loss, g = oracle.mlp_grad(P, S)
    assert abs(loss - oracle.mlp_loss(P, S)) < 1e-12
    rng = np.random.default_rng(3)
    # every block of parameters, a few entries each (FP32 parameters: step 2^-10
    # on values of magnitude <= 1 is exactly representable around them)
    idx = np.concatenate([rng.choice(np.arange(a, b), size=min(6, b - a), replace=False)
                          for a, b in ((W1, B1), (B1, W2), (W2, B2), (B2, W3), (W3, B3), (B3, NP))])
    h = 2.0 ** -10
    for j in idx:
        Pp, Pm = P.copy(), P.copy()
        Pp[j] += h
        Pm[j] -= h
        fd = (oracle.mlp_loss(Pp, S) - oracle.mlp_loss(Pm, S)) / (float(Pp[j]) - float(Pm[j]))
        assert abs(fd - g[j]) <= 1e-6 + 1e-3 * abs(g[j]), (j, fd, g[j])


def test_grad_zero_network_closed_form():
    # all weights zero: z = b3 for every sample, so only b3 has a gradient,
    # mean(w * (sigmoid(b3) - y)), and the loss is mean(w * BCE(b3, y))
    P = np.zeros(NP, np.float32)
    P[B3] = 0.3
    S = rand_samples(50, 4)
    loss, g = oracle.mlp_grad(P, S)
    sig = 1 / (1 + math.exp(-float(P[B3])))
    y, w = S[:, 32].astype(np.float64), S[:, 33].astype(np.float64)
    assert g[B3] == pytest.approx(np.mean(w * (sig - y)), rel=1e-12)
    assert np.count_nonzero(g[:B3]) == 0
    bce = np.where(y == 1, -math.log(sig), -math.log(1 - sig))
    assert loss == pytest.approx(np.mean(w * bce), rel=1e-12)


def test_logits_hand_network():
    # h1_j = relu(x_j) for j < 32 (identity W1 block), h2_0 = relu(sum_j h1_j),
    # z = 2 * h2_0 - 1  ->  z = 2 * sum(relu(x)) - 1
    P = np.zeros(NP, np.float32)
    for j in range(32):
        P[W1 + j * 64 + j] = 1.0
        P[W2 + j * 64 + 0] = 1.0
    P[W3 + 0] = 2.0
    P[B3] = -1.0
    x = np.random.default_rng(5).uniform(-1, 1, (10, 32)).astype(np.float32)
    want = 2 * np.maximum(x.astype(np.float64), 0).sum(axis=1) - 1
    assert np.allclose(oracle.mlp_logits(P, x), want, rtol=0, atol=1e-12)
    # the same network in BF16 mode: x rounded to BF16, each h1 rounded (already
    # BF16 since h1 = relu(bf16(x))), weights exact in BF16
    xb = np.array([[oracle.bf16_round(float(v)) for v in row] for row in x], np.float64)
    wantb = 2 * np.maximum(xb, 0).sum(axis=1) - 1
    assert np.allclose(oracle.mlp_logits(P, x, bf16=True), wantb, rtol=0, atol=1e-12)


def test_bf16_round_values():
    assert oracle.bf16_round(1.0) == 1.0
    # 1 + 2^-8 is halfway between 1 and 1 + 2^-7: ties to even (1.0)
    assert oracle.bf16_round(1.0 + 2.0 ** -8) == 1.0
    # 1 + 3 * 2^-8 is halfway between 1 + 2^-7 and 1 + 2^-6: ties to even (1 + 2^-6)
    assert oracle.bf16_round(1.0 + 3 * 2.0 ** -8) == 1.0 + 2.0 ** -6
    assert oracle.bf16_round(1.0 + 2.0 ** -8 + 2.0 ** -20) == 1.0 + 2.0 ** -7
    assert oracle.bf16_round(-3.0) == -3.0
    assert oracle.bf16_round(1.0 / 3.0) == pytest.approx(0.333984375, abs=0)


def test_features_hand_node():
    # child 0 = [0,1]x[0,1]x[0,1] (a leaf of 3 triangles), child 1 = [1,3]x[0,2]x[0,1] (internal,
    # 5 triangles); node box [0,3]x[0,2]x[0,1], largest extent 3
    nd = np.zeros(16, np.float32)
    nd[0:4] = [0, 1, 0, 1]          # c0.lo.x c0.hi.x c0.lo.y c0.hi.y
    nd[4:8] = [0, 1, 1, 3]          # c0.lo.z c0.hi.z c1.lo.x c1.hi.x
    nd[8:12] = [0, 2, 0, 1]         # c1.lo.y c1.hi.y c1.lo.z c1.hi.z
    refs = np.array([0x80000000 | (10 << 3) | 2, 7, 0, 0], np.uint32)
    nd[12:16] = refs.view(np.float32)
    aux = np.array([10, 13, 18, 4], np.uint32)
    f = oracle.order_features(nd, aux, octant=0b101, any_hit=1)
    assert list(f[0:4]) == [-1, 1, -1, 1]
    assert f[4] == pytest.approx(4 / 32)
    assert np.allclose(f[5:8], [0, 0, 0]) and np.allclose(f[8:11], [1 / 3, 1 / 3, 1 / 3])
    assert np.allclose(f[11:14], [1 / 3, 0, 0]) and np.allclose(f[14:17], [1, 2 / 3, 1 / 3])
    # centres (0.5,0.5,0.5) and (2,1,0.5): offset (1.5,0.5,0)/3, signs (-1,1,-1)
    assert np.allclose(f[17:20], [0.5, 1 / 6, 0])
    assert f[20] == pytest.approx(-0.5 + 1 / 6)
    node_area = 2 * (3 * 2 + 2 * 1 + 1 * 3)
    assert f[21] == pytest.approx(6 / node_area) and f[22] == pytest.approx(2 * (4 + 2 + 2) / node_area)
    assert f[23] == pytest.approx(2 / 24) and f[24] == pytest.approx(math.log2(6) / 24)
    assert (f[25], f[26]) == (1, 0)
    assert np.allclose(f[27:30], [1, 2 / 3, 1 / 3])
    assert (f[30], f[31]) == (0, 0)
