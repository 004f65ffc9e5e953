"""FP64 NumPy restatement of the tile-configuration cost model (the
multilayer perceptron of torque_amd/csrc/tq_costmodel.hip) and of its
features. TEST INFRASTRUCTURE ONLY: imported by tests/ and the benchmark,
never by product code.

No reference counterpart: the reference has no tile choice (it calls
cublas{S,D}gemmStridedBatched per A block, src/gpu/block_sparse.cuh:870-930).

Pinning: `grad` is checked against central finite differences of `loss`
(tests/test_costmodel.py), which the mathematics fixes independently of
the backward formulas; the GPU gradients are then checked against `grad`.
"""
import numpy as np

F = 34   # TQ_COST_FEATURES
H = 64   # TQ_COST_HIDDEN

OFF_W1 = 0
OFF_B1 = OFF_W1 + F * H
OFF_W2 = OFF_B1 + H
OFF_B2 = OFF_W2 + H * H
OFF_W3 = OFF_B2 + H
OFF_B3 = OFF_W3 + H
N_PARAMS = OFF_B3 + 1


def unpack(p):
    p = np.asarray(p, np.float64)
    return (p[OFF_W1:OFF_B1].reshape(F, H), p[OFF_B1:OFF_W2], p[OFF_W2:OFF_B2].reshape(H, H),
            p[OFF_B2:OFF_W3], p[OFF_W3:OFF_B3], p[OFF_B3])


def pack(w1, b1, w2, b2, w3, b3):
    return np.concatenate([w1.ravel(), b1, w2.ravel(), b2, w3, [b3]])


def forward(p, x):
    """Network output for inputs x [n, F]; also the intermediates."""
    w1, b1, w2, b2, w3, b3 = unpack(p)
    x = np.asarray(x, np.float64)
    z1 = x @ w1 + b1
    h1 = np.maximum(z1, 0.0)
    z2 = h1 @ w2 + b2
    h2 = np.maximum(z2, 0.0)
    out = h2 @ w3 + b3
    return out, (x, z1, h1, z2, h2)


def loss(p, x, y):
    """Mean squared error."""
    out, _ = forward(p, x)
    return float(np.mean((out - np.asarray(y, np.float64)) ** 2))


def grad(p, x, y):
    """Gradient of `loss` by backpropagation (ReLU derivative 0 at 0, as in
    the kernels)."""
    w1, b1, w2, b2, w3, b3 = unpack(p)
    out, (x, z1, h1, z2, h2) = forward(p, x)
    n = x.shape[0]
    d = 2.0 * (out - np.asarray(y, np.float64)) / n          # dL/dout [n]
    g_w3 = h2.T @ d
    g_b3 = d.sum()
    dz2 = np.outer(d, w3) * (z2 > 0)                         # [n, H]
    g_w2 = h1.T @ dz2
    g_b2 = dz2.sum(0)
    dz1 = (dz2 @ w2.T) * (z1 > 0)
    g_w1 = x.T @ dz1
    g_b1 = dz1.sum(0)
    return pack(g_w1, g_b1, g_w2, g_b2, g_w3, g_b3)


def finite_difference(p, x, y, idx, h=1e-6):
    """Central difference of `loss` along parameter idx."""
    p = np.asarray(p, np.float64)
    e = np.zeros_like(p)
    e[idx] = h
    return (loss(p + e, x, y) - loss(p - e, x, y)) / (2 * h)


def _bucket(d):
    b, t = 0, 16
    while b < 7 and d >= t:
        b += 1
        t <<= 1
    return b


def features(m, n, k, bm, bn, bk, wave_tile, n_cu=256, min_waves=1):
    """The feature vector of tq_cost_features for pairs (m, n, k) under a
    tile of bm x bn (k step bk)."""
    f = np.zeros(F)
    m, n, k = (np.asarray(v, np.float64) for v in (m, n, k))
    if m.size == 0:
        return f
    fl = 2 * m * n * k
    tot = fl.sum()
    for j, v in enumerate((m, n, k)):
        for q in range(v.size):
            f[8 * j + _bucket(int(v[q]))] += fl[q]
    f[:24] /= tot
    tiles = (np.ceil(m / bm) * np.ceil(n / bn)).sum()
    padded = (2 * np.ceil(m / bm) * bm * np.ceil(n / bn) * bn * np.ceil(k / bk) * bk).sum()
    f[24] = np.log10(tot)
    f[25] = np.log10(m.size)
    f[26] = np.log10(tiles)
    f[27] = tot / padded
    f[28], f[29], f[30] = np.log2(bm), np.log2(bn), np.log2(bk)
    f[31] = 2 * wave_tile + min_waves
    f[32] = np.log10(tiles / (n_cu * (16 if wave_tile else 4)))
    f[33] = np.log10((8 * m * n).sum())
    return f
