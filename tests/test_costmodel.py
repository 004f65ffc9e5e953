"""Cost-model oracle pinned by the mathematics, and the library's host-side
features against the oracle's (no GPU needed).

The oracle's backpropagation (oracle/mlp.py) must agree with central finite
differences of its own loss: a dropped term, wrong sign or transposed weight
in the backward formulas fails this. The C++ feature extraction must agree
with the oracle's restatement for every configuration."""
import numpy as np
import pytest

from oracle import mlp as M
from torque_amd import gemm_configs
from torque_amd.costmodel import N_FEATURES, shape_features


def random_params(rng, scale=0.3):
    return rng.normal(0, scale, M.N_PARAMS)


def test_layout_matches_library():
    from torque_amd.costmodel import CostModel
    assert M.F == N_FEATURES
    assert CostModel.n_params() == M.N_PARAMS


@pytest.mark.parametrize("seed", [0, 1])
def test_backprop_matches_finite_differences(seed):
    rng = np.random.default_rng(seed)
    p = random_params(rng)
    x = rng.normal(size=(23, M.F))
    y = rng.normal(size=23)
    g = M.grad(p, x, y)
    # every parameter group: W1, b1, W2, b2, w3, b3
    idx = np.concatenate([rng.integers(M.OFF_W1, M.OFF_B1, 40), rng.integers(M.OFF_B1, M.OFF_W2, 8),
                          rng.integers(M.OFF_W2, M.OFF_B2, 40), rng.integers(M.OFF_B2, M.OFF_W3, 8),
                          rng.integers(M.OFF_W3, M.OFF_B3, 8), [M.OFF_B3]])
    fd = np.array([M.finite_difference(p, x, y, i) for i in idx])
    assert np.allclose(g[idx], fd, rtol=1e-5, atol=1e-8), np.abs(g[idx] - fd).max()


def test_backprop_closed_form_linear_case():
    # with all pre-activations positive the network is linear: out = x A + c,
    # and dL/db3 = mean(2 (out - y)) exactly
    rng = np.random.default_rng(3)
    w1 = np.abs(rng.normal(size=(M.F, M.H)))
    w2 = np.abs(rng.normal(size=(M.H, M.H)))
    p = M.pack(w1, np.ones(M.H), w2, np.ones(M.H), rng.normal(size=M.H), 0.5)
    x = np.abs(rng.normal(size=(9, M.F)))
    y = rng.normal(size=9)
    out, _ = M.forward(p, x)
    assert np.allclose(out, x @ w1 @ w2 @ M.unpack(p)[4] + (np.ones(M.H) @ w2 + 1) @ M.unpack(p)[4]
                       + 0.5)
    g = M.grad(p, x, y)
    assert np.isclose(g[M.OFF_B3], np.mean(2 * (out - y)))


def test_features_match_oracle():
    rng = np.random.default_rng(5)
    m = rng.integers(1, 900, 60)
    n = rng.integers(1, 700, 60)
    k = rng.integers(9, 1500, 60)
    for q, (bm, bn, bk, wave, min_waves) in enumerate(gemm_configs(with_kind=True)):
        got = shape_features(m, n, k, q, n_cu=256)
        ref = M.features(m, n, k, bm, bn, bk, wave, n_cu=256, min_waves=min_waves)
        assert np.allclose(got, ref, rtol=1e-5, atol=1e-6), (q, np.abs(got - ref).max())
    assert not shape_features([], [], [], 0).any()
