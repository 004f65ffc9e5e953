"""The cost model's GPU kernels against the FP64 oracle (oracle/mlp.py),
and what the trained model gains over the analytic configuration choice.

Gradients and forward values from the HIP kernels (FP32) must match the
oracle's to a relative error of 1e-4 (the oracle's backpropagation is pinned
by finite differences in test_costmodel.py). The gain test trains the model
on autotuner timings of one set of contractions and compares, on held-out
contractions, the regret (time of the chosen configuration over the measured
best, minus one) of the model's choice with the analytic model's."""
import numpy as np
import pytest
import torch

from helpers import W
from oracle import mlp as M
from torque_amd import BlockSparseTensor
from torque_amd.costmodel import CostModel, record

pytestmark = pytest.mark.gpu

TOL = 1e-4


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / max(np.linalg.norm(b), 1e-30))


@pytest.mark.parametrize("n", [1, 7, 300, 1500])
def test_gradient_matches_oracle(cuda, n):
    rng = np.random.default_rng(n)
    model = CostModel(seed=n)
    p = rng.normal(0, 0.2, M.N_PARAMS).astype(np.float32)
    model.set_params(p)
    assert np.array_equal(model.params(), p)
    x = rng.normal(size=(n, M.F)).astype(np.float32)
    y = rng.normal(size=n).astype(np.float32)
    g, loss = model.gradient(x, y)
    g_ref = M.grad(p.astype(np.float64), x, y)
    assert rel(g, g_ref) < TOL
    # every parameter group on its own
    for lo, hi in [(M.OFF_W1, M.OFF_B1), (M.OFF_B1, M.OFF_W2), (M.OFF_W2, M.OFF_B2),
                   (M.OFF_B2, M.OFF_W3), (M.OFF_W3, M.OFF_B3), (M.OFF_B3, M.N_PARAMS)]:
        assert rel(g[lo:hi], g_ref[lo:hi]) < TOL, (lo, hi)
    assert abs(loss - M.loss(p, x, y)) <= TOL * M.loss(p, x, y)
    out = model.predict(x, raw=True)
    ref, _ = M.forward(p, x)
    assert rel(out, ref) < TOL


def test_fit_learns_a_smooth_function(cuda):
    rng = np.random.default_rng(11)
    x = rng.normal(size=(800, M.F)).astype(np.float32)
    y = (np.sin(x[:, 0]) + 0.5 * x[:, 1] * x[:, 2] - x[:, 3] ** 2).astype(np.float32)
    model = CostModel(seed=1)
    final = model.fit(x, y, steps=1500, lr=3e-3)
    assert final < 0.05  # normalised MSE (1.0 = predicting the mean)
    pred = model.predict(x)
    assert np.mean((pred - y) ** 2) / np.var(y) < 0.05


def timed_contractions(cuda, n, seed0):
    """Autotuner records of n U(1) rank-4 x rank-4 contractions whose leg
    dimensions span 8..160: (plan, analytic choice, x, y, times)."""
    ranges = [(8, 40), (16, 64), (24, 80), (40, 96), (60, 120), (12, 160)]
    rows = []
    for s in range(n):
        lo, hi = ranges[s % len(ranges)]
        a, b, legs = W.batched(1, seed=seed0 + s, dim_lo=lo, dim_hi=hi)[0]
        A = BlockSparseTensor(a.dims, a.data, a.begin, a.end, device=cuda)
        B = BlockSparseTensor(b.dims, b.data, b.begin, b.end, device=cuda)
        plan = A.plan_contract(B, legs)
        c = torch.empty(max(plan.out_n_elem, 1), dtype=torch.float64, device=cuda)
        heuristic = plan.config()  # the analytic choice, before any timing
        x, y, times = record(plan, A.data, B.data, c, reps=3)
        rows.append((plan, heuristic, x, y, times))
    return rows


def test_trained_model_beats_analytic_choice(cuda):
    rows = timed_contractions(cuda, 140, 1000)
    train, test = rows[:100], rows[100:]
    x = np.concatenate([r[2] for r in train])
    y = np.concatenate([r[3] for r in train])
    model = CostModel(seed=0)
    model.fit(x, y, steps=1000, lr=3e-3)
    reg_model, reg_heur = [], []
    for plan, heuristic, _, _, times in test:
        chosen = model.choose(plan)
        for cls in range(2):
            t = np.array(times[cls])
            if not (t > 0).any():
                continue
            best = t[t > 0].min()
            reg_model.append(t[chosen[cls]] / best - 1)
            reg_heur.append(t[heuristic[cls]] / best - 1)
    mm, mh = float(np.mean(reg_model)), float(np.mean(reg_heur))
    print("mean regret: cost model %.3f, analytic %.3f over %d held-out tile classes"
          % (mm, mh, len(reg_model)))
    assert len(reg_model) >= 40
    assert mm < 0.6 * mh, (mm, mh)
    assert mm < 0.3, (mm, mh)
