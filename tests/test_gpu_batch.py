"""Batched contractions (one work list for many plans) and HIP-graph
capture: the batch must reproduce each member plan's output bit for bit
(same per-element summation order), the member outputs must match the
oracle, and every graph replay must equal the eager execute bit for bit."""
import numpy as np
import pytest
import torch

from helpers import O, W, from_workload, rel_fro
from torque_amd import BatchPlan, BlockSparseTensor, default_handle

pytestmark = pytest.mark.gpu


def make_batch(cuda, n, seed, precision="native", dim_lo=8, dim_hi=40):
    pairs = W.batched(n, seed=seed, dim_lo=dim_lo, dim_hi=dim_hi)
    ga, gb, plans, refs = [], [], [], []
    for a, b, legs in pairs:
        A = BlockSparseTensor(a.dims, a.data, a.begin, a.end, device=cuda)
        B = BlockSparseTensor(b.dims, b.data, b.begin, b.end, device=cuda)
        ga.append(A)
        gb.append(B)
        plans.append(A.plan_contract(B, legs, precision))
        refs.append((a, b, legs))
    a_off = np.cumsum([0] + [t.data.numel() for t in ga])[:-1]
    b_off = np.cumsum([0] + [t.data.numel() for t in gb])[:-1]
    c_off = np.cumsum([0] + [p.out_n_elem for p in plans])[:-1]
    a_all = torch.cat([t.data for t in ga])
    b_all = torch.cat([t.data for t in gb])
    c_total = int(sum(p.out_n_elem for p in plans))
    batch = BatchPlan(default_handle(), plans, a_off, b_off, c_off)
    return ga, gb, plans, refs, a_all, b_all, c_total, c_off, batch


@pytest.mark.parametrize("precision", ["native", "bf16x3"])
def test_batch_matches_members_and_oracle(cuda, precision):
    ga, gb, plans, refs, a_all, b_all, c_total, c_off, batch = make_batch(cuda, 12, 3, precision)
    c = torch.full((c_total,), float("nan"), dtype=torch.float64, device=cuda)
    batch.execute(a_all, b_all, c)
    torch.cuda.synchronize()
    got = c.cpu().numpy()
    assert not np.isnan(got).any()
    for i, p in enumerate(plans):
        out = torch.empty(p.out_n_elem, dtype=torch.float64, device=cuda)
        p.execute(ga[i].data, gb[i].data, out)
        mine = got[c_off[i]: c_off[i] + p.out_n_elem]
        assert np.array_equal(mine, out.cpu().numpy()), i
        a, b, legs = refs[i]
        (A, da), (B, db) = from_workload(a), from_workload(b)
        ref, _ = O.contract(A, da, B, db, legs)
        assert rel_fro(mine, ref) < (1e-12 if precision == "native" else 1e-6), i


def test_batch_every_config(cuda):
    ga, gb, plans, refs, a_all, b_all, c_total, c_off, batch = make_batch(cuda, 6, 8)
    c0 = torch.empty(c_total, dtype=torch.float64, device=cuda)
    batch.execute(a_all, b_all, c0)
    from torque_amd import gemm_configs
    for cfg in range(len(gemm_configs())):
        batch.set_config(cfg)
        c = torch.empty_like(c0)
        batch.execute(a_all, b_all, c)
        torch.cuda.synchronize()
        assert torch.equal(c, c0), cfg


def test_graph_replay_equals_eager(cuda):
    ga, gb, plans, refs, a_all, b_all, c_total, c_off, batch = make_batch(cuda, 8, 5)
    eager = torch.empty(c_total, dtype=torch.float64, device=cuda)
    batch.execute(a_all, b_all, eager)
    torch.cuda.synchronize()
    c = torch.zeros(c_total, dtype=torch.float64, device=cuda)
    batch.capture(a_all, b_all, c)
    for r in range(3):
        c.zero_()
        torch.cuda.synchronize()
        batch.graph_launch()
        torch.cuda.synchronize()
        assert torch.equal(c, eager), r
        # eager work between replays
        tmp = torch.empty_like(eager)
        batch.execute(a_all, b_all, tmp)
        torch.cuda.synchronize()
    # a single plan captures too
    p = plans[0]
    out = torch.zeros(p.out_n_elem, dtype=torch.float64, device=cuda)
    ref = torch.empty_like(out)
    p.execute(ga[0].data, gb[0].data, ref)
    p.capture(ga[0].data, gb[0].data, out)
    for _ in range(3):
        out.zero_()
        p.graph_launch()
        torch.cuda.synchronize()
        assert torch.equal(out, ref)


def test_setters_release_captured_graph(cuda):
    # a captured graph holds the tile lists' addresses: changing the
    # configuration (which may rebuild them) releases it, and a replay is
    # refused until the plan is captured again
    from torque_amd import TorqueError
    ga, gb, plans, refs, a_all, b_all, c_total, c_off, batch = make_batch(cuda, 6, 11)
    eager = torch.empty(c_total, dtype=torch.float64, device=cuda)
    batch.execute(a_all, b_all, eager)
    c = torch.zeros(c_total, dtype=torch.float64, device=cuda)
    for change in (lambda: batch.set_config(0), lambda: batch.set_class_config(1, 3),
                   lambda: batch.set_small_k(4),
                   lambda: batch.autotune(a_all, b_all, c, reps=1)):
        batch.capture(a_all, b_all, c)
        change()
        with pytest.raises(TorqueError, match="not captured"):
            batch.graph_launch()
        now = torch.empty_like(eager)
        batch.execute(a_all, b_all, now)
        torch.cuda.synchronize()
        assert torch.allclose(now, eager, rtol=1e-12, atol=1e-12)
        batch.capture(a_all, b_all, c)
        c.zero_()
        batch.graph_launch()
        torch.cuda.synchronize()
        assert torch.equal(c, now)
