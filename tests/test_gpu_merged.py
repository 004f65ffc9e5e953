"""Merged-output contraction plans (tq_bs_contract_plan_merged) against the
oracle: one output block per distinct output block, holding the sum of the
oracle's unmerged per-pair blocks with those begin and end points, in the
order of first occurrence; the dense image equals the unmerged one.
FP64 to 1e-12, the FP32 tier to 1e-5 and split-BF16 to 1e-6 (relative
Frobenius, overall and per block)."""
import numpy as np
import pytest
import torch

from helpers import O, W, from_workload, permuted_strides, random_blocks, rel_fro, tensor, \
    torch_tensor

pytestmark = pytest.mark.gpu

TOL = {"native": 1e-12, "fp32": 1e-5, "bf16x3": 1e-6}


def merged_keys(ref):
    """Group of each oracle pair and the first pair of each group, by equal
    output begin and end points, groups in order of first occurrence."""
    keys = np.concatenate([ref["out_begin"], ref["out_end"]], axis=1)
    _, first, inv = np.unique(keys, axis=0, return_index=True, return_inverse=True)
    order = np.argsort(first, kind="stable")
    rank = np.empty_like(order)
    rank[order] = np.arange(order.size)
    return rank[inv.reshape(-1)], first[order]


def merged_reference(ref_c, ref):
    """Sum of the unmerged blocks of each group, blocks concatenated."""
    grp, first = merged_keys(ref)
    sizes = (ref["m"] * ref["n"])[first]
    offs = np.concatenate([[0], np.cumsum(sizes)])
    out = np.zeros(int(offs[-1]))
    for p, g in enumerate(grp):
        o = ref["out_offsets"][p]
        out[offs[g]:offs[g + 1]] += ref_c[o:o + sizes[g]]
    return out, grp, first, offs


def check_merged(plan, got, ref_c, ref, tol):
    want, grp, first, offs = merged_reference(ref_c, ref)
    assert plan.n_pairs == first.size and plan.n_matched == ref["m"].size
    assert np.array_equal(plan.out_begin, ref["out_begin"][first])
    assert np.array_equal(plan.out_end, ref["out_end"][first])
    assert np.array_equal(plan.out_offsets, offs[:-1])
    assert plan.out_n_elem == offs[-1]
    p = plan.pairs()
    for k in ("a_block", "b_block", "a_offset", "b_offset", "m", "n", "k"):
        assert np.array_equal(p[k], ref[k]), k
    assert rel_fro(got[:want.size], want) <= tol
    worst = max(rel_fro(got[offs[g]:offs[g + 1]], want[offs[g]:offs[g + 1]])
                for g in range(first.size))
    assert worst <= tol
    return want


def run(a, da, b, db, legs, cuda, precision="native"):
    ga, gb = torch_tensor(a, da, cuda), torch_tensor(b, db, cuda)
    plan = ga.plan_contract(gb, legs, precision, merged=True)
    out = torch.full((max(plan.out_n_elem, 1),), float("nan"), dtype=torch.float64,
                     device=cuda)
    plan.execute(ga.data, gb.data, out)
    torch.cuda.synchronize()
    return plan, out.cpu().numpy()


@pytest.mark.parametrize("precision", ["native", "fp32", "bf16x3"])
def test_c1_merged(cuda, precision):
    a, b, legs = W.c1(0)
    A, da = from_workload(a)
    B, db = from_workload(b)
    ref_c, ref = O.contract(A, da, B, db, legs)
    plan, got = run(A, da, B, db, legs, cuda, precision)
    # 45 matched pairs write 19 distinct output blocks
    assert plan.n_pairs == 19 and plan.n_matched == 45
    check_merged(plan, got, ref_c, ref, TOL[precision])
    # dense cross-check against the dense contraction of the densified inputs
    out = O.BS(plan.out_dims, plan.out_begin, plan.out_end)
    dense_ref, _ = O.dense_contract(a.dims, W.densify(a), b.dims, W.densify(b), legs)
    assert rel_fro(O.densify(out, got), dense_ref) <= TOL[precision]


@pytest.mark.parametrize("seed", range(4))
def test_overlapping_blocks_merged(cuda, seed):
    # general blocks: overlapping, partly matched, stored in random leg
    # orders; equal output keys arise from A (or B) blocks with equal free
    # extents
    rng = np.random.default_rng(900 + seed)
    dims_a, dims_b = [4, 5, 3, 6], [3, 6, 5]
    ba, ea = random_blocks(rng, dims_a, 9)
    ba[1::3, :2], ea[1::3, :2] = ba[0::3, :2][:ba[1::3].shape[0]], ea[0::3, :2][:ea[1::3].shape[0]]
    bb, eb = random_blocks(rng, dims_b, 7)
    sa = permuted_strides(ba, ea, rng) if seed % 2 else None
    a, da = tensor(dims_a, ba, ea, rng, sa)
    b, db = tensor(dims_b, bb, eb, rng)
    legs = [[2, 0], [3, 1]]
    ref_c, ref = O.contract(a, da, b, db, legs)
    if ref["m"].size == 0:
        pytest.skip("no matching blocks for this seed")
    for precision in ("native", "bf16x3"):
        plan, got = run(a, da, b, db, legs, cuda, precision)
        check_merged(plan, got, ref_c, ref, TOL[precision])
        dense = O.densify(O.BS(plan.out_dims, plan.out_begin, plan.out_end), got)
        dense_ref = O.densify(O.BS(plan.out_dims, ref["out_begin"], ref["out_end"]), ref_c)
        assert rel_fro(dense, dense_ref) <= TOL[precision]


def test_peps_d6_chi64_merged_tiers(cuda):
    E, A, Ac, l1, l2 = W.peps(6, 64, seed=2)
    TA, dA = from_workload(A)
    ref_aa, p1 = O.contract(TA, dA, TA, dA, l1)
    ta = O.BS(np.concatenate([A.dims[1:], A.dims[1:]]), p1["out_begin"], p1["out_end"])
    TE, dE = from_workload(E)
    ref_c, ref = O.contract(TE, dE, ta, ref_aa, l2)
    for precision in ("native", "fp32", "bf16x3"):
        plan, got = run(TE, dE, ta, ref_aa, l2, cuda, precision)
        assert plan.n_pairs < plan.n_matched
        check_merged(plan, got, ref_c, ref, TOL[precision])


def test_peps_d8_chi128_merged_sizes(cuda):
    # the chi = 128 absorption E.a: 44,872 matched pairs write 32,411
    # distinct blocks, 479,145,728 elements instead of 716,363,520; values
    # against the oracle's per-pair blocks summed on the GPU
    E, A, Ac, l1, l2 = W.peps(8, 128, seed=0, log10_scale_range=(-3.0, 0.0))
    TA, dA = from_workload(A)
    ref_aa, p1 = O.contract(TA, dA, TA, Ac.data, l1, threads=16)
    ta = O.BS(np.concatenate([A.dims[1:], A.dims[1:]]), p1["out_begin"], p1["out_end"])
    TE, dE = from_workload(E)
    ref_c, ref = O.contract(TE, dE, ta, ref_aa, l2, threads=16)
    grp, first = merged_keys(ref)
    sizes = (ref["m"] * ref["n"])[first]
    offs = np.concatenate([[0], np.cumsum(sizes)])
    assert first.size == 32411 and offs[-1] == 479145728
    # element e of pair p's block lands at offs[grp[p]] + e of the merged output
    shift = torch.as_tensor(offs[grp] - ref["out_offsets"], device=cuda)
    pair_sizes = torch.as_tensor(ref["m"] * ref["n"], device=cuda)
    dst = torch.repeat_interleave(shift, pair_sizes)
    dst += torch.arange(ref_c.size, device=cuda)
    want = torch.zeros(int(offs[-1]), dtype=torch.float64, device=cuda)
    want.index_add_(0, dst, torch.from_numpy(ref_c).to(cuda))
    del dst
    gE = torch_tensor(TE, dE, cuda)
    ga = torch_tensor(ta, ref_aa, cuda)
    lens = torch.as_tensor(sizes, device=cuda)
    for precision in ("native", "bf16x3"):
        plan = gE.plan_contract(ga, l2, precision, merged=True)
        assert plan.n_pairs == 32411 and plan.out_n_elem == 479145728
        out = torch.empty(plan.out_n_elem, dtype=torch.float64, device=cuda)
        plan.execute(gE.data, ga.data, out)
        torch.cuda.synchronize()
        d2 = torch.segment_reduce((out - want).square_(), "sum", lengths=lens)
        r2 = torch.segment_reduce(want.square(), "sum", lengths=lens)
        overall = float(torch.sqrt(d2.sum() / r2.sum()))
        worst = float(torch.sqrt((d2 / r2).max()))
        print("merged D=8 chi=128 %s: rel err %.2e, block max %.2e" % (precision, overall, worst))
        assert overall <= TOL[precision] and worst <= TOL[precision]
        plan.destroy()
        del out


def shared_output_blocks(n_members, w, rng, m=37, n=70):
    """A (m x K) and B (K x n) whose blocks pair up along K only: every pair
    writes the same output block, so one group of n_members pairs, each with
    w contracted values (pairs stored in opposite leg orders)."""
    K = n_members * w
    ba = np.array([[0, i * w] for i in range(n_members)], np.int64)
    ea = np.array([[m - 1, i * w + w - 1] for i in range(n_members)], np.int64)
    bb = np.array([[i * w, 0] for i in range(n_members)], np.int64)
    eb = np.array([[i * w + w - 1, n - 1] for i in range(n_members)], np.int64)
    a, da = tensor([m, K], ba, ea, rng)
    b, db = tensor([K, n], bb, eb, rng)
    return a, da, b, db


@pytest.mark.parametrize("n_members,w,streamed", [(1, 8, True), (2, 5, True), (2, 8, True),
                                                  (3, 4, False), (2, 9, False)])
def test_streamed_groups(cuda, n_members, w, streamed):
    # a merged plan streams a group (no MFMA tiles, no group panels) when
    # its pairs are short (K <= 8) and at most two; otherwise it runs the
    # K-concatenated group GEMM; either way the group's block is the sum of
    # its pairs' products
    rng = np.random.default_rng(40 + 10 * n_members + w)
    a, da, b, db = shared_output_blocks(n_members, w, rng)
    legs = [[1, 0]]
    ref_c, ref = O.contract(a, da, b, db, legs)
    for precision in ("native", "fp32"):
        plan, got = run(a, da, b, db, legs, cuda, precision)
        assert plan.n_pairs == 1 and plan.n_matched == n_members
        tiles, segs = plan.work()
        assert (tiles == 0) == streamed, (tiles, segs)
        check_merged(plan, got, ref_c, ref, TOL[precision])
    # the FP32 data path
    ga = torch_tensor(a, da, cuda, torch.float32)
    gb = torch_tensor(b, db, cuda, torch.float32)
    plan = ga.plan_contract(gb, legs, merged=True)
    out = torch.full((plan.out_n_elem,), float("nan"), dtype=torch.float32, device=cuda)
    plan.execute(ga.data, gb.data, out)
    torch.cuda.synchronize()
    want, _, _, _ = merged_reference(ref_c, ref)
    assert rel_fro(out.cpu().numpy().astype(np.float64), want) <= 1e-5


def test_streamed_groups_refuse_batch(cuda):
    # a merged plan that streams its groups keeps its own member list: it
    # does not join a batch
    from torque_amd import BatchPlan, TorqueError, default_handle
    rng = np.random.default_rng(77)
    a, da, b, db = shared_output_blocks(2, 4, rng)
    ga, gb = torch_tensor(a, da, cuda), torch_tensor(b, db, cuda)
    plan = ga.plan_contract(gb, [[1, 0]], merged=True)
    with pytest.raises(TorqueError, match="streams its groups"):
        BatchPlan(default_handle(), [plan], [0], [0], [0])
