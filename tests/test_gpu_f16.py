"""FP16 data (the reference's `half` instantiation): contractions on FP16
MFMA with FP32 accumulation and FP16 output, against the oracle on the same
FP16-rounded inputs (relative Frobenius error 1e-3: the output rounding is
2^-11 relative); permutations bit-exact; the network scalar in FP64."""
import numpy as np
import pytest
import torch

from helpers import O, W, from_workload, permuted_strides, random_blocks, rel_fro, tensor
from torque_amd import BlockSparseTensor, TorqueError, gemm_configs

pytestmark = pytest.mark.gpu

TOL16 = 1e-3


def f16(x):
    return np.asarray(x, np.float64).astype(np.float16).astype(np.float64)


def run16(a, da, b, db, legs, cuda, cfg=None, merged=False, split=None):
    ga = BlockSparseTensor(a.dims, torch.tensor(f16(da)), a.begin, a.end, a.strides,
                           dtype=torch.float16, device=cuda)
    gb = BlockSparseTensor(b.dims, torch.tensor(f16(db)), b.begin, b.end, b.strides,
                           dtype=torch.float16, device=cuda)
    plan = ga.plan_contract(gb, legs, merged=merged)
    if cfg is not None:
        plan.set_config(cfg)
    if split is not None:
        plan.set_class_split(0, split)
        plan.set_class_split(1, split)
    out = torch.empty(max(plan.out_n_elem, 1), dtype=torch.float16, device=cuda)
    plan.execute(ga.data, gb.data, out)
    torch.cuda.synchronize()
    return plan, out.cpu().double().numpy()


def test_c1_every_configuration(cuda):
    a, b, legs = W.c1(5)
    (A, da), (B, db) = from_workload(a), from_workload(b)
    ref_c, ref = O.contract(A, f16(da), B, f16(db), legs)
    n_ok = 0
    for cfg in range(len(gemm_configs())):
        try:
            plan, got = run16(A, da, B, db, legs, cuda, cfg=cfg)
        except TorqueError as e:
            assert "FP16" in str(e)  # wave-tile and short-k-step configurations
            continue
        n_ok += 1
        assert np.array_equal(plan.out_begin, ref["out_begin"])
        assert rel_fro(got, ref_c) < TOL16, cfg
    assert n_ok == 8


@pytest.mark.parametrize("seed", range(3))
def test_overlapping_blocks_and_layouts(cuda, seed):
    rng = np.random.default_rng(1600 + seed)
    dims_a, dims_b = [9, 6, 12, 7], [12, 7, 10]
    ba, ea = random_blocks(rng, dims_a, 7)
    bb, eb = random_blocks(rng, dims_b, 6)
    a, da = tensor(dims_a, ba, ea, rng, permuted_strides(ba, ea, rng) if seed else None)
    b, db = tensor(dims_b, bb, eb, rng)
    legs = [[2, 0], [3, 1]]
    ref_c, ref = O.contract(a, f16(da), b, f16(db), legs)
    for small_k in (None,):
        plan, got = run16(a, da, b, db, legs, cuda)
        if plan.n_pairs:
            assert rel_fro(got, ref_c) < TOL16
    # merged output and split-K in FP16
    plan, got = run16(a, da, b, db, legs, cuda, merged=True)
    if plan.n_pairs:
        dense = O.densify(O.BS(plan.out_dims, plan.out_begin, plan.out_end), got)
        dref = O.densify(O.BS(plan.out_dims, ref["out_begin"], ref["out_end"]), ref_c)
        assert rel_fro(dense, dref) < TOL16
    plan, got = run16(a, da, b, db, legs, cuda, split=4)
    if plan.n_pairs:
        assert rel_fro(got, ref_c) < TOL16


def test_large_k_blocks(cuda):
    # block-diagonal GEMMs with K up to 400 through the MFMA tiles
    rng = np.random.default_rng(17)
    sizes = [(40, 400, 72), (130, 300, 150), (7, 90, 5), (64, 64, 64)]
    r, k, n = (np.array(x) for x in zip(*sizes))

    def mat(rows, cols):
        b0 = np.concatenate([[0], np.cumsum(rows)[:-1]])
        b1 = np.concatenate([[0], np.cumsum(cols)[:-1]])
        begin = np.stack([b0, b1], 1).astype(np.int64)
        end = np.stack([b0 + rows - 1, b1 + cols - 1], 1).astype(np.int64)
        return O.BS([rows.sum(), cols.sum()], begin, end), rng.standard_normal(
            int((rows * cols).sum()))

    (A, da), (B, db) = mat(r, k), mat(k, n)
    ref_c, _ = O.contract(A, f16(da), B, f16(db), [[1, 0]])
    plan, got = run16(A, da, B, db, [[1, 0]], cuda)
    assert plan.work()[0] > 0
    assert rel_fro(got, ref_c) < TOL16


def test_permute_bit_exact_and_scalar(cuda):
    rng = np.random.default_rng(5)
    dims = [5, 6, 7, 3]
    ba, ea = random_blocks(rng, dims, 6)
    a, da = tensor(dims, ba, ea, rng)
    t = BlockSparseTensor(dims, torch.tensor(f16(da)), ba, ea, dtype=torch.float16,
                          device=cuda)
    perm = [2, 0, 3, 1]
    pt = t.hard_transpose(perm)
    ref = O.bs_permute(a, f16(da), perm)
    assert np.array_equal(pt.data.cpu().double().numpy()[:ref.size], ref)
    # the network scalar reduces in FP64 and is rounded once to FP16
    s, _ = O.contract(a, f16(da), a, f16(da), [[0, 0], [1, 1], [2, 2], [3, 3]])
    got = t.contract(t, [[0, 0], [1, 1], [2, 2], [3, 3]]).to_number()
    assert got == float(np.float16(s[0]))
