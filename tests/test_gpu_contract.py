"""Block-sparse contraction on the GPU against the oracle.

Integer maps (matched pairs, output begin/end, sub-block offsets, GEMM
shapes, output offsets) must be bit-exact. FP64 data must match the oracle
to a relative Frobenius error of 1e-12 (summation order differs); the FP32
tier to 1e-5; FP32 data to 1e-5.
"""
import numpy as np
import pytest
import torch

from helpers import O, W, dense_of, from_workload, permuted_strides, random_blocks, rel_fro, \
    tensor, torch_tensor
from torque_amd import BlockSparseTensor, DenseTensor, TorqueError, dot_f64, gemm_configs, \
    normalize_f64, nrm2_f64, scale_f64

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-12


def gpu_contract(a, da, b, db, legs, cuda, precision="native", cfg=None, dtype=torch.float64,
                 small_k=None):
    ga = torch_tensor(a, da, cuda, dtype)
    gb = torch_tensor(b, db, cuda, dtype)
    plan = ga.plan_contract(gb, legs, precision)
    if small_k is not None:
        plan.set_small_k(small_k)
    if cfg is not None:
        plan.set_config(cfg)
    out = torch.empty(max(plan.out_n_elem, 1), dtype=dtype, device=cuda)
    plan.execute(ga.data, gb.data, out)
    torch.cuda.synchronize()
    return plan, out.cpu().double().numpy()[: max(plan.out_n_elem, 1)]


def check_maps(plan, ref):
    p = plan.pairs()
    for k in ("a_block", "b_block", "a_offset", "b_offset", "m", "n", "k"):
        assert np.array_equal(p[k], ref[k]), k
    assert np.array_equal(plan.out_begin, ref["out_begin"])
    assert np.array_equal(plan.out_end, ref["out_end"])
    if plan.out_rank > 0:
        assert np.array_equal(plan.out_offsets, ref["out_offsets"])


# ---- the reference's GPU known answers (test/gpu/block_sparse_test.cu) ---

def bst(dims, blocks, cuda, dtype=torch.float32):
    t = BlockSparseTensor(dims, dtype=dtype, device=cuda)
    for b, e, v, s in blocks:
        t.append_block(v, b, e, s)
    return t


def test_known_answers_gpu(cuda):
    t = bst([5], [([0], [1], [0, 1], [1]), ([2], [4], [2, 3, 4], [1])], cuda)
    assert t.contract(t, [[0, 0]]).to_number() == 30
    vec = np.array([0, 1, 0, 0, 2, 3, 0, 0, 0, 0, 4, 5, 0, 0, 6, 7], np.float32)
    dense = DenseTensor([4, 4], vec, dtype=torch.float32, device=cuda)
    sq = dense.contract(dense, [[1, 0]])
    s = bst([4, 4], [([0, 0], [1, 1], [0, 1, 2, 3], [1, 2]),
                     ([2, 2], [3, 3], [4, 5, 6, 7], [1, 2])], cuda)
    ss = s.contract(s, [[1, 0]])
    for i in range(4):
        for j in range(4):
            assert sq.query([i, j]) == ss.query([i, j])
    t = bst([3, 3], [([0, 0], [1, 1], [0, 1, 2, 3], [1, 2]),
                     ([1, 1], [2, 2], [0, 7, 8, 9], [1, 2])], cuda)
    ref_a = np.array([[0, 1, 0], [2, 3, 7], [0, 8, 9]], np.float64).T
    assert t.contract(t, [[1, 0], [0, 1]]).to_number() == np.sum(ref_a * ref_a.T)
    t = bst([2, 2, 2], [([0, 0, 0], [1, 1, 0], [0, 1, 2, 3], [1, 2, 4]),
                        ([0, 0, 1], [1, 1, 1], [4, 5, 6, 7], [1, 2, 4])], cuda)
    v = bst([2], [([0], [1], [1, 2], [1])], cuda)
    c = t.contract(v, [[0, 0]])
    assert [c.query(i) for i in ([0, 0], [1, 0], [0, 1], [1, 1])] == [2, 8, 14, 20]
    c = t.contract(v, [[1, 0]])
    assert [c.query(i) for i in ([0, 0], [1, 0], [0, 1], [1, 1])] == [4, 7, 16, 19]
    assert t.contract(v, [[2, 0]]).query([1, 0]) == 11
    t = bst([2, 2, 3], [([0, 0, 2], [1, 1, 2], [8, 9, 10, 11], [1, 2, 4]),
                        ([0, 0, 1], [1, 1, 1], [4, 5, 6, 7], [1, 2, 4]),
                        ([0, 0, 0], [1, 1, 0], [0, 1, 2, 3], [1, 2, 4])], cuda)
    m = bst([2, 2], [([0, 0], [1, 1], [1, 2, 3, 4], [1, 2])], cuda)
    c = t.contract(m, [[0, 1], [1, 0]])
    assert [c.query([i]) for i in range(3)] == [19, 59, 99]


def test_dense_contract_known_and_random(cuda):
    rng = np.random.default_rng(2)
    a = rng.standard_normal(3 * 4 * 5)
    b = rng.standard_normal(5 * 3 * 2)
    A = DenseTensor([3, 4, 5], a, device=cuda)
    B = DenseTensor([5, 3, 2], b, device=cuda)
    C = A.contract(B, [[2, 0], [0, 1]])
    ref, od = O.dense_contract([3, 4, 5], a, [5, 3, 2], b, [[2, 0], [0, 1]])
    assert list(C.dimension) == od
    assert rel_fro(C.data.cpu().numpy(), ref) < FP64_TOL
    v = DenseTensor([5], np.arange(5), device=cuda)
    assert v.contract(v, [[0, 0]]).to_number() == 30


# ---- integer maps and FP64 data against the oracle -----------------------

def test_c1_config(cuda):
    a, b, legs = W.c1(0)
    A, da = from_workload(a)
    B, db = from_workload(b)
    ref_c, ref = O.contract(A, da, B, db, legs)
    plan, got = gpu_contract(A, da, B, db, legs, cuda)
    assert plan.n_pairs == 45 and plan.out_n_elem == 44544
    assert plan.flops == 2834432
    check_maps(plan, ref)
    assert rel_fro(got, ref_c) < FP64_TOL
    # dense cross-check (config 1)
    out = O.BS(plan.out_dims, plan.out_begin, plan.out_end)
    dense_ref, _ = O.dense_contract(a.dims, W.densify(a), b.dims, W.densify(b), legs)
    assert rel_fro(O.densify(out, got), dense_ref) < FP64_TOL


@pytest.mark.parametrize("seed", range(6))
def test_random_overlapping_blocks(cuda, seed):
    rng = np.random.default_rng(500 + seed)
    dims_a = [4, 6, 3, 5]
    dims_b = [3, 5, 6]
    ba, ea = random_blocks(rng, dims_a, 7)
    bb, eb = random_blocks(rng, dims_b, 6)
    sa = permuted_strides(ba, ea, rng) if seed % 2 else None
    a, da = tensor(dims_a, ba, ea, rng, sa)
    b, db = tensor(dims_b, bb, eb, rng)
    legs = [[2, 0], [3, 1]] if seed < 3 else [[3, 1], [1, 2]]
    ref_c, ref = O.contract(a, da, b, db, legs)
    plan, got = gpu_contract(a, da, b, db, legs, cuda)
    check_maps(plan, ref)
    if plan.n_pairs:
        assert rel_fro(got, ref_c) < FP64_TOL


def test_every_gemm_config_agrees(cuda):
    a, b, legs = W.c1(3)
    A, da = from_workload(a)
    B, db = from_workload(b)
    ref_c, _ = O.contract(A, da, B, db, legs)
    for cfg in range(len(gemm_configs())):
        _, got = gpu_contract(A, da, B, db, legs, cuda, cfg=cfg)
        assert rel_fro(got, ref_c) < FP64_TOL, cfg
        # the FP32 tier and FP32 data run the same configurations
        _, got = gpu_contract(A, da, B, db, legs, cuda, cfg=cfg, precision="fp32")
        assert rel_fro(got, ref_c) < 1e-5, cfg
        _, got = gpu_contract(A, da, B, db, legs, cuda, cfg=cfg, dtype=torch.float32)
        assert rel_fro(got, ref_c) < 1e-5, cfg


@pytest.mark.parametrize("small_k", [0, 4, 8])
def test_small_k_stream_and_mfma_agree(cuda, small_k):
    # the same contraction through MFMA tiles only, the default split, and
    # the streaming kernel for every pair
    rng = np.random.default_rng(31)
    dims_a = [7, 9, 4, 3]
    dims_b = [4, 3, 5]
    ba, ea = random_blocks(rng, dims_a, 8)
    bb, eb = random_blocks(rng, dims_b, 6)
    a, da = tensor(dims_a, ba, ea, rng)
    b, db = tensor(dims_b, bb, eb, rng)
    legs = [[2, 0], [3, 1]]
    ref_c, ref = O.contract(a, da, b, db, legs)
    plan, got = gpu_contract(a, da, b, db, legs, cuda, small_k=small_k)
    check_maps(plan, ref)
    assert rel_fro(got, ref_c) < FP64_TOL
    _, got32 = gpu_contract(a, da, b, db, legs, cuda, small_k=small_k, precision="fp32")
    assert rel_fro(got32, ref_c) < 1e-5
    # the split-BF16 tier streams short pairs from its split panels
    _, got3 = gpu_contract(a, da, b, db, legs, cuda, small_k=small_k, precision="bf16x3")
    assert rel_fro(got3, ref_c) < 1e-6
    with pytest.raises(TorqueError, match="small-K"):
        plan.set_small_k(9)


@pytest.mark.parametrize("layout", ["k_fast", "n_fast", "mixed"])
def test_small_k_wide_blocks_are_sliced(cuda, layout):
    # streamed pairs (K <= 8) wider than 64 columns are cut into 64-column
    # slices, each with its own B and C offsets; every tier must match the
    # oracle whatever the B layout (k fastest, n fastest, or per block). The
    # streaming kernels walk columns with one loop per K: every K from 1 to 8
    # occurs.
    rng = np.random.default_rng(17)
    nb = 12
    r = rng.integers(40, 151, nb)
    k = rng.permutation(np.concatenate([np.arange(1, 9), rng.integers(1, 9, nb - 8)]))
    n = rng.integers(65, 201, nb)

    def diag(rows, cols):
        b0 = np.concatenate([[0], np.cumsum(rows)[:-1]])
        b1 = np.concatenate([[0], np.cumsum(cols)[:-1]])
        begin = np.stack([b0, b1], 1).astype(np.int64)
        end = np.stack([b0 + rows - 1, b1 + cols - 1], 1).astype(np.int64)
        return [int(rows.sum()), int(cols.sum())], begin, end

    da_, ba, ea = diag(r, k)
    a, dta = tensor(da_, ba, ea, rng)
    if layout == "n_fast":
        db_, bb, eb = diag(n, k)
        legs = [[1, 1]]
        b, dtb = tensor(db_, bb, eb, rng)
    else:
        db_, bb, eb = diag(k, n)
        legs = [[1, 0]]
        strides = permuted_strides(bb, eb, rng) if layout == "mixed" else None
        b, dtb = tensor(db_, bb, eb, rng, strides)
    ref_c, ref = O.contract(a, dta, b, dtb, legs)
    plan, got = gpu_contract(a, dta, b, dtb, legs, cuda)
    check_maps(plan, ref)
    assert rel_fro(got, ref_c) < FP64_TOL
    n_tiles, n_segments = plan.work()
    assert n_tiles == 0
    assert n_segments == int((np.ceil(r / 64) * np.ceil(n / 64)).sum())
    _, got32 = gpu_contract(a, dta, b, dtb, legs, cuda, precision="fp32")
    assert rel_fro(got32, ref_c) < 1e-5
    _, got3 = gpu_contract(a, dta, b, dtb, legs, cuda, precision="bf16x3")
    assert rel_fro(got3, ref_c) < 1e-6


def test_peps_absorption_d6_chi64(cuda):
    E, A, Ac, l1, l2 = W.peps(6, 64, seed=1)
    TA, dA = from_workload(A)
    ref_aa, ref1 = O.contract(TA, dA, TA, dA, l1)
    plan1, aa = gpu_contract(TA, dA, TA, dA, l1, cuda)
    check_maps(plan1, ref1)
    assert rel_fro(aa, ref_aa) < FP64_TOL
    ta = O.BS(plan1.out_dims, plan1.out_begin, plan1.out_end)
    TE, dE = from_workload(E)
    ref_c, ref2 = O.contract(TE, dE, ta, ref_aa, l2)
    plan2, got = gpu_contract(TE, dE, ta, ref_aa, l2, cuda)
    check_maps(plan2, ref2)
    assert rel_fro(got, ref_c) < FP64_TOL
    assert plan2.a_layout == 2 and plan2.b_layout == 2  # both permuted to GEMM layout
    # split-BF16 tier: every E.a pair has K <= 4 and is streamed from split panels
    plan3, got3 = gpu_contract(TE, dE, ta, ref_aa, l2, cuda, precision="bf16x3")
    check_maps(plan3, ref2)
    assert plan3.work()[0] == 0 and plan3.work()[1] > 0
    assert rel_fro(got3, ref_c) < 1e-6


def test_batched_pairs_large_k(cuda):
    pairs = W.batched(3, seed=9, dim_lo=40, dim_hi=96)
    for a, b, legs in pairs:
        A, da = from_workload(a)
        B, db = from_workload(b)
        ref_c, ref = O.contract(A, da, B, db, legs)
        plan, got = gpu_contract(A, da, B, db, legs, cuda)
        check_maps(plan, ref)
        assert rel_fro(got, ref_c) < FP64_TOL


def test_tile_classes(cuda):
    # pairs on both MFMA tile classes (min(M, N) >= 128 and below), tuned and
    # then forced to unrelated configurations per class
    a, b, legs = W.batched(1, seed=9, dim_lo=60, dim_hi=96)[0]
    A, da = from_workload(a)
    B, db = from_workload(b)
    ref_c, _ = O.contract(A, da, B, db, legs)
    ga = torch_tensor(A, da, cuda, torch.float64)
    gb = torch_tensor(B, db, cuda, torch.float64)
    plan = ga.plan_contract(gb, legs)
    p = plan.pairs()
    mfma = p["k"] > 8
    large = mfma & (p["m"] >= 128) & (p["n"] >= 128)
    assert large.any() and (mfma & ~large).any()
    out = torch.empty(plan.out_n_elem, dtype=torch.float64, device=cuda)
    best, times = plan.autotune(ga.data, gb.data, out, reps=1)
    assert plan.config() == best
    assert all(t > 0 for t in times[0]) and all(t > 0 for t in times[1])
    assert best[0] == int(np.argmin(times[0])) and best[1] == int(np.argmin(times[1]))
    nc = len(gemm_configs())
    for c0, c1 in [(6, 3), (3, 6), (0, nc - 1), (nc - 1, 0)]:
        plan.set_class_config(0, c0)
        plan.set_class_config(1, c1)
        assert plan.config() == [c0, c1]
        out.zero_()
        plan.execute(ga.data, gb.data, out)
        torch.cuda.synchronize()
        assert rel_fro(out.cpu().numpy(), ref_c) < FP64_TOL, (c0, c1)
    plan.set_config(-1)
    with pytest.raises(TorqueError):
        plan.set_class_config(2, 0)


def test_rank0_outer_and_empty(cuda):
    rng = np.random.default_rng(8)
    ba, ea = random_blocks(rng, [6, 5], 5)
    a, da = tensor([6, 5], ba, ea, rng)
    ref_c, _ = O.contract(a, da, a, da, [[0, 0], [1, 1]])
    plan, got = gpu_contract(a, da, a, da, [[0, 0], [1, 1]], cuda)
    assert plan.out_rank == 0 and abs(got[0] - ref_c[0]) <= 1e-12 * abs(ref_c[0])
    ref_c, ref = O.contract(a, da, a, da, np.zeros((0, 2)))
    plan, got = gpu_contract(a, da, a, da, np.zeros((0, 2)), cuda)
    check_maps(plan, ref)
    assert rel_fro(got, ref_c) < FP64_TOL
    # no matching block: empty output (the reference returns an empty tensor)
    b, db = tensor([6, 5], np.array([[0, 0]]), np.array([[0, 0]]), rng)
    c, dc = tensor([6, 5], np.array([[5, 4]]), np.array([[5, 4]]), rng)
    plan, _ = gpu_contract(b, db, c, dc, [[0, 0]], cuda)
    assert plan.n_pairs == 0


def test_dimension_mismatch_error(cuda):
    t = BlockSparseTensor([5], [0, 1], [[0]], [[1]], device=cuda)
    u = BlockSparseTensor([4], [0, 1], [[0]], [[1]], device=cuda)
    with pytest.raises(TorqueError, match="do not match"):
        t.contract(u, [[0, 0]])


def test_worklist_is_cost_sorted_permutation(cuda):
    pairs = W.batched(2, seed=4)
    a, b, legs = pairs[0]
    A, da = from_workload(a)
    B, db = from_workload(b)
    plan, _ = gpu_contract(A, da, B, db, legs, cuda)
    order = plan.worklist()
    assert sorted(order.tolist()) == list(range(plan.n_pairs))
    p = plan.pairs()
    # streamed pairs (K <= 8) by M * N * (K + 4), tiled pairs by K + 4
    cost = np.where(p["k"] <= 8, p["m"] * p["n"] * (p["k"] + 4), p["k"] + 4)
    lg = np.floor(np.log2(np.maximum(cost[order], 1)))
    assert np.all(np.diff(lg) <= 0)
    # stable inside a bucket
    for v in np.unique(lg):
        idx = order[lg == v]
        assert np.all(np.diff(idx) > 0)


def test_fp32_tier_and_fp32_data(cuda):
    a, b, legs = W.c1(5)
    A, da = from_workload(a)
    B, db = from_workload(b)
    ref_c, _ = O.contract(A, da, B, db, legs)
    _, got = gpu_contract(A, da, B, db, legs, cuda, precision="fp32")
    assert rel_fro(got, ref_c) < 1e-5
    _, got = gpu_contract(A, da, B, db, legs, cuda, dtype=torch.float32)
    assert rel_fro(got, ref_c) < 1e-5


def test_reductions(cuda):
    rng = np.random.default_rng(1)
    x = rng.standard_normal(1_000_003)
    y = rng.standard_normal(1_000_003)
    gx = torch.tensor(x, device=cuda)
    gy = torch.tensor(y, device=cuda)
    assert abs(dot_f64(gx, gy) - np.dot(x, y)) < 1e-10 * np.sqrt(len(x))
    assert abs(nrm2_f64(gx) - np.linalg.norm(x)) < 1e-12 * np.linalg.norm(x)
    scale_f64(gx, 1.0 / np.linalg.norm(x))
    assert abs(nrm2_f64(gx) - 1.0) < 1e-14
    # repeated reductions agree bitwise
    assert dot_f64(gx, gy) == dot_f64(gx, gy)
    # device-side normalisation: the norm stays on the device and equals
    # tq_nrm2_f64 bitwise (same reduction order); x / ||x|| has norm 1
    gx = torch.tensor(x * 3.5, device=cuda)
    want = nrm2_f64(gx)
    nrm = normalize_f64(gx)
    torch.cuda.synchronize()
    assert nrm.is_cuda and float(nrm[0]) == want
    assert abs(float(nrm[0]) - 3.5 * np.linalg.norm(x)) < 1e-12 * 3.5 * np.linalg.norm(x)
    assert np.array_equal(gx.cpu().numpy(), torch.tensor(x * 3.5).numpy() / want)
    assert abs(nrm2_f64(gx) - 1.0) < 1e-14
    # a zero vector keeps its values and reports norm 0
    z = torch.zeros(1000, dtype=torch.float64, device=cuda)
    assert float(normalize_f64(z)[0]) == 0.0 and not z.any()
    e = torch.zeros(1, dtype=torch.float64, device=cuda)
    assert float(normalize_f64(e[:0])[0]) == 0.0


@pytest.mark.parametrize("maker", ["c1", "batched", "overlap"])
def test_bf16x3_tier(cuda, maker):
    # relative Frobenius error against the FP64 oracle stays below 1e-6
    if maker == "c1":
        a, b, legs = W.c1(6)
        (A, da), (B, db) = from_workload(a), from_workload(b)
    elif maker == "batched":
        a, b, legs = W.batched(1, seed=13, dim_lo=40, dim_hi=96)[0]
        (A, da), (B, db) = from_workload(a), from_workload(b)
    else:
        rng = np.random.default_rng(77)
        ba, ea = random_blocks(rng, [6, 5, 7], 6)
        bb, eb = random_blocks(rng, [7, 5, 4], 5)
        A, da = tensor([6, 5, 7], ba, ea, rng)
        B, db = tensor([7, 5, 4], bb, eb, rng)
        legs = [[2, 0], [1, 1]]
    ref_c, ref = O.contract(A, da, B, db, legs)
    for cfg in range(len(gemm_configs("bf16x3"))):
        plan, got = gpu_contract(A, da, B, db, legs, cuda, precision="bf16x3", cfg=cfg)
        check_maps(plan, ref)
        if plan.n_pairs:
            assert rel_fro(got, ref_c) < 1e-6, cfg


def block_diagonal_pair(rng, n_blocks, lo, hi):
    """A (R x K) and B (K x N) block-diagonal matrices whose i-th blocks are
    r_i x k_i and k_i x n_i with sizes uniform in [lo, hi]: n_blocks
    independent GEMMs of uneven shapes."""
    r, k, n = (rng.integers(lo, hi + 1, n_blocks) for _ in range(3))

    def mat(rows, cols):
        b0 = np.concatenate([[0], np.cumsum(rows)[:-1]])
        b1 = np.concatenate([[0], np.cumsum(cols)[:-1]])
        begin = np.stack([b0, b1], 1).astype(np.int64)
        end = np.stack([b0 + rows - 1, b1 + cols - 1], 1).astype(np.int64)
        return W.Tensor([rows.sum(), cols.sum()], begin, end,
                        rng.standard_normal(int((rows * cols).sum())))

    return mat(r, k), mat(k, n), [[1, 0]]


def test_bf16x3_many_tiles_per_workgroup(cuda):
    # more tiles than resident workgroups, so every persistent workgroup runs
    # several tiles and the double-buffered kernel's loads cross from one
    # tile's last k step to the next tile's first; K is not a multiple of 32
    rng = np.random.default_rng(91)
    a, b, legs = block_diagonal_pair(rng, 96, 150, 420)
    (A, da), (B, db) = from_workload(a), from_workload(b)
    ref_c, ref = O.contract(A, da, B, db, legs, threads=8)
    plan, c64 = gpu_contract(A, da, B, db, legs, cuda)
    check_maps(plan, ref)
    assert rel_fro(c64, ref_c) < FP64_TOL
    first = None
    for cfg in range(len(gemm_configs("bf16x3"))):
        _, got = gpu_contract(A, da, B, db, legs, cuda, precision="bf16x3", cfg=cfg)
        assert rel_fro(got, ref_c) < 1e-6, cfg
        # every configuration adds the piece products in one order
        if first is None:
            first = got
        assert np.array_equal(got, first), cfg


def test_tile_counters_repeat_and_replay(cuda):
    # the 64 x 64 wave-tile configurations hand out tiles from per-XCD
    # counters that a small launch clears before every GEMM launch: repeated
    # executes, and graph replays with eager executes in between, must give
    # the same bits as each other and as a configuration without counters
    rng = np.random.default_rng(29)
    a, b, legs = block_diagonal_pair(rng, 64, 130, 300)
    (A, da), (B, db) = from_workload(a), from_workload(b)
    ref_c, _ = O.contract(A, da, B, db, legs, threads=8)
    ga, gb = torch_tensor(A, da, cuda), torch_tensor(B, db, cuda)
    configs = gemm_configs(with_kind=True)
    counted = [q for q, c in enumerate(configs) if c[3] == 1 and c[0] == 64 and c[1] == 64]
    plain = [q for q, c in enumerate(configs) if c[3] == 1 and c[0] * c[1] < 64 * 64]
    assert counted and plain
    plan = ga.plan_contract(gb, legs, "native")
    out = torch.empty(plan.out_n_elem, dtype=torch.float64, device=cuda)
    plan.set_class_split(0, 1)  # whole pairs in every configuration
    plan.set_class_split(1, 1)
    plan.set_config(plain[0])
    plan.execute(ga.data, gb.data, out)
    torch.cuda.synchronize()
    base = out.clone()
    assert rel_fro(base.cpu().numpy(), ref_c) < FP64_TOL
    for cfg in counted:
        plan.set_config(cfg)
        for r in range(3):
            out.fill_(float("nan"))
            plan.execute(ga.data, gb.data, out)
            torch.cuda.synchronize()
            assert torch.equal(out, base), (cfg, r)
        assert plan.work()[0] >= 32 * 8  # enough tiles for a grid that is a multiple of 8
        got = torch.zeros_like(out)
        plan.capture(ga.data, gb.data, got)
        for r in range(3):
            got.fill_(float("nan"))
            torch.cuda.synchronize()
            plan.graph_launch()
            torch.cuda.synchronize()
            assert torch.equal(got, base), (cfg, r)
            plan.execute(ga.data, gb.data, out)  # eager work between replays
            torch.cuda.synchronize()
            assert torch.equal(out, base), (cfg, r)


def test_bf16x3_tile_counters_repeat_and_replay(cuda):
    # the single-buffered split-BF16 configurations draw tiles from per-XCD
    # counters (thread 0 posts the workgroup's next tile in LDS): repeated
    # executes and graph replays give the bits of the double-buffered
    # configurations, which walk a fixed sequence of tiles
    rng = np.random.default_rng(43)
    a, b, legs = block_diagonal_pair(rng, 64, 130, 300)
    (A, da), (B, db) = from_workload(a), from_workload(b)
    ref_c, _ = O.contract(A, da, B, db, legs, threads=8)
    ga, gb = torch_tensor(A, da, cuda), torch_tensor(B, db, cuda)
    plan = ga.plan_contract(gb, legs, "bf16x3")
    plan.set_class_split(0, 1)
    plan.set_class_split(1, 1)
    out = torch.empty(plan.out_n_elem, dtype=torch.float64, device=cuda)
    n_cfg = len(gemm_configs("bf16x3"))
    plan.set_config(n_cfg - 1)  # a double-buffered configuration
    plan.execute(ga.data, gb.data, out)
    torch.cuda.synchronize()
    base = out.clone()
    assert rel_fro(base.cpu().numpy(), ref_c) < 1e-6
    for cfg in (0, 3, 8):  # single-buffered: 64 x 64, 128 x 128, 128 x 128 two per SIMD
        plan.set_config(cfg)
        for r in range(2):
            out.fill_(float("nan"))
            plan.execute(ga.data, gb.data, out)
            torch.cuda.synchronize()
            assert torch.equal(out, base), (cfg, r)
        got = torch.zeros_like(out)
        plan.capture(ga.data, gb.data, got)
        for r in range(3):
            got.fill_(float("nan"))
            torch.cuda.synchronize()
            plan.graph_launch()
            torch.cuda.synchronize()
            assert torch.equal(got, base), (cfg, r)
            plan.execute(ga.data, gb.data, out)
            torch.cuda.synchronize()
            assert torch.equal(out, base), (cfg, r)


@pytest.mark.parametrize("permuted", [False, True])
def test_runs_match_pair_tiles(cuda, permuted):
    # U(1) pairs that share an A block and write adjacent output blocks are
    # tiled as one GEMM by the wave-tile and split-BF16 kernels (B columns
    # found per lane in the run's pairs); the workgroup-tile kernel tiles
    # every pair alone. The two must give the same bits, with B in place or
    # packed (blocks stored in random leg orders).
    rng = np.random.default_rng(57)
    a, b, legs = W.batched(1, seed=23, dim_lo=24, dim_hi=72)[0]
    A = O.BS(a.dims, a.begin, a.end, permuted_strides(a.begin, a.end, rng) if permuted else None)
    B = O.BS(b.dims, b.begin, b.end, permuted_strides(b.begin, b.end, rng) if permuted else None)
    da, db = a.data, b.data
    ref_c, ref = O.contract(A, da, B, db, legs)
    configs = gemm_configs(with_kind=True)
    wave = [q for q, c in enumerate(configs) if c[3] == 1]
    plan, pair_tiles = gpu_contract(A, da, B, db, legs, cuda, cfg=0)
    check_maps(plan, ref)
    assert rel_fro(pair_tiles, ref_c) < FP64_TOL
    for cfg in wave:
        _, got = gpu_contract(A, da, B, db, legs, cuda, cfg=cfg)
        assert np.array_equal(got, pair_tiles), cfg
    first = None
    for cfg in range(len(gemm_configs("bf16x3"))):
        _, got = gpu_contract(A, da, B, db, legs, cuda, precision="bf16x3", cfg=cfg)
        assert rel_fro(got, ref_c) < 1e-6, cfg
        if first is None:
            first = got
        assert np.array_equal(got, first), cfg


def restride(bs, data, strides):
    """The same logical blocks stored under other index tables."""
    out = np.empty_like(data)
    for b in range(bs.begin.shape[0]):
        dims = bs.end[b] - bs.begin[b] + 1
        n = int(np.prod(dims))
        idx = np.unravel_index(np.arange(n), tuple(dims), order="F")
        pos = sum(i * int(s) for i, s in zip(idx, strides[b]))
        out[bs.offsets[b] + pos] = data[bs.offsets[b]:bs.offsets[b] + n]
    return out


def test_bf16x3_in_place_equals_split_panels(cuda):
    # plain operands are read in place by the split-BF16 GEMM and split on the
    # way to LDS (no pack pass); the same logical tensors stored in random leg
    # orders go through the split pack and panels. Both stage the same BF16
    # pieces and add the products in one order: the same bits in every
    # configuration, with small-K pairs (streamed) and runs of pairs included.
    rng = np.random.default_rng(58)
    a, b, legs = W.batched(1, seed=29, dim_lo=8, dim_hi=80)[0]
    A, da = from_workload(a)
    B, db = from_workload(b)
    ref_c, ref = O.contract(A, da, B, db, legs)
    sa = permuted_strides(a.begin, a.end, rng)
    sb = permuted_strides(b.begin, b.end, rng)
    Ap = O.BS(a.dims, a.begin, a.end, sa)
    Bp = O.BS(b.dims, b.begin, b.end, sb)
    dap, dbp = restride(A, da, sa), restride(B, db, sb)
    p = torch_tensor(A, da, cuda).plan_contract(torch_tensor(B, db, cuda), legs, "bf16x3")
    assert (p.a_layout, p.b_layout) != (2, 2)  # read in place
    assert p.pairs()["k"].min() <= 8 < p.pairs()["k"].max()
    for cfg in range(len(gemm_configs("bf16x3"))):
        plan, direct = gpu_contract(A, da, B, db, legs, cuda, precision="bf16x3", cfg=cfg)
        check_maps(plan, ref)
        assert rel_fro(direct, ref_c) < 1e-6, cfg
        plan_p, packed = gpu_contract(Ap, dap, Bp, dbp, legs, cuda, precision="bf16x3", cfg=cfg)
        assert plan_p.a_layout == 2 and plan_p.b_layout == 2
        assert np.array_equal(direct, packed), cfg


def test_split_operand_forms(cuda):
    # a split-BF16 plan with plain operands reads them in place or packs them
    # into split panels, switchable in either direction with the same bits;
    # the autotuner keeps one of the two forms; a plan whose operands are not
    # plain refuses the in-place form
    a, b, legs = W.batched(1, seed=31, dim_lo=16, dim_hi=90)[0]
    A, da = from_workload(a)
    B, db = from_workload(b)
    ref_c, _ = O.contract(A, da, B, db, legs)
    ga, gb = torch_tensor(A, da, cuda), torch_tensor(B, db, cuda)
    plan = ga.plan_contract(gb, legs, "bf16x3")
    # whole pairs (the autotuner below keeps them on this small contraction),
    # so the forms and the autotuned result compare bit for bit
    plan.set_class_split(0, 1)
    plan.set_class_split(1, 1)
    out = torch.empty(plan.out_n_elem, dtype=torch.float64, device=cuda)
    outs = {}
    for form in ("in_place", "panels", "in_place", "panels"):
        plan.set_split_operands(form)
        assert (plan.a_layout == 2) == (form == "panels")
        out.fill_(float("nan"))
        plan.execute(ga.data, gb.data, out)
        torch.cuda.synchronize()
        got = out.cpu().numpy()
        assert rel_fro(got, ref_c) < 1e-6, form
        if form in outs:
            assert np.array_equal(got, outs[form])
        outs[form] = got
    assert np.array_equal(outs["in_place"], outs["panels"])
    plan.set_split_operands("auto")
    plan.autotune(ga.data, gb.data, out, reps=1)
    assert plan.a_layout in (0, 1, 2)
    plan.execute(ga.data, gb.data, out)
    torch.cuda.synchronize()
    assert np.array_equal(out.cpu().numpy(), outs["in_place"])
    rng = np.random.default_rng(4)
    Ap = O.BS(a.dims, a.begin, a.end, permuted_strides(a.begin, a.end, rng))
    pp = torch_tensor(Ap, restride(A, da, Ap.strides), cuda).plan_contract(gb, legs, "bf16x3")
    assert pp.a_layout == 2
    with pytest.raises(TorqueError, match="plain"):
        pp.set_split_operands("in_place")
    pp.set_split_operands("panels")
    with pytest.raises(TorqueError, match="split-BF16"):
        ga.plan_contract(gb, legs).set_split_operands("panels")


def thin_pairs(rng, shapes):
    """Block-diagonal A (sum m x sum k) and B (sum k x sum n) whose i-th
    blocks are m_i x k_i and k_i x n_i."""
    r, k, n = (np.array(x, np.int64) for x in zip(*shapes))

    def mat(rows, cols):
        b0 = np.concatenate([[0], np.cumsum(rows)[:-1]])
        b1 = np.concatenate([[0], np.cumsum(cols)[:-1]])
        begin = np.stack([b0, b1], 1).astype(np.int64)
        end = np.stack([b0 + rows - 1, b1 + cols - 1], 1).astype(np.int64)
        return W.Tensor([rows.sum(), cols.sum()], begin, end,
                        rng.standard_normal(int((rows * cols).sum())))

    return mat(r, k), mat(k, n), [[1, 0]]


@pytest.mark.parametrize("precision", ["native", "fp32", "bf16x3"])
def test_split_k(cuda, precision):
    # few output tiles, long contracted ranges: each pair's K is cut into
    # slices whose partial products are added in slice order; every split
    # matches the oracle, repeats bit for bit, and splits of both tile
    # classes combine with every configuration class
    rng = np.random.default_rng(71)
    shapes = [(48, 3001, 40), (16, 2048, 64), (130, 1500, 140), (64, 33, 64), (8, 700, 5)]
    a, b, legs = thin_pairs(rng, shapes)
    (A, da), (B, db) = from_workload(a), from_workload(b)
    ref_c, ref = O.contract(A, da, B, db, legs)
    tol = {"native": FP64_TOL, "fp32": 1e-5, "bf16x3": 1e-6}[precision]
    ga, gb = torch_tensor(A, da, cuda), torch_tensor(B, db, cuda)
    plan = ga.plan_contract(gb, legs, precision)
    out = torch.empty(plan.out_n_elem, dtype=torch.float64, device=cuda)
    for s0, s1 in [(1, 1), (2, 2), (3, 5), (8, 1), (64, 64)]:
        plan.set_class_split(0, s0)
        plan.set_class_split(1, s1)
        assert plan.split() == [s0, s1]
        runs = []
        for _ in range(2):
            out.fill_(float("nan"))
            plan.execute(ga.data, gb.data, out)
            torch.cuda.synchronize()
            runs.append(out.cpu().numpy())
        assert np.array_equal(runs[0], runs[1]), (s0, s1)
        assert rel_fro(runs[0], ref_c) <= tol, (s0, s1)
    best, _ = plan.autotune(ga.data, gb.data, out, reps=1)
    assert all(s in (1, 2, 4, 8, 16) for s in plan.split())
    plan.execute(ga.data, gb.data, out)
    torch.cuda.synchronize()
    assert rel_fro(out.cpu().numpy(), ref_c) <= tol
    with pytest.raises(TorqueError, match="split"):
        plan.set_class_split(0, 65)


def test_default_split_for_few_tiles(cuda):
    # a matrix times a vector (M = 4096, N = 1, K = 4096) gives one column of
    # tiles: the new plan splits K by default (slots filled, 256 k per slice
    # at least), matches the FP64 product and repeats bit for bit; a plan with
    # many tiles keeps whole pairs; an explicit split survives re-picking
    rng = np.random.default_rng(11)
    m = k = 4096
    xa, xb = rng.standard_normal(m * k), rng.standard_normal(k)
    A = BlockSparseTensor([m, k], xa, np.zeros((1, 2), np.int64), np.array([[m - 1, k - 1]]),
                          device=cuda)
    B = BlockSparseTensor([k, 1], xb, np.zeros((1, 2), np.int64), np.array([[k - 1, 0]]),
                          device=cuda)
    plan = A.plan_contract(B, [[1, 0]])
    split = plan.split()
    assert split[0] > 1 and split[1] == 1, split
    out = torch.empty(plan.out_n_elem, dtype=torch.float64, device=cuda)
    plan.execute(A.data, B.data, out)
    first = out.cpu().numpy()
    want = xa.reshape(k, m).T @ xb
    assert rel_fro(first, want) < FP64_TOL
    plan.execute(A.data, B.data, out)
    assert np.array_equal(out.cpu().numpy(), first)
    plan.set_class_split(0, 2)
    plan.set_small_k(8)  # re-picks the defaults
    assert plan.split()[0] == 2
    a, b, legs = W.c1(0)
    assert BlockSparseTensor(a.dims, a.data, a.begin, a.end, device=cuda).plan_contract(
        BlockSparseTensor(b.dims, b.data, b.begin, b.end, device=cuda), legs).split() == [1, 1]


def test_split_k_cost_model_choice(cuda):
    # the committed cost model predicts split classes from the shapes of
    # their slices; whatever it picks must compute the same contraction
    from torque_amd import costmodel
    model = costmodel.load_default()
    if model is None:
        pytest.skip("no committed cost model")
    rng = np.random.default_rng(72)
    a, b, legs = thin_pairs(rng, [(64, 8192, 64), (32, 4096, 48)])
    (A, da), (B, db) = from_workload(a), from_workload(b)
    ref_c, _ = O.contract(A, da, B, db, legs)
    ga, gb = torch_tensor(A, da, cuda), torch_tensor(B, db, cuda)
    plan = ga.plan_contract(gb, legs)
    model.choose(plan)
    assert plan.split()[0] in (1, 2, 4, 8, 16)
    out = torch.empty(plan.out_n_elem, dtype=torch.float64, device=cuda)
    plan.execute(ga.data, gb.data, out)
    torch.cuda.synchronize()
    assert rel_fro(out.cpu().numpy(), ref_c) <= FP64_TOL


@pytest.mark.parametrize("form", ["panels", "in_place"])
def test_split_domain_flag(cuda, form):
    # values outside the split-BF16 domain (|x| >= 2^127, 0 < |x| < 2^-100,
    # NaN) raise the plan's flag in the pack, the GEMM's in-place loads and
    # the small-K stream; the check clears it; zeros and ordinary values do
    # not raise it; other tiers never do
    rng = np.random.default_rng(81)
    a, b, legs = thin_pairs(rng, [(40, 48, 36), (20, 4, 30)])  # an MFMA and a streamed pair
    (A, da), (B, db) = from_workload(a), from_workload(b)
    ga, gb = torch_tensor(A, da, cuda), torch_tensor(B, db, cuda)
    plan = ga.plan_contract(gb, legs, "bf16x3")
    plan.set_split_operands(form)
    out = torch.empty(plan.out_n_elem, dtype=torch.float64, device=cuda)

    def run(tensor, index, value):
        saved = float(tensor.data[index])
        tensor.data[index] = value
        plan.execute(ga.data, gb.data, out)
        torch.cuda.synchronize()
        tensor.data[index] = saved
        return plan.split_range_check()

    plan.execute(ga.data, gb.data, out)
    assert not plan.split_range_check()
    n_mfma_a = 40 * 48
    for tensor, index in ((ga, 5), (gb, 40), (ga, n_mfma_a + 3)):  # MFMA A, MFMA B, streamed A
        for value in (1e300, -2.0 ** 127, 1e-320, 2.0 ** -101, float("nan"), float("inf")):
            assert run(tensor, index, value), (tensor is ga, index, value)
            assert not plan.split_range_check()  # cleared
        for value in (0.0, 2.0 ** 126, 2.0 ** -99, -1e-20):
            assert not run(tensor, index, value), (index, value)
    p64 = ga.plan_contract(gb, legs)
    ga.data[5] = 1e300
    p64.execute(ga.data, gb.data, out)
    assert not p64.split_range_check()
