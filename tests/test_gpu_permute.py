"""Permutation kernels against the oracle: bit-exact (pure copies)."""
import numpy as np
import pytest
import torch

from helpers import O, W, permuted_strides, random_blocks, tensor, torch_tensor
from torque_amd import BlockSparseTensor, DenseTensor

pytestmark = pytest.mark.gpu


def test_known_block_sparse_transpositions(cuda):
    # test/gpu/block_sparse_test.cu:35-60
    t = BlockSparseTensor([4, 3], [1, 2, 3, 4], [[2, 0]], [[3, 1]], dtype=torch.float32, device=cuda)
    t = t.hard_transpose([1, 0])
    assert t.query([0, 3]) == 2 and t.query([1, 2]) == 3
    back = t.hard_transpose([1, 0])
    assert back.query([3, 0]) == 2 and back.query([2, 1]) == 3
    # test/gpu/block_sparse_test.cu:76-91: overlapping blocks, different shapes
    t = BlockSparseTensor([2, 2, 3], dtype=torch.float32, device=cuda)
    t.append_block([1, 2, 3, 4, 5, 6, 0, 0, 0, 0, 0, 0], [0, 0, 0], [1, 1, 2], [1, 2, 4])
    t.append_block([7, 8, 9, 10, 11, 12], [1, 0, 0], [1, 1, 2], [1, 1, 2])
    tt = t.hard_transpose([1, 0, 2])
    assert [tt.query(i) for i in ([1, 0, 0], [1, 0, 1], [1, 1, 2])] == [3, 0, 12]


def test_known_dense_transpositions(cuda):
    # test/gpu/dense_test.cu:27-94
    d = DenseTensor([4, 3], np.arange(1, 13), dtype=torch.float32, device=cuda)
    assert d.query([1, 2]) == 10
    d = d.hard_transpose([1, 0])
    assert d.query([1, 2]) == 7
    back = d.hard_transpose([1, 0])
    for i in range(3):
        for j in range(4):
            assert d.query([i, j]) == back.query([j, i])
    t = DenseTensor([2, 2, 3], np.arange(1, 13), device=cuda)
    t = t.hard_transpose([0, 2, 1])
    assert t.query([0, 2, 0]) == 9 and t.query([0, 2, 1]) == 11
    t2 = t.hard_transpose([1, 0, 2])
    assert t2.query([2, 0, 0]) == 9 and t2.query([2, 0, 1]) == 11


CASES = [
    ([7], [0]),
    ([5, 9], [1, 0]),
    ([3, 4, 5], [2, 0, 1]),
    ([64, 80], [1, 0]),                 # tiled, square-ish
    ([10, 20, 25, 8], [3, 2, 1, 0]),    # tiled, small fastest leg
    ([33, 2, 65], [2, 1, 0]),
    ([6, 5, 4, 3, 2, 7], [5, 4, 3, 2, 1, 0]),
    ([6, 5, 4, 3, 2, 7], [0, 2, 1, 4, 3, 5]),  # fastest leg kept
    ([16, 1, 24, 1, 9], [2, 1, 4, 0, 3]),      # unit legs
    ([128, 3, 130], [2, 1, 0]),
]


@pytest.mark.parametrize("dims,perm", CASES)
@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_dense_permute_bit_exact(cuda, dims, perm, dtype):
    rng = np.random.default_rng(len(dims) * 31 + sum(dims))
    x = rng.standard_normal(int(np.prod(dims)))
    d = DenseTensor(dims, x, dtype=dtype, device=cuda)
    got = d.hard_transpose(perm).data.cpu().numpy()
    ref = O.dense_permute(dims, d.data.cpu().double().numpy(), perm)
    assert np.array_equal(got.astype(np.float64), ref)


def test_dense_permute_after_soft_transpose(cuda):
    rng = np.random.default_rng(4)
    dims = [6, 7, 8]
    x = rng.standard_normal(336)
    d = DenseTensor(dims, x, device=cuda)
    d.soft_transpose([2, 0, 1])
    got = d.hard_transpose([1, 2, 0]).data.cpu().numpy()
    # element (i0,i1,i2) of the soft view is x at sum i_l * table[l]
    table = np.array([42, 1, 6])
    sview = np.zeros((8, 6, 7))
    for a in range(8):
        for b in range(6):
            for c in range(7):
                sview[a, b, c] = x[a * table[0] + b * table[1] + c * table[2]]
    ref = np.transpose(sview, [1, 2, 0])
    assert np.array_equal(got.reshape(ref.shape, order="F"), ref)


@pytest.mark.parametrize("seed", range(4))
def test_block_sparse_permute_bit_exact(cuda, seed):
    rng = np.random.default_rng(seed)
    dims = [5, 7, 4, 6][: 2 + seed % 3]
    ba, ea = random_blocks(rng, dims, 9)
    t, d = tensor(dims, ba, ea, rng, permuted_strides(ba, ea, rng))
    perm = list(rng.permutation(len(dims)))
    gt = torch_tensor(t, d, cuda)
    out = gt.hard_transpose(perm)
    ref = O.bs_permute(t, d, perm)
    assert np.array_equal(out.data.cpu().numpy()[: ref.size], ref)
    assert np.array_equal(out.begin_points, t.begin[:, perm])
    assert np.array_equal(out.block_offsets, O.nest_sum(t.n_elem))


def test_block_sparse_permute_chunk_edges(cuda):
    # blocks whose sizes sit on and around the gather's 512-element work items
    # (one wavefront each), mixed with one-element blocks, so that items end
    # inside, at and just past block ends; rank 3 with permuted storage
    rng = np.random.default_rng(5)
    shapes = [(1, 1, 511), (8, 64, 1), (3, 9, 19), (1, 1, 1), (31, 33, 1), (2, 16, 32),
              (5, 5, 41), (1, 1, 1), (64, 32, 1), (7, 11, 27), (1, 2, 1), (9, 7, 48)]
    shapes = np.array(shapes, np.int64)
    b0 = np.cumsum(np.vstack([np.zeros((1, 3), np.int64), shapes[:-1]]), axis=0)
    begin, end = b0, b0 + shapes - 1
    dims = list(end.max(axis=0) + 1)
    t, d = tensor(dims, begin, end, rng, permuted_strides(begin, end, rng))
    for perm in ([2, 0, 1], [1, 2, 0], [0, 2, 1]):
        out = torch_tensor(t, d, cuda).hard_transpose(perm)
        ref = O.bs_permute(t, d, perm)
        assert np.array_equal(out.data.cpu().numpy()[: ref.size], ref), perm


def test_block_sparse_permute_large_blocks(cuda):
    from torque_amd import workloads as W
    E, A, Ac, _, _ = W.peps(8, 128)
    t = O.BS(E.dims, E.begin, E.end)
    gt = torch_tensor(t, E.data, cuda)
    out = gt.hard_transpose([0, 3, 1, 2])
    ref = O.bs_permute(t, E.data, [0, 3, 1, 2])
    assert np.array_equal(out.data.cpu().numpy()[: ref.size], ref)


# ---- slice (src/gpu/block_sparse.cuh:1124-1176) ---------------------------

def expected_sub_blocks(begin, end, divisor):
    """Sub-block begin/end points in the order of the reference's loop:
    block-major, sub-block j decoded leg 0 fastest over the divisor."""
    divisor = np.asarray(divisor)
    n_sub = int(np.prod(divisor))
    bs, es = [], []
    for b, e in zip(begin, end):
        sub = (e - b + 1) // divisor
        for j in range(n_sub):
            q = np.array(np.unravel_index(j, tuple(divisor), order="F"))
            bs.append(b + q * sub)
            es.append(b + q * sub + sub - 1)
    return np.array(bs, np.int64), np.array(es, np.int64)


@pytest.mark.parametrize("seed", [0, 1])
def test_slice_keeps_values(cuda, seed):
    rng = np.random.default_rng(40 + seed)
    dims = [12, 8, 6]
    # block extents divisible by the divisor (2, 2, 3); overlapping blocks and
    # permuted index tables on the second seed
    n = 5
    ext = np.array([[2, 2, 3], [4, 2, 6], [6, 4, 3], [2, 8, 6], [12, 4, 3]], np.int64)[:n]
    begin = np.array([[int(rng.integers(0, d - x + 1)) for d, x in zip(dims, row)]
                      for row in ext], np.int64)
    end = begin + ext - 1
    strides = permuted_strides(begin, end, rng) if seed else None
    t, d = tensor(dims, begin, end, rng, strides)
    g = torch_tensor(t, d, cuda)
    s = g.slice([2, 2, 3])
    eb, ee = expected_sub_blocks(begin, end, [2, 2, 3])
    assert np.array_equal(s.begin_points, eb) and np.array_equal(s.end_points, ee)
    # every sub-block holds the values of its region: same dense image, and
    # with compact tables each sub-block equals the parent's region
    so = O.BS(dims, s.begin_points, s.end_points)
    got = s.data.cpu().numpy()
    assert np.array_equal(O.densify(so, got), O.densify(t, d))
    parent = O.densify(O.BS(dims, begin[:1], end[:1], None if strides is None else strides[:1]),
                       d[:int(np.prod(ext[0]))]).reshape(dims, order="F")
    for j in range(12):
        b, e = eb[j], ee[j]
        region = parent[b[0]:e[0] + 1, b[1]:e[1] + 1, b[2]:e[2] + 1]
        off = int(s.block_offsets[j])
        assert np.array_equal(got[off:off + region.size], region.reshape(-1, order="F"))


def test_slice_then_contract(cuda):
    # block_sparse_benchmark.cu:73-83: contracting sliced tensors gives the
    # same dense result
    a, b, legs = W.c1(3)
    ga = torch_tensor(O.BS(a.dims, a.begin, a.end), a.data, cuda)
    gb = torch_tensor(O.BS(b.dims, b.begin, b.end), b.data, cuda)
    ref = ga.contract(gb, legs)
    sa, sb = ga.slice([2, 2, 2, 2]), gb.slice([2, 2, 2, 2])
    got = sa.contract(sb, legs)
    dref = O.densify(O.BS(ref.dimension, ref.begin_points, ref.end_points),
                     ref.data.cpu().numpy())
    dgot = O.densify(O.BS(got.dimension, got.begin_points, got.end_points),
                     got.data.cpu().numpy())
    assert np.abs(dgot - dref).max() <= 1e-12 * np.abs(dref).max()


def test_slice_residue_error(cuda):
    from torque_amd import TorqueError
    t = BlockSparseTensor([6, 6], np.arange(9.0), [[0, 0]], [[2, 2]], device=cuda)
    with pytest.raises(TorqueError, match="residues"):
        t.slice([2, 1])
