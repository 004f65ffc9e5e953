"""GPU element-sparse contraction and transposition (tq_sparse_contract,
tq_sparse_permute_indices) against the oracle: the reference's GPU test
sections (test/gpu/sparse_test.cu:45-333), the PEPS index fixture
(test/gpu/peps_test.cu:29-768), random tensors with soft-transposed index
tables, multi-pass sorts and empty results. Output indices must match
exactly; FP64 values to 1e-12 relative, FP32 data to 1e-5."""
import os

import numpy as np
import pytest
import torch

from helpers import O

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "sparse", "peps_pattern.npz")


def sp(dims, values, indices, dev, table=None, dtype=torch.float64):
    from torque_amd import SparseTensor
    return SparseTensor(dims, np.asarray(values, np.float64), np.asarray(indices, np.int64),
                        table, dtype=dtype, device=dev)


def host(t):
    return t.indices.cpu().numpy(), t.data.cpu().numpy().astype(np.float64)


def rel(x, ref):
    n = np.linalg.norm(ref)
    return np.linalg.norm(x - ref) / (n if n > 0 else 1.0)


def test_reference_sections(cuda):
    from torque_amd import SparseTensor
    # sparse matrix initialization (sparse_test.cu:45-62)
    t = SparseTensor([4, 3], device=cuda)
    assert t.query([1, 1]) == 0
    t.initialize(np.arange(1.0, 13.0), np.arange(12))
    assert t.query([1, 2]) == 10
    # sparse matrix transposition (:64-95)
    t.soft_transpose([1, 0])
    assert t.query([1, 2]) == 7
    tt = t.hard_transpose([1, 0])
    for i in range(3):
        for j in range(4):
            assert t.query([i, j]) == tt.query([j, i])
    # sparse tensor initialization and transposition (:97-134)
    t3 = sp([2, 2, 3], np.arange(1.0, 13.0), np.arange(12), cuda)
    assert t3.query([0, 0, 2]) == 9 and t3.query([0, 1, 2]) == 11
    t3.soft_transpose([0, 2, 1])
    assert t3.query([0, 2, 0]) == 9 and t3.query([0, 2, 1]) == 11
    t4 = t3.hard_transpose([1, 0, 2])
    assert t4.query([2, 0, 0]) == 9 and t4.query([2, 0, 1]) == 11
    # scalar (:136-145)
    s = SparseTensor([], device=cuda)
    assert s.rank == 0 and s.query([]) == 0
    s.modify([], 1)
    assert s.query([]) == 1
    # vector contraction (:147-165)
    v = sp([5], np.arange(5.0), np.arange(5), cuda)
    assert v.contract(v, [[0, 0]]).to_number() == 30
    # matrix multiplication (:167-197)
    m = sp([3, 3], np.arange(9.0), np.arange(9), cuda)
    sq = m.contract(m, [[1, 0]])
    ref = np.array([[15, 42, 69], [18, 54, 90], [21, 66, 111]])
    for i in range(3):
        for j in range(3):
            assert sq.query([i, j]) == ref[i, j]
    # matrix inner product (:199-227)
    a = np.arange(9.0).reshape(3, 3, order="F")
    assert m.contract(m, [[1, 0], [0, 1]]).to_number() == np.sum(a * a.T)
    # tensor-vector contraction (:229-284)
    t8 = sp([2, 2, 2], np.arange(8.0), np.arange(8), cuda)
    vec = sp([2], [1.0, 2.0], [0, 1], cuda)
    c = t8.contract(vec, [[0, 0]])
    assert [c.query([0, 0]), c.query([1, 0]), c.query([0, 1]), c.query([1, 1])] == [2, 8, 14, 20]
    c = t8.contract(vec, [[1, 0]])
    assert [c.query([0, 0]), c.query([1, 0]), c.query([0, 1]), c.query([1, 1])] == [4, 7, 16, 19]
    assert t8.contract(vec, [[2, 0]]).query([1, 0]) == 11
    # tensor-matrix contraction (:286-333)
    t12 = sp([2, 2, 3], np.arange(12.0), np.arange(12), cuda)
    mat = sp([2, 2], [1.0, 2, 3, 4], np.arange(4), cuda)
    c = t12.contract(mat, [[0, 1], [1, 0]])
    assert [c.query([0]), c.query([1]), c.query([2])] == [19, 59, 99]


def test_modify_semantics(cuda):
    t = sp([3, 3], [5.0, 6.0], [0, 4], cuda)
    t.modify([2, 2], 7.0)  # append
    assert t.query([2, 2]) == 7 and t.nnz == 3
    t.modify([0, 0], 0.0)  # remove: the last element moves into the hole
    assert t.query([0, 0]) == 0 and t.nnz == 2
    assert t.indices.cpu().tolist() == [8, 4]
    t.modify([1, 1], 9.0)  # overwrite
    assert t.query([1, 1]) == 9


def test_peps_fixture(cuda):
    z = np.load(GOLDEN)
    a = sp(z["a_dims"], z["a_values"], z["a_indices"], cuda)
    b = sp(z["b_dims"], z["b_values"], z["b_indices"], cuda)
    c = a.contract(b, z["legs"])
    i, v = host(c)
    assert c.dimension.tolist() == z["out_dims"].tolist()
    assert np.array_equal(i, z["out_indices"])
    assert np.array_equal(v, z["out_values"])


@pytest.mark.parametrize("case", range(4))
def test_random_vs_oracle(cuda, case):
    rng = np.random.default_rng(100 + case)
    dims_a0 = np.array([[40, 30, 20], [12, 9, 7], [6, 5, 4], [1000, 7, 1000]][case], np.int64)
    perm = rng.permutation(3) if case in (1, 2) else np.arange(3)
    dims_a = dims_a0[perm]
    table_a = O.index_table(dims_a0)[perm]
    na = [5000, 300, 60, 3000][case]
    ia = rng.choice(int(np.prod(dims_a0)), na, replace=False)  # unsorted
    va = rng.standard_normal(na)
    if case < 3:
        # B contracts A's legs 1 and 2 (in the current, possibly permuted, order)
        dims_b = np.array([dims_a[2], 11, dims_a[1]], np.int64)
        legs = [[1, 2], [2, 0]]
    else:
        # one contracted leg of 7: 10^12 output indices, five radix passes
        dims_b = np.array([1000, 1000, 7], np.int64)
        legs = [[1, 2]]
    nb = min([8000, 400, 100, 3000][case], int(np.prod(dims_b)) // 2)
    ib = rng.choice(int(np.prod(dims_b)), nb, replace=False)
    vb = rng.standard_normal(nb)
    ri, rv = O.sparse_contract(dims_a, ia, va, dims_b, ib, vb, legs, table_a=table_a)
    a = sp(dims_a, va, ia, cuda, table=table_a)
    b = sp(dims_b, vb, ib, cuda)
    gi, gv = host(a.contract(b, legs))
    print("case %d: %d products keys, %d output elements" % (case, ri.size, gi.size))
    assert np.array_equal(gi, ri)
    assert rel(gv, rv) < 1e-12


def test_fp32_data(cuda):
    rng = np.random.default_rng(7)
    ia = rng.choice(600, 200, replace=False)
    ib = rng.choice(300, 150, replace=False)
    va, vb = rng.standard_normal(200), rng.standard_normal(150)
    ri, rv = O.sparse_contract([10, 6, 10], ia, va.astype(np.float32), [10, 30], ib,
                               vb.astype(np.float32), [[2, 0]])
    a = sp([10, 6, 10], va, ia, cuda, dtype=torch.float32)
    b = sp([10, 30], vb, ib, cuda, dtype=torch.float32)
    gi, gv = host(a.contract(b, [[2, 0]]))
    assert np.array_equal(gi, ri)
    assert rel(gv, rv) < 1e-5


def test_empty_and_scalar_results(cuda):
    # no contracted coordinate matches: empty result of the right shape
    a = sp([3, 3], [1.0], [0], cuda)
    b = sp([3, 3], [1.0], [8], cuda)
    c = a.contract(b, [[1, 1]])
    assert c.nnz == 0 and c.dimension.tolist() == [3, 3]
    # full contraction with no match is the scalar 0
    assert a.contract(b, [[0, 0], [1, 1]]).to_number() == 0
    # an operand with no elements
    e = sp([3, 3], np.zeros(0), np.zeros(0, np.int64), cuda)
    assert a.contract(e, [[1, 0]]).nnz == 0
    # full contraction over many products: the FP64 tree sum
    rng = np.random.default_rng(9)
    x = rng.standard_normal(20000)
    t = sp([200, 100], x, np.arange(20000), cuda)
    got = t.contract(t, [[0, 0], [1, 1]]).to_number()
    assert abs(got - np.dot(x, x)) / np.dot(x, x) < 1e-13


def test_hard_transpose_vs_oracle(cuda):
    rng = np.random.default_rng(4)
    dims0 = np.array([5, 1, 7, 6], np.int64)
    perm0 = [3, 0, 2, 1]
    idx = rng.choice(210, 90, replace=False)
    t = sp(dims0, rng.standard_normal(90), idx, cuda)
    t.soft_transpose(perm0)
    for perm in ([1, 0, 2, 3], [3, 2, 1, 0], [2, 3, 0, 1]):
        g = t.hard_transpose(perm)
        ref = O.sparse_permute_indices(t.dimension, idx, perm, table=t.index_table)
        assert np.array_equal(g.indices.cpu().numpy(), ref)
        assert g.index_table.tolist() == O.index_table(t.dimension[perm]).tolist()


def test_dimension_mismatch(cuda):
    from torque_amd import TorqueError
    a = sp([3, 4], [1.0], [0], cuda)
    b = sp([5, 4], [1.0], [0], cuda)
    with pytest.raises(TorqueError, match="do not match"):
        a.contract(b, [[0, 0]])
