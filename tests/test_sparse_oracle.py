"""Pins the oracle's element-sparse contraction and transposition
(src/tensor/sparse.h:244-405 restated in oracle/oracle.cpp): the
reference's own known answers (test/sparse_tensor_test.cpp:112-258,
test/gpu/sparse_test.cu:145-333), numpy.einsum on dense images (also under
soft-transposed index tables), numpy.transpose for the index re-encoding,
and the PEPS index fixture of test/gpu/peps_test.cu:29-768."""
import os

import numpy as np
import pytest

from helpers import O

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "sparse", "peps_pattern.npz")


def full(n):
    return np.arange(n, dtype=np.int64)


def as_dict(idx, val):
    return {int(i): float(v) for i, v in zip(idx, val)}


def test_known_vector_contraction():
    # sparse_tensor_test.cpp:112-125: {0..4} . {0..4} = 30
    i, v = O.sparse_contract([5], full(5), np.arange(5.0), [5], full(5), np.arange(5.0),
                             [[0, 0]])
    assert i.tolist() == [0] and v.tolist() == [30.0]


def test_known_matrix_multiplication():
    # sparse_tensor_test.cpp:127-151: A(3x3) column-major 0..8, A^2 over (1, 0)
    i, v = O.sparse_contract([3, 3], full(9), np.arange(9.0), [3, 3], full(9),
                             np.arange(9.0), [[1, 0]])
    got = np.zeros(9)
    got[i] = v
    ref = np.array([[15, 42, 69], [18, 54, 90], [21, 66, 111]], np.float64)
    assert np.array_equal(got.reshape(3, 3, order="F"), ref)


def test_known_matrix_inner_product():
    # sparse_tensor_test.cpp:153-175: sum(A % A^T)
    i, v = O.sparse_contract([3, 3], full(9), np.arange(9.0), [3, 3], full(9),
                             np.arange(9.0), [[1, 0], [0, 1]])
    a = np.arange(9.0).reshape(3, 3, order="F")
    assert v.tolist() == [float(np.sum(a * a.T))]


def test_known_tensor_vector_contraction():
    # sparse_tensor_test.cpp:177-221
    t = np.arange(8.0)
    for legs, expect in [([[0, 0]], {(0, 0): 2, (1, 0): 8, (0, 1): 14, (1, 1): 20}),
                         ([[1, 0]], {(0, 0): 4, (1, 0): 7, (0, 1): 16, (1, 1): 19}),
                         ([[2, 0]], {(1, 0): 11})]:
        i, v = O.sparse_contract([2, 2, 2], full(8), t, [2], full(2), np.array([1.0, 2.0]),
                                 legs)
        d = as_dict(i, v)
        for (x, y), val in expect.items():
            assert d[x + 2 * y] == val


def test_known_tensor_matrix_contraction():
    # sparse_tensor_test.cpp:224-258: {19, 59, 99}
    i, v = O.sparse_contract([2, 2, 3], full(12), np.arange(12.0), [2, 2], full(4),
                             np.array([1.0, 2, 3, 4]), [[0, 1], [1, 0]])
    assert i.tolist() == [0, 1, 2] and v.tolist() == [19.0, 59.0, 99.0]


def test_only_receiving_indices_are_listed():
    # a product that sums to zero keeps its index; an index with no product
    # is absent (reduce_by_key over the matching products)
    # A(0,0)=1, A(1,0)=1; B(0,0)=1, B(1,0)=-1 over leg 0: output (0,0)
    # receives 1*1 + 1*(-1)
    i, v = O.sparse_contract([2, 2], np.array([0, 1]), np.array([1.0, 1.0]), [2, 2],
                             np.array([0, 1]), np.array([1.0, -1.0]), [[0, 0]])
    assert i.tolist() == [0]
    assert v.tolist() == [0.0]
    i, v = O.sparse_contract([2, 2], np.array([0]), np.array([3.0]), [2, 2],
                             np.array([3]), np.array([5.0]), [[1, 1]])
    assert i.size == 0


def dense(dims, idx, val):
    d = np.zeros(int(np.prod(dims)))
    np.add.at(d, idx, val)
    return d.reshape(tuple(dims), order="F")


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_contract_vs_einsum_soft_transposed(seed):
    rng = np.random.default_rng(seed)
    dims_a0 = np.array([4, 3, 5], np.int64)
    perm = rng.permutation(3)
    ia = np.sort(rng.choice(60, 25, replace=False))
    va = rng.standard_normal(25)
    # soft_transpose: dimension and table permuted, indices unchanged
    table_a = O.index_table(dims_a0)[perm]
    dims_a = dims_a0[perm]
    ta = np.transpose(dense(dims_a0, ia, va), perm)
    db = np.array([dims_a[2], 2, dims_a[0]], np.int64)
    nb = int(np.prod(db))
    ib = np.sort(rng.choice(nb, nb // 2, replace=False))
    vb = rng.standard_normal(ib.size)
    oi, ov = O.sparse_contract(dims_a, ia, va, db, ib, vb, [[2, 0], [0, 2]], table_a=table_a)
    ref = np.einsum("ijk,kli->jl", ta, dense(db, ib, vb))
    got = np.zeros(ref.size)
    got[oi] = ov
    assert np.allclose(got.reshape(ref.shape, order="F"), ref, rtol=1e-13, atol=1e-14)
    assert np.all(np.diff(oi) > 0)


def test_permute_indices_vs_numpy_transpose():
    rng = np.random.default_rng(3)
    dims = np.array([3, 1, 4, 5], np.int64)
    idx = rng.choice(60, 30, replace=False)
    val = rng.standard_normal(30)
    for perm in ([2, 0, 3, 1], [3, 2, 1, 0], [1, 0, 2, 3]):
        out = O.sparse_permute_indices(dims, idx, perm)
        assert np.array_equal(dense(dims[perm], out, val),
                              np.transpose(dense(dims, idx, val), perm))
    # after a soft transpose the decode follows the permuted table
    # (sparse_tensor_test.cpp:78-98)
    d3 = np.array([2, 2, 3], np.int64)
    t = O.index_table(d3)[[0, 2, 1]]
    out = O.sparse_permute_indices(d3[[0, 2, 1]], full(12), [1, 0, 2], table=t)
    nd = d3[[0, 2, 1]][[1, 0, 2]]
    got = dense(nd, out, np.arange(1.0, 13.0))
    assert got[2, 0, 0] == 9 and got[2, 0, 1] == 11


def test_peps_fixture():
    z = np.load(GOLDEN)
    oi, ov = O.sparse_contract(z["a_dims"], z["a_indices"], z["a_values"], z["b_dims"],
                               z["b_indices"], z["b_values"], z["legs"])
    assert np.array_equal(oi, z["out_indices"])
    assert np.array_equal(ov, z["out_values"])
