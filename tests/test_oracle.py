"""Pins the CPU oracle: the reference's own known answers, closed forms of
the index math, brute-force enumeration of block matching, the U(1) closed
form, and numpy.einsum on dense images (an independent library routine)."""
import itertools

import numpy as np
import pytest

from helpers import O, W, dense_of, permuted_strides, random_blocks, rel_fro, tensor


def bs(dims, blocks):
    """blocks: list of (begin, end, values, strides-or-None)."""
    r = len(dims)
    begin = np.array([b for b, _, _, _ in blocks], np.int64).reshape(-1, r)
    end = np.array([e for _, e, _, _ in blocks], np.int64).reshape(-1, r)
    strides = None
    if any(s is not None for _, _, _, s in blocks):
        rows = []
        for b, e, _, s in blocks:
            ext = np.array(e) - np.array(b) + 1
            rows.append(s if s is not None else O.index_table(ext))
        strides = np.array(rows, np.int64).reshape(-1, r)
    data = np.concatenate([np.asarray(v, np.float64) for _, _, v, _ in blocks]) if blocks else np.zeros(0)
    return O.BS(dims, begin, end, strides), data


def query(t, data, idx):
    d = O.densify(t, data)
    table = O.index_table(t.dims)
    return d[int(np.dot(idx, table))] if len(idx) else d[0]


# ---- a1: index math (src/util/space.h) -----------------------------------

def test_index_table_closed_form():
    dims = [3, 4, 5, 2]
    assert O.index_table(dims).tolist() == [1, 3, 12, 60]


def test_index_to_indices_roundtrip_bruteforce():
    dims = [3, 4, 2]
    t = O.index_table(dims)
    for i, idx in enumerate(itertools.product(*[range(d) for d in reversed(dims)])):
        idx = list(reversed(idx))
        assert O.index_to_indices(i, t).tolist() == idx
    # permuted table (a soft-transposed tensor): decode follows sort order
    tp = t[[2, 0, 1]]
    si = np.argsort(tp)
    for i in range(24):
        ind = O.index_to_indices(i, tp, si)
        assert int(np.dot(ind, tp)) == i


def test_nest_sum_is_exclusive_cumsum():
    v = np.array([3, 0, 5, 7, 1])
    assert O.nest_sum(v).tolist() == [0, 3, 3, 8, 15]
    assert O.nest_sum(v, 10).tolist() == [10, 13, 13, 18, 25]


def test_sort_into_matrix_matches_numpy_transpose():
    rng = np.random.default_rng(1)
    dims = [2, 3, 4, 5]
    block = rng.standard_normal(120)
    arr = block.reshape(dims, order="F")
    m = O.sort_into_matrix(block, dims, O.index_table(dims), [3, 1])
    ref = np.transpose(arr, (0, 2, 3, 1)).reshape(2 * 4, 5 * 3, order="F")
    assert np.array_equal(m.reshape(8, 15, order="F"), ref)


# ---- known answers of test/block_sparse_tensor_test.cpp -----------------

def test_known_vector_contraction():
    # test/block_sparse_tensor_test.cpp:104-116
    t, d = bs([5], [([0], [1], [0, 1], None), ([2], [4], [2, 3, 4], None)])
    c, _ = O.contract(t, d, t, d, [[0, 0]])
    assert c[0] == 30


def test_known_matrix_multiplication():
    # test/block_sparse_tensor_test.cpp:118-141
    vec = np.array([0, 1, 0, 0, 2, 3, 0, 0, 0, 0, 4, 5, 0, 0, 6, 7], np.float64)
    t, d = bs([4, 4], [([0, 0], [1, 1], [0, 1, 2, 3], None),
                       ([2, 2], [3, 3], [4, 5, 6, 7], None)])
    c, p = O.contract(t, d, t, d, [[1, 0]])
    out = O.BS([4, 4], p["out_begin"], p["out_end"])
    dense_ref, _ = O.dense_contract([4, 4], vec, [4, 4], vec, [[1, 0]])
    assert np.array_equal(O.densify(out, c), dense_ref)


def test_known_matrix_inner_product():
    # test/block_sparse_tensor_test.cpp:143-165: overlapping blocks
    t, d = bs([3, 3], [([0, 0], [1, 1], [0, 1, 2, 3], None),
                       ([1, 1], [2, 2], [0, 7, 8, 9], None)])
    c, _ = O.contract(t, d, t, d, [[1, 0], [0, 1]])
    ref_a = np.array([[0, 1, 0], [2, 3, 7], [0, 8, 9]], np.float64).T
    assert c[0] == np.sum(ref_a * ref_a.T)


def test_known_tensor_vector_contraction():
    # test/block_sparse_tensor_test.cpp:167-204
    t, d = bs([2, 2, 2], [([0, 0, 0], [1, 1, 0], [0, 1, 2, 3], None),
                          ([0, 0, 1], [1, 1, 1], [4, 5, 6, 7], None)])
    v, dv = bs([2], [([0], [1], [1, 2], None)])
    c, p = O.contract(t, d, v, dv, [[0, 0]])
    out = O.BS([2, 2], p["out_begin"], p["out_end"])
    assert [query(out, c, i) for i in ([0, 0], [1, 0], [0, 1], [1, 1])] == [2, 8, 14, 20]
    c, p = O.contract(t, d, v, dv, [[1, 0]])
    out = O.BS([2, 2], p["out_begin"], p["out_end"])
    assert [query(out, c, i) for i in ([0, 0], [1, 0], [0, 1], [1, 1])] == [4, 7, 16, 19]
    c, p = O.contract(t, d, v, dv, [[2, 0]])
    out = O.BS([2, 2], p["out_begin"], p["out_end"])
    assert query(out, c, [1, 0]) == 11


def test_known_tensor_matrix_contraction():
    # test/block_sparse_tensor_test.cpp:206-231
    t, d = bs([2, 2, 3], [([0, 0, 2], [1, 1, 2], [8, 9, 10, 11], None),
                          ([0, 0, 1], [1, 1, 1], [4, 5, 6, 7], None),
                          ([0, 0, 0], [1, 1, 0], [0, 1, 2, 3], None)])
    m, dm = bs([2, 2], [([0, 0], [1, 1], [1, 2, 3, 4], None)])
    c, p = O.contract(t, d, m, dm, [[0, 1], [1, 0]])
    out = O.BS([3], p["out_begin"], p["out_end"])
    assert [query(out, c, [i]) for i in range(3)] == [19, 59, 99]


def test_known_transpositions():
    # test/block_sparse_tensor_test.cpp:35-60 (hard_transpose back)
    t, d = bs([4, 3], [([2, 0], [3, 1], [1, 2, 3, 4], None)])
    out = O.bs_permute(t, d, [1, 0])
    tt = O.BS([3, 4], t.begin[:, [1, 0]], t.end[:, [1, 0]])
    assert query(tt, out, [0, 3]) == 2 and query(tt, out, [1, 2]) == 3
    # test/block_sparse_tensor_test.cpp:76-91 (unit leg, table {0,1,2})
    t, d = bs([2, 2, 3], [([0, 0, 0], [0, 1, 2], [1, 2, 3, 4, 5, 6], [0, 1, 2]),
                          ([1, 0, 0], [1, 1, 2], [7, 8, 9, 10, 11, 12], [0, 1, 2])])
    out = O.bs_permute(t, d, [1, 0, 2])
    tt = O.BS([2, 2, 3], t.begin[:, [1, 0, 2]], t.end[:, [1, 0, 2]])
    assert query(tt, out, [1, 0, 0]) == 2
    assert query(tt, out, [1, 0, 1]) == 4
    assert query(tt, out, [1, 1, 2]) == 12
    # test/gpu/block_sparse_test.cu:76-91: overlapping blocks of different shapes
    t, d = bs([2, 2, 3], [([0, 0, 0], [1, 1, 2], [1, 2, 3, 4, 5, 6] + [0] * 6, [1, 2, 4]),
                          ([1, 0, 0], [1, 1, 2], [7, 8, 9, 10, 11, 12], [1, 1, 2])])
    out = O.bs_permute(t, d, [1, 0, 2])
    tt = O.BS([2, 2, 3], t.begin[:, [1, 0, 2]], t.end[:, [1, 0, 2]])
    assert [query(tt, out, i) for i in ([1, 0, 0], [1, 0, 1], [1, 1, 2])] == [3, 0, 12]


def test_known_dense():
    # test/dense_tensor_test.cpp / test/gpu/dense_test.cu:78-94
    a = np.arange(1, 13, dtype=np.float64)
    t1 = O.dense_permute([2, 2, 3], a, [0, 2, 1])
    assert t1.reshape((2, 3, 2), order="F")[0, 2, 0] == 9
    assert t1.reshape((2, 3, 2), order="F")[0, 2, 1] == 11
    t2 = O.dense_permute([2, 3, 2], t1, [1, 0, 2])
    assert t2.reshape((3, 2, 2), order="F")[2, 0, 0] == 9
    assert t2.reshape((3, 2, 2), order="F")[2, 0, 1] == 11
    m = O.dense_permute([4, 3], a, [1, 0])
    assert m.reshape((3, 4), order="F")[1, 2] == 7


# ---- a2/a3: matching and index maps -------------------------------------

def brute_pairs(a, b, legs):
    out = []
    for i in range(a.begin.shape[0]):
        for j in range(b.begin.shape[0]):
            ok = all(max(a.begin[i, x], b.begin[j, y]) <= min(a.end[i, x], b.end[j, y])
                     for x, y in legs)
            if ok:
                out.append((i, j))
    return out


@pytest.mark.parametrize("seed", range(6))
def test_matching_bruteforce(seed):
    rng = np.random.default_rng(seed)
    dims_a = [5, 6, 4]
    dims_b = [6, 3, 5]
    ba, ea = random_blocks(rng, dims_a, 7)
    bb, eb = random_blocks(rng, dims_b, 9)
    a, _ = tensor(dims_a, ba, ea, rng)
    b, _ = tensor(dims_b, bb, eb, rng)
    legs = [[1, 0], [0, 2]]
    p = O.plan(a, b, legs)
    assert list(zip(p["a_block"], p["b_block"])) == brute_pairs(a, b, legs)
    # invariants of the output maps
    assert np.all(np.diff(p["out_offsets"]) == (p["m"] * p["n"])[:-1])
    assert p["out_begin"].shape == (len(p["m"]), 2)
    for q, (i, j) in enumerate(zip(p["a_block"], p["b_block"])):
        assert p["out_begin"][q].tolist() == [a.begin[i, 2], b.begin[j, 1]]
        assert p["out_end"][q].tolist() == [a.end[i, 2], b.end[j, 1]]


def test_u1_closed_form_c1():
    a, b, legs = W.c1(0)
    A = O.BS(a.dims, a.begin, a.end)
    B = O.BS(b.dims, b.begin, b.end)
    p = O.plan(A, B, legs)
    # match <=> equal sector on both contracted legs (disjoint sectors)
    q = {0: -1, 4: 0, 12: 1}
    for i, j in zip(p["a_block"], p["b_block"]):
        assert q[a.begin[i, 2]] == q[b.begin[j, 0]]
        assert q[a.begin[i, 3]] == q[b.begin[j, 1]]
    assert (a.n_blocks, b.n_blocks, len(p["m"])) == (19, 19, 45)
    assert int((2 * p["m"] * p["n"] * p["k"]).sum()) == 2834432
    assert int((p["m"] * p["n"]).sum()) == 44544


def test_peps_pair_count():
    E, A, Ac, l1, l2 = W.peps(8, 128)
    TA = O.BS(A.dims, A.begin, A.end)
    p1 = O.plan(TA, TA, l1)
    assert (A.n_blocks, E.n_blocks, len(p1["m"])) == (160, 85, 12800)
    ab = np.array([np.concatenate([A.begin[i, 1:], A.begin[j, 1:]]) for i, j in
                   zip(p1["a_block"], p1["b_block"])])
    ae = np.array([np.concatenate([A.end[i, 1:], A.end[j, 1:]]) for i, j in
                   zip(p1["a_block"], p1["b_block"])])
    assert np.array_equal(ab, p1["out_begin"]) and np.array_equal(ae, p1["out_end"])
    a2 = O.BS(np.concatenate([A.dims[1:], A.dims[1:]]), ab, ae)
    TE = O.BS(E.dims, E.begin, E.end)
    assert O.plan(TE, a2, l2)["m"].size == 44872


# ---- a5/a12/a13: arithmetic against numpy.einsum --------------------------

def einsum_ref(da, db, legs):
    ra, rb = da.ndim, db.ndim
    la = list(range(ra))
    lb = list(range(ra, ra + rb))
    for x, y in legs:
        lb[y] = la[x]
    out = [l for l in la if l not in [la[x] for x, _ in legs]] + \
          [l for i, l in enumerate(lb) if i not in [y for _, y in legs]]
    return np.einsum(da, la, db, lb, out)


@pytest.mark.parametrize("seed", range(5))
def test_contract_vs_einsum_random_overlapping(seed):
    rng = np.random.default_rng(100 + seed)
    dims_a = [4, 5, 3, 6]
    dims_b = [3, 6, 4]
    ba, ea = random_blocks(rng, dims_a, 6)
    bb, eb = random_blocks(rng, dims_b, 5)
    a, da = tensor(dims_a, ba, ea, rng, permuted_strides(ba, ea, rng))
    b, db = tensor(dims_b, bb, eb, rng)
    legs = [[3, 1], [2, 0]]
    c, p = O.contract(a, da, b, db, legs)
    out = O.BS([4, 5, 4], p["out_begin"], p["out_end"])
    got = dense_of(out, c)
    ref = einsum_ref(dense_of(a, da), dense_of(b, db), legs)
    assert rel_fro(got, ref) < 1e-13


def test_contract_rank0_and_outer():
    rng = np.random.default_rng(7)
    ba, ea = random_blocks(rng, [4, 5], 4)
    a, da = tensor([4, 5], ba, ea, rng)
    c, _ = O.contract(a, da, a, da, [[0, 0], [1, 1]])
    d = dense_of(a, da)
    assert abs(c[0] - np.sum(d * d)) < 1e-12 * np.sum(d * d)
    c, p = O.contract(a, da, a, da, np.zeros((0, 2)))
    out = O.BS([4, 5, 4, 5], p["out_begin"], p["out_end"])
    assert rel_fro(dense_of(out, c), np.einsum("ab,cd->abcd", d, d)) < 1e-14


def test_dense_contract_and_permute_vs_numpy():
    rng = np.random.default_rng(3)
    a = rng.standard_normal(3 * 4 * 5)
    b = rng.standard_normal(5 * 3 * 2)
    out, od = O.dense_contract([3, 4, 5], a, [5, 3, 2], b, [[2, 0], [0, 1]])
    ref = np.einsum("ijk,kil->jl", a.reshape(3, 4, 5, order="F"), b.reshape(5, 3, 2, order="F"))
    assert od == [4, 2] and np.allclose(out.reshape(4, 2, order="F"), ref, rtol=1e-13)
    x = rng.standard_normal(2 * 3 * 4 * 5)
    perm = [2, 0, 3, 1]
    got = O.dense_permute([2, 3, 4, 5], x, perm)
    ref = np.transpose(x.reshape(2, 3, 4, 5, order="F"), perm)
    assert np.array_equal(got.reshape(ref.shape, order="F"), ref)


def test_bs_permute_vs_dense_transpose():
    rng = np.random.default_rng(11)
    dims = [3, 4, 5]
    ba, ea = random_blocks(rng, dims, 5)
    t, d = tensor(dims, ba, ea, rng, permuted_strides(ba, ea, rng))
    perm = [2, 0, 1]
    out = O.bs_permute(t, d, perm)
    tt = O.BS(np.array(dims)[perm], t.begin[:, perm], t.end[:, perm])
    assert np.array_equal(dense_of(tt, out), np.transpose(dense_of(t, d), perm))


def test_sparse_contract_vs_einsum():
    rng = np.random.default_rng(5)
    da, db = [4, 5, 3], [5, 3, 2]
    ia = np.sort(rng.choice(60, 20, replace=False))
    ib = np.sort(rng.choice(30, 12, replace=False))
    va, vb = rng.standard_normal(20), rng.standard_normal(12)
    oi, ov = O.sparse_contract(da, ia, va, db, ib, vb, [[1, 0], [2, 1]])
    A = np.zeros(60)
    A[ia] = va
    B = np.zeros(30)
    B[ib] = vb
    ref = np.einsum("ijk,jkl->il", A.reshape(da, order="F"), B.reshape(db, order="F"))
    got = np.zeros(8)
    got[oi] = ov
    assert np.allclose(got.reshape(4, 2, order="F"), ref, rtol=1e-13, atol=1e-15)
    assert np.all(np.diff(oi) > 0)


def test_dimension_mismatch_is_reported():
    t, d = bs([5], [([0], [1], [0, 1], None)])
    u, e = bs([4], [([0], [1], [0, 1], None)])
    with pytest.raises(ValueError, match="do not match"):
        O.contract(t, d, u, e, [[0, 0]])


# ---- golden fixtures (numpy.einsum vectors committed under tests/golden) --

def test_golden_fixtures():
    import glob
    import os
    files = sorted(glob.glob(os.path.join(os.path.dirname(__file__), "golden", "*.npz")))
    assert files, "golden fixtures missing"
    for f in files:
        z = np.load(f)
        a = O.BS(z["a_dims"], z["a_begin"], z["a_end"], z["a_strides"])
        b = O.BS(z["b_dims"], z["b_begin"], z["b_end"], z["b_strides"])
        c, p = O.contract(a, z["a_data"], b, z["b_data"], z["legs"])
        out = O.BS(z["out_dims"], p["out_begin"], p["out_end"])
        assert rel_fro(dense_of(out, c), z["dense_out"]) < 1e-13, f
