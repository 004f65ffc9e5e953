"""Writes tests/golden/sparse/peps_pattern.npz: the element-sparse PEPS
fixture of the reference's GPU test (test/gpu/peps_test.cu:29-768, the same
list as test/gpu/dense_peps_test.cu:25-764): 5,627 sorted linear indices of
a rank-4 tensor of dims {10, 10, 10, 40}, contracted over legs (0,0),(1,1)
with a 10x10 diagonal matrix (indices 0, 11, ..., 99; peps_test.cu:807-808).

The index list is read from the reference's source file (data only: a list
of integers). Values are small integers from a fixed seed, so the expected
result is exact in FP64; it is computed with numpy.einsum on the dense
images, and the expected index set (every output index that receives a
product) from numpy set operations on the coordinates -- both independent
of the oracle. Run from the repository root with the reference at
/root/reference: python tests/golden/sparse/make_sparse_golden.py
"""
import os
import re

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = "/root/reference/test/gpu/peps_test.cu"


def read_indices():
    text = open(SRC).read()
    m = re.search(r"arma::uvec indices = \{([^}]*)\}", text)
    vals = np.array([int(x) for x in m.group(1).replace("\n", " ").split(",") if x.strip()],
                    np.int64)
    return vals


def main():
    idx = read_indices()
    assert idx.size == 5627 and np.all(np.diff(idx) > 0)
    dims_a = np.array([10, 10, 10, 40], np.int64)
    dims_b = np.array([10, 10], np.int64)
    idx_b = np.arange(10, dtype=np.int64) * 11
    rng = np.random.default_rng(5627)
    val_a = rng.integers(-4, 5, idx.size).astype(np.float64)
    val_b = rng.integers(1, 5, idx_b.size).astype(np.float64)
    A = np.zeros(int(np.prod(dims_a)))
    A[idx] = val_a
    B = np.zeros(int(np.prod(dims_b)))
    B[idx_b] = val_b
    dense = np.einsum("abcd,ab->cd", A.reshape(tuple(dims_a), order="F"),
                      B.reshape(tuple(dims_b), order="F"))
    # coordinates of the stored elements; output (c, d) receives a product
    # iff some stored (a, b, c, d) has (a, b) stored in B
    ca = np.stack(np.unravel_index(idx, tuple(dims_a), order="F"), 1)
    cb = np.stack(np.unravel_index(idx_b, tuple(dims_b), order="F"), 1)
    bset = {(int(a), int(b)) for a, b in cb}
    hit = np.array([(int(a), int(b)) in bset for a, b, _, _ in ca])
    out_idx = np.unique(ca[hit, 2] + 10 * ca[hit, 3])
    out_val = dense.reshape(-1, order="F")[out_idx]
    np.savez_compressed(os.path.join(HERE, "peps_pattern.npz"), a_dims=dims_a, a_indices=idx,
                        a_values=val_a, b_dims=dims_b, b_indices=idx_b, b_values=val_b,
                        legs=np.array([[0, 0], [1, 1]], np.int32), out_dims=np.array([10, 40]),
                        out_indices=out_idx, out_values=out_val)
    print("peps_pattern: %d A elements, %d output elements" % (idx.size, out_idx.size))


if __name__ == "__main__":
    main()
