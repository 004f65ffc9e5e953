"""Writes the golden fixtures of tests/golden/*.npz.

Each fixture holds two block-sparse tensors (metadata + data), the
contracted leg pairs and the dense result computed by numpy.einsum on the
dense images (sums over overlapping blocks), a library routine independent
of the oracle. Run from the repository root: python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from torque_amd import workloads as W  # noqa: E402


def dense(dims, begin, end, strides, data):
    out = np.zeros(tuple(dims), order="F")
    off = 0
    for b, e, s in zip(begin, end, strides):
        ext = e - b + 1
        n = int(np.prod(ext))
        for lin in range(n):
            idx = np.unravel_index(lin, tuple(ext), order="F")
            src = off + int(np.dot(idx, s))
            out[tuple(np.array(idx) + b)] += data[src]
        off += n
    return out


def einsum(da, db, legs):
    ra, rb = da.ndim, db.ndim
    la = list(range(ra))
    lb = list(range(ra, ra + rb))
    for x, y in legs:
        lb[y] = la[x]
    out = [l for l in la if l not in [la[x] for x, _ in legs]] + \
          [l for i, l in enumerate(lb) if i not in [y for _, y in legs]]
    return np.einsum(da, la, db, lb, out)


def canonical(begin, end):
    bd = end - begin + 1
    return np.cumprod(np.concatenate([np.ones((bd.shape[0], 1), np.int64), bd[:, :-1]], 1), 1)


def random_case(rng, dims_a, dims_b, na, nb, legs):
    def blocks(dims, n):
        b = np.zeros((n, len(dims)), np.int64)
        e = np.zeros_like(b)
        for i in range(n):
            for l, d in enumerate(dims):
                ext = int(rng.integers(1, d + 1))
                lo = int(rng.integers(0, d - ext + 1))
                b[i, l], e[i, l] = lo, lo + ext - 1
        return b, e
    ba, ea = blocks(dims_a, na)
    bb, eb = blocks(dims_b, nb)
    sa, sb = canonical(ba, ea), canonical(bb, eb)
    xa = rng.integers(-4, 5, int(np.prod(ea - ba + 1, axis=1).sum())).astype(np.float64)
    xb = rng.integers(-4, 5, int(np.prod(eb - bb + 1, axis=1).sum())).astype(np.float64)
    return (dims_a, ba, ea, sa, xa), (dims_b, bb, eb, sb, xb), legs


def save(name, A, B, legs):
    da = dense(*A)
    db = dense(*B)
    ref = einsum(da, db, legs)
    out_dims = list(ref.shape)
    np.savez_compressed(
        os.path.join(HERE, name + ".npz"),
        a_dims=np.array(A[0], np.int64), a_begin=A[1], a_end=A[2], a_strides=A[3], a_data=A[4],
        b_dims=np.array(B[0], np.int64), b_begin=B[1], b_end=B[2], b_strides=B[3], b_data=B[4],
        legs=np.array(legs, np.int32), out_dims=np.array(out_dims, np.int64),
        dense_out=ref)


def main():
    rng = np.random.default_rng(20240601)
    save("overlap_rank3", *random_case(rng, [4, 5, 3], [3, 5, 2], 4, 3, [[2, 0], [1, 1]]))
    save("overlap_rank4", *random_case(rng, [3, 4, 2, 5], [5, 2, 3], 5, 4, [[3, 0], [0, 2]]))
    a, b, legs = W.c1(1)
    A = (list(a.dims), a.begin, a.end, canonical(a.begin, a.end), np.round(a.data * 4))
    B = (list(b.dims), b.begin, b.end, canonical(b.begin, b.end), np.round(b.data * 4))
    save("u1_c1_int", A, B, legs)


if __name__ == "__main__":
    main()
