"""Shared test helpers: random block-sparse structures and conversions."""
import numpy as np

from oracle import oracle as O
from torque_amd import workloads as W


def random_blocks(rng, dims, n_blocks, max_ext=None, overlap=True):
    """Random (possibly overlapping, unaligned) blocks inside dims."""
    r = len(dims)
    begin = np.zeros((n_blocks, r), np.int64)
    end = np.zeros((n_blocks, r), np.int64)
    for b in range(n_blocks):
        for l, d in enumerate(dims):
            ext = int(rng.integers(1, (max_ext or d) + 1))
            ext = min(ext, d)
            lo = int(rng.integers(0, d - ext + 1))
            begin[b, l], end[b, l] = lo, lo + ext - 1
    return begin, end


def permuted_strides(begin, end, rng):
    """Index tables that store each block in a random leg order (the layout
    a soft_transpose leaves behind), plus the matching data length."""
    bd = end - begin + 1
    tables = np.zeros_like(bd)
    for b in range(bd.shape[0]):
        order = rng.permutation(bd.shape[1])
        acc = 1
        for l in order:
            tables[b, l] = acc
            acc *= bd[b, l]
    return tables


def tensor(dims, begin, end, rng, strides=None):
    bd = end - begin + 1
    n = int(np.prod(bd, axis=1).sum()) if len(bd) else 0
    data = rng.standard_normal(n)
    return O.BS(dims, begin, end, strides), data


def torch_tensor(bs, data, device, dtype=None):
    """torque_amd.BlockSparseTensor from oracle metadata and host data."""
    import torch
    from torque_amd import BlockSparseTensor
    dtype = dtype or torch.float64
    t = BlockSparseTensor(bs.dims, np.asarray(data) if len(data) else np.zeros(1),
                          bs.begin, bs.end, bs.strides, dtype=dtype, device=device)
    return t


def from_workload(t):
    return O.BS(t.dims, t.begin, t.end), t.data


def rel_fro(x, ref):
    x = np.asarray(x, np.float64)
    ref = np.asarray(ref, np.float64)
    den = np.linalg.norm(ref)
    return np.linalg.norm(x - ref) / (den if den > 0 else 1.0)


def dense_of(bs, data):
    return O.densify(bs, data).reshape(tuple(bs.dims), order="F") if bs.dims.size else \
        O.densify(bs, data)


__all__ = ["random_blocks", "permuted_strides", "tensor", "torch_tensor", "from_workload",
           "rel_fro", "dense_of", "W", "O"]
