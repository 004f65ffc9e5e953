"""Charge-aware front end for U(1)-symmetric block-sparse tensors.

A leg carries its U(1) charge sectors (charge, size) in a fixed order and a
direction (+1 or -1); a tensor holds one block per sector tuple whose signed
charges sum to the tensor's flux, and its blocks are addressed by their charge
tuples. Contractions check that each contracted pair of legs carries the same
sectors with opposite directions and run the merged-output device plan
(tq_bs_contract_plan_merged), so the result again has one block per output
charge tuple: the sum over the contracted charges of the matching products.
The blocks are a plain BlockSparseTensor (column-major blocks, leg 0
fastest), which every other operation of the package takes as is.

The reference's tensors take their blocks as given begin and end points
(src/gpu/block_sparse.cuh:190-283) and keep no charge information; this
module derives the blocks from the charges and addresses them by charge.
"""
import itertools

import numpy as np
import torch

from . import _lib
from ._lib import TorqueError
from .tensor import BlockSparseTensor


class U1Leg:
    """Charge sectors (charge, size) of one leg, in the order they occupy
    the leg's index range, and the leg's direction (+1 or -1)."""

    def __init__(self, charges, sizes, sign=1):
        charges = [int(q) for q in charges]
        sizes = [int(n) for n in sizes]
        if len(charges) != len(sizes) or not charges:
            raise TorqueError(_lib.TQ_ERROR_INVALID_VALUE,
                              "a leg needs one size per charge sector")
        if len(set(charges)) != len(charges):
            raise TorqueError(_lib.TQ_ERROR_INVALID_VALUE, "leg charges must be distinct")
        if any(n < 1 for n in sizes):
            raise TorqueError(_lib.TQ_ERROR_INVALID_VALUE, "sector sizes must be positive")
        if sign not in (1, -1):
            raise TorqueError(_lib.TQ_ERROR_INVALID_VALUE, "a leg's sign is +1 or -1")
        self.charges = tuple(charges)
        self.sizes = tuple(sizes)
        self.sign = int(sign)
        self.begins = tuple(int(b) for b in np.concatenate([[0], np.cumsum(sizes)[:-1]]))
        self.dim = int(sum(sizes))
        self._at = {b: i for i, b in enumerate(self.begins)}

    def dual(self):
        """The same sectors with the opposite direction (the leg it contracts with)."""
        return U1Leg(self.charges, self.sizes, -self.sign)

    def same_sectors(self, other):
        return self.charges == other.charges and self.sizes == other.sizes

    def charge_at(self, begin):
        """Charge of the sector starting at index `begin`."""
        return self.charges[self._at[int(begin)]]

    def __eq__(self, other):
        return isinstance(other, U1Leg) and self.same_sectors(other) and self.sign == other.sign

    def __repr__(self):
        return "U1Leg(%s, %s, sign=%+d)" % (list(self.charges), list(self.sizes), self.sign)


def structure(legs, flux=0):
    """Charge tuples and begin/end points of every charge-conserving block
    (sum over legs of sign * charge == flux), leg 0's sector fastest."""
    keys, begin, end = [], [], []
    ranges = [range(len(l.charges)) for l in legs]
    # itertools.product varies the last factor fastest: reverse for leg 0
    for rev in itertools.product(*ranges[::-1]):
        idx = rev[::-1]
        q = sum(l.sign * l.charges[i] for l, i in zip(legs, idx))
        if q != flux:
            continue
        keys.append(tuple(l.charges[i] for l, i in zip(legs, idx)))
        begin.append([l.begins[i] for l, i in zip(legs, idx)])
        end.append([l.begins[i] + l.sizes[i] - 1 for l, i in zip(legs, idx)])
    r = len(legs)
    return (keys, np.asarray(begin, np.int64).reshape(-1, r),
            np.asarray(end, np.int64).reshape(-1, r))


class U1Tensor:
    """A U(1)-symmetric tensor: legs (U1Leg), flux, and the blocks of every
    charge-conserving sector tuple, keyed by charge tuple."""

    def __init__(self, legs, flux=0, data=None, dtype=torch.float64, device=None):
        self.legs = list(legs)
        self.flux = int(flux)
        keys, begin, end = structure(self.legs, self.flux)
        n = int(np.prod(end - begin + 1, axis=1).sum()) if len(keys) else 0
        if data is None:
            data = torch.zeros(max(n, 1), dtype=dtype,
                               device=device if device is not None else "cuda")
        self.bs = BlockSparseTensor([l.dim for l in self.legs], data, begin, end, dtype=dtype,
                                    device=device)
        self._set_keys(keys)

    @classmethod
    def random(cls, legs, flux=0, generator=None, dtype=torch.float64, device=None):
        """Blocks of N(0, 1) entries (drawn on the tensor's device)."""
        keys, begin, end = structure(list(legs), int(flux))
        n = int(np.prod(end - begin + 1, axis=1).sum()) if len(keys) else 0
        dev = torch.device(device) if device is not None else torch.device("cuda")
        data = torch.randn(max(n, 1), dtype=dtype, device=dev, generator=generator)
        return cls(legs, flux, data, dtype=dtype, device=dev)

    @classmethod
    def _wrap(cls, bs, legs, flux):
        t = cls.__new__(cls)
        t.legs = list(legs)
        t.flux = int(flux)
        t.bs = bs
        keys = [tuple(l.charge_at(b) for l, b in zip(t.legs, row)) for row in bs.begin_points]
        t._set_keys(keys)
        return t

    def _set_keys(self, keys):
        self.keys = list(keys)
        self._index = {k: i for i, k in enumerate(self.keys)}
        if len(self._index) != len(self.keys):
            raise TorqueError(_lib.TQ_ERROR_INVALID_VALUE, "two blocks share a charge tuple")

    @property
    def rank(self):
        return len(self.legs)

    @property
    def dims(self):
        return [l.dim for l in self.legs]

    @property
    def n_blocks(self):
        return len(self.keys)

    def block(self, key):
        """The block of charge tuple `key` as a view of the tensor's data,
        indexed [i0, i1, ...] (stored column-major, leg 0 fastest)."""
        i = self._index.get(tuple(int(q) for q in key))
        if i is None:
            raise KeyError(key)
        shape = [int(x) for x in self.bs.blocks_dimension[i]]
        off, n = int(self.bs.block_offsets[i]), int(self.bs.block_n_elem[i])
        flat = self.bs.data[off:off + n]
        return flat.view(shape[::-1]).permute(*range(len(shape) - 1, -1, -1))

    def contract(self, other, pairs, precision="native"):
        """Contracts legs pairs[i][0] of this tensor with legs pairs[i][1] of
        `other`. Each pair must carry the same sectors with opposite signs.
        The result's legs are this tensor's free legs then other's, its flux
        the sum of the fluxes, one block per output charge tuple."""
        pairs = [(int(a), int(b)) for a, b in pairs]
        for a, b in pairs:
            if not (0 <= a < self.rank and 0 <= b < other.rank):
                raise TorqueError(_lib.TQ_ERROR_INVALID_VALUE, "contracted leg out of range")
            la, lb = self.legs[a], other.legs[b]
            if not la.same_sectors(lb):
                raise TorqueError(_lib.TQ_ERROR_INVALID_VALUE,
                                  "contracted legs %d and %d carry different sectors" % (a, b))
            if la.sign != -lb.sign:
                raise TorqueError(_lib.TQ_ERROR_INVALID_VALUE,
                                  "contracted legs %d and %d must point in opposite "
                                  "directions" % (a, b))
        out = self.bs.contract(other.bs, pairs, precision=precision, merged=True)
        ca = {a for a, _ in pairs}
        cb = {b for _, b in pairs}
        legs = ([l for i, l in enumerate(self.legs) if i not in ca] +
                [l for j, l in enumerate(other.legs) if j not in cb])
        return U1Tensor._wrap(out, legs, self.flux + other.flux)

    def transpose(self, permutation):
        """Leg i of the result is leg permutation[i] (a hard transpose of
        every block)."""
        perm = [int(p) for p in permutation]
        out = self.bs.hard_transpose(perm)
        return U1Tensor._wrap(out, [self.legs[p] for p in perm], self.flux)

    def norm(self):
        """Frobenius norm (FP64 sum on the GPU)."""
        n = int(self.bs.block_n_elem.sum())
        return float(torch.linalg.vector_norm(self.bs.data[:n].double()))

    def to_dense(self):
        """The dense tensor, indexed [i0, i1, ...], on the tensor's device."""
        dense = torch.zeros(self.dims, dtype=self.bs.dtype, device=self.bs.data.device)
        for k, b, e in zip(self.keys, self.bs.begin_points, self.bs.end_points):
            sl = tuple(slice(int(x), int(y) + 1) for x, y in zip(b, e))
            dense[sl] += self.block(k)
        return dense

    def __repr__(self):
        return "U1Tensor(legs=%s, flux=%d, blocks=%d)" % (self.legs, self.flux, self.n_blocks)
