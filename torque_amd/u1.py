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

Groups of legs fuse into one leg (fuse_legs, U1Tensor.fuse) and fused legs
split back into their parts (U1Tensor.split): each block of the unfused
tensor is one box of a block of the fused tensor, and the device moves the
boxes in one launch (tq_copy_plan_create / tq_copy_execute), planned once per
block structure. A fused block-diagonal product is then a few large GEMMs
where the unfused contraction is many small ones.

The reference's tensors take their blocks as given begin and end points
(src/gpu/block_sparse.cuh:190-283) and keep no charge information; this
module derives the blocks from the charges and addresses them by charge.
"""
import itertools

import numpy as np
import torch

from . import _lib
from ._lib import TorqueError
from .tensor import BOX_DTYPE, BlockSparseTensor, CopyPlan, default_handle


class U1Leg:
    """Charge sectors (charge, size) of one leg, in the order they occupy
    the leg's index range, and the leg's direction (+1 or -1)."""

    def __init__(self, charges, sizes, sign=1, parts=None):
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
        self._index = {q: i for i, q in enumerate(self.charges)}
        # the legs a fused leg was made of (fuse_legs), None for a plain leg
        self.parts = tuple(parts) if parts is not None else None

    def dual(self):
        """The same sectors with the opposite direction (the leg it contracts
        with); a fused leg's parts are dualised with it."""
        parts = None if self.parts is None else [p.dual() for p in self.parts]
        return U1Leg(self.charges, self.sizes, -self.sign, parts)

    def same_sectors(self, other):
        """Same sectors, and for fused legs the same parts in the same order
        (so the sector tuples sit at the same places inside each sector)."""
        if self.charges != other.charges or self.sizes != other.sizes:
            return False
        if (self.parts is None) != (other.parts is None):
            return False
        return self.parts is None or (
            len(self.parts) == len(other.parts) and
            all(p.same_sectors(q) for p, q in zip(self.parts, other.parts)))

    def sector(self, charge):
        """Index of the sector of `charge`."""
        return self._index[int(charge)]

    def charge_at(self, begin):
        """Charge of the sector starting at index `begin`."""
        return self.charges[self._at[int(begin)]]

    def __eq__(self, other):
        return (isinstance(other, U1Leg) and self.same_sectors(other) and
                self.sign == other.sign and
                (self.parts is None or all(p == q for p, q in zip(self.parts, other.parts))))

    def __repr__(self):
        if self.parts is not None:
            return "U1Leg(fused %s, sign=%+d)" % (list(self.parts), self.sign)
        return "U1Leg(%s, %s, sign=%+d)" % (list(self.charges), list(self.sizes), self.sign)


def fuse_legs(legs):
    """The leg fusing `legs` (legs[0] fastest) and the place of each sector
    tuple in it. The fused leg points like legs[0] (sign s) and has one
    sector per distinct charge Q = s * sum_l s_l q_l, in ascending order;
    Q's sector holds the sector tuples of charge Q one after another (leg 0's
    sector fastest), each as a run of prod(sizes) indices ordered leg 0
    fastest. Conservation is kept: s * Q is the tuple's signed charge, and
    the dual legs fuse into the dual of the fused leg.

    Returns (fused leg, {sector-index tuple: (Q, offset in Q's sector)})."""
    legs = list(legs)
    if not legs:
        raise TorqueError(_lib.TQ_ERROR_INVALID_VALUE, "nothing to fuse")
    s = legs[0].sign
    combos = {}
    for rev in itertools.product(*[range(len(l.charges)) for l in legs[::-1]]):
        idx = rev[::-1]
        q = s * sum(l.sign * l.charges[i] for l, i in zip(legs, idx))
        combos.setdefault(q, []).append(idx)
    place, sizes = {}, []
    charges = sorted(combos)
    for q in charges:
        off = 0
        for idx in combos[q]:
            place[idx] = (q, off)
            off += int(np.prod([l.sizes[i] for l, i in zip(legs, idx)]))
        sizes.append(off)
    return U1Leg(charges, sizes, s, parts=legs), place


def structure(legs, flux=0):
    """Charge tuples and begin/end points of every charge-conserving block
    (sum over legs of sign * charge == flux), leg 0's sector fastest."""
    r = len(legs)
    counts = [len(l.charges) for l in legs]
    # every sector tuple, leg 0 fastest: row-major indices over the reversed legs
    idx = np.indices(counts[::-1]).reshape(r, -1)[::-1].T if r else np.zeros((1, 0), np.int64)
    q = np.zeros(len(idx), np.int64)
    for l, leg in enumerate(legs):
        q += leg.sign * np.asarray(leg.charges, np.int64)[idx[:, l]]
    idx = idx[q == flux]
    ch = np.stack([np.asarray(l.charges, np.int64)[idx[:, k]] for k, l in enumerate(legs)],
                  axis=1) if r else np.zeros((len(idx), 0), np.int64)
    begin = np.stack([np.asarray(l.begins, np.int64)[idx[:, k]] for k, l in enumerate(legs)],
                     axis=1) if r else np.zeros((len(idx), 0), np.int64)
    size = np.stack([np.asarray(l.sizes, np.int64)[idx[:, k]] for k, l in enumerate(legs)],
                    axis=1) if r else np.zeros((len(idx), 0), np.int64)
    keys = [tuple(int(x) for x in row) for row in ch]
    n = len(idx)
    return keys, begin.reshape(n, r), (begin + size - 1).reshape(n, r)


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
            if not _opposite(la, lb):
                raise TorqueError(_lib.TQ_ERROR_INVALID_VALUE,
                                  "contracted legs %d and %d must point in opposite "
                                  "directions" % (a, b))
        out = self.bs.contract(other.bs, pairs, precision=precision, merged=True)
        ca = {a for a, _ in pairs}
        cb = {b for _, b in pairs}
        legs = ([l for i, l in enumerate(self.legs) if i not in ca] +
                [l for j, l in enumerate(other.legs) if j not in cb])
        return U1Tensor._wrap(out, legs, self.flux + other.flux)

    # -- leg fusion --
    @classmethod
    def _from_data(cls, legs, flux, data, layout=None):
        """The tensor holding every conserving block of legs / flux, over
        `data` as is (no copy); layout = structure(legs, flux) if known."""
        keys, begin, end = layout if layout is not None else structure(legs, flux)
        t = cls.__new__(cls)
        t.legs = list(legs)
        t.flux = int(flux)
        t.bs = BlockSparseTensor._adopt([l.dim for l in t.legs], begin, end, data, data.dtype)
        t._set_keys(keys)
        return t

    def fusion_plan(self, groups, handle=None):
        """The plan of fuse(groups) for tensors of this block structure
        (cached per handle)."""
        groups = [[int(l) for l in g] for g in groups]
        if any(not g for g in groups) or \
                sorted(l for g in groups for l in g) != list(range(self.rank)):
            raise TorqueError(_lib.TQ_ERROR_INVALID_VALUE,
                              "the groups must name every leg exactly once")
        h = handle or default_handle(self.bs.data.device.index)
        key = ("fuse", self.bs.structure_key(), _legs_sig(self.legs), self.flux,
               str(self.bs.dtype), tuple(tuple(g) for g in groups))
        plan = h.plans.get(key)
        if plan is None:
            fused, places = [], []
            for g in groups:
                if len(g) == 1:
                    leg = self.legs[g[0]]
                    fused.append(leg)
                    places.append({(i,): (q, 0) for i, q in enumerate(leg.charges)})
                else:
                    leg, place = fuse_legs([self.legs[l] for l in g])
                    fused.append(leg)
                    places.append(place)
            layout = structure(fused, self.flux)
            fus = U1Tensor._from_data(fused, self.flux, self.bs.data[:1], layout)
            boxes, covered, _ = _fusion_boxes(self, fus, groups, places)
            n_out = int(fus.bs.block_n_elem.sum())
            copy = CopyPlan(h, self.bs.dtype, boxes, int(self.bs.block_n_elem.sum()), n_out)
            # a fused block the blocks of this tensor do not fill is zeroed first
            zero = bool(np.any(covered != fus.bs.block_n_elem))
            plan = FusionPlan(copy, fused, self.flux, layout, n_out, zero)
            h.plans[key] = plan
        return plan

    def fuse(self, groups):
        """Fuses each group of legs into one leg (fuse_legs): leg g of the
        result fuses legs groups[g] of this tensor, groups[g][0] fastest; a
        group of one leg keeps its leg. Every block of this tensor becomes a
        box of one block of the result (one device copy launch)."""
        return self.fusion_plan(groups).run(self)

    def split_plan(self, handle=None):
        """The plan of split() for tensors of this block structure (cached
        per handle)."""
        h = handle or default_handle(self.bs.data.device.index)
        key = ("split", self.bs.structure_key(), _legs_sig(self.legs), self.flux,
               str(self.bs.dtype))
        plan = h.plans.get(key)
        if plan is None:
            groups, legs, places = [], [], []
            for leg in self.legs:
                parts = list(leg.parts) if leg.parts is not None else [leg]
                groups.append(list(range(len(legs), len(legs) + len(parts))))
                legs += parts
                if leg.parts is None:
                    places.append({(i,): (q, 0) for i, q in enumerate(leg.charges)})
                else:
                    again, place = fuse_legs(parts)
                    if not again.same_sectors(leg) or again.sign != leg.sign:
                        raise TorqueError(_lib.TQ_ERROR_INVALID_VALUE,
                                          "a fused leg does not match the fusion of its parts")
                    places.append(place)
            layout = structure(legs, self.flux)
            unf = U1Tensor._from_data(legs, self.flux, self.bs.data[:1], layout)
            boxes, _, kept = _fusion_boxes(unf, self, groups, places)
            # the boxes move fused -> unfused
            boxes["src_off"], boxes["dst_off"] = boxes["dst_off"].copy(), boxes["src_off"].copy()
            boxes["src_strides"], boxes["dst_strides"] = (boxes["dst_strides"].copy(),
                                                          boxes["src_strides"].copy())
            n_out = int(unf.bs.block_n_elem.sum())
            copy = CopyPlan(h, self.bs.dtype, boxes, int(self.bs.block_n_elem.sum()), n_out)
            # blocks whose fused block this tensor does not hold are zero
            plan = FusionPlan(copy, legs, self.flux, layout, n_out, not bool(kept.all()))
            h.plans[key] = plan
        return plan

    def split(self):
        """Splits every fused leg back into its parts (the inverse of fuse):
        the result's legs are the parts of each fused leg in order, plain
        legs kept, one block per conserving sector tuple."""
        if all(l.parts is None for l in self.legs):
            return self
        return self.split_plan().run(self)

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


def _opposite(la, lb):
    """Whether la and lb point in opposite directions, and so do the parts
    of fused legs pairwise (the sector tuples then sit at the same places)."""
    if la.sign != -lb.sign:
        return False
    return la.parts is None or all(_opposite(p, q) for p, q in zip(la.parts, lb.parts))


def _legs_sig(legs):
    """Hashable description of legs (sectors, directions and parts)."""
    return tuple((l.charges, l.sizes, l.sign,
                  None if l.parts is None else _legs_sig(l.parts)) for l in legs)


def _fusion_boxes(unf, fus, groups, places):
    """Box copies between the blocks of `unf` and those of `fus`, whose leg g
    fuses legs groups[g] of `unf` (sector tuples placed by places[g]).

    Returns (boxes moving unf -> fus, elements covered per fus block, which
    unf blocks have a fus block). A box iterates the unfused block's legs
    group by group (leg groups[0][0] fastest), so the fused side is written
    (or read) in runs along its first leg."""
    nb = unf.n_blocks
    order = [l for g in groups for l in g]
    r = len(order)
    if r > _lib.TQ_MAX_RANK:
        raise TorqueError(_lib.TQ_ERROR_UNSUPPORTED, "rank above TQ_MAX_RANK")
    ubd, ut, uoff = unf.bs.blocks_dimension, unf.bs.index_tables, unf.bs.block_offsets
    ft, foff = fus.bs.index_tables, fus.bs.block_offsets
    src_off = np.zeros(nb, np.int64)
    dst_off = np.zeros(nb, np.int64)
    dims = np.ones((nb, _lib.TQ_MAX_RANK), np.int64)
    sstr = np.zeros((nb, _lib.TQ_MAX_RANK), np.int64)
    dstr = np.zeros((nb, _lib.TQ_MAX_RANK), np.int64)
    kept = np.zeros(nb, bool)
    covered = np.zeros(fus.n_blocks, np.int64)
    for b, key in enumerate(unf.keys):
        fkey, offs = [], []
        for g, place in zip(groups, places):
            q, o = place[tuple(unf.legs[l].sector(key[l]) for l in g)]
            fkey.append(q)
            offs.append(o)
        fb = fus._index.get(tuple(fkey))
        if fb is None:
            continue
        kept[b] = True
        src_off[b] = uoff[b]
        dst_off[b] = foff[fb] + sum(o * int(ft[fb, g]) for g, o in enumerate(offs))
        k = 0
        for g, ls in enumerate(groups):
            c = 1
            for l in ls:
                dims[b, k] = ubd[b, l]
                sstr[b, k] = ut[b, l]
                dstr[b, k] = c * int(ft[fb, g])
                c *= int(ubd[b, l])
                k += 1
        covered[fb] += int(unf.bs.block_n_elem[b])
    boxes = np.zeros(int(kept.sum()), BOX_DTYPE)
    boxes["src_off"], boxes["dst_off"] = src_off[kept], dst_off[kept]
    boxes["rank"] = r
    boxes["dims"], boxes["src_strides"], boxes["dst_strides"] = dims[kept], sstr[kept], dstr[kept]
    return boxes, covered, kept


class FusionPlan:
    """The box copies that move a tensor of one block structure into the
    fused layout of a leg grouping, or back (U1Tensor.fusion_plan,
    U1Tensor.split_plan): planned once, run on any data of that structure."""

    def __init__(self, copy, legs, flux, layout, n_out, zero):
        self.copy = copy
        self.legs, self.flux, self.layout = list(legs), int(flux), layout
        self.n_out = int(n_out)
        self.zero = bool(zero)

    def new_output(self, dtype, device):
        alloc = torch.zeros if self.zero else torch.empty
        return alloc(max(self.n_out, 1), dtype=dtype, device=device)

    def execute(self, src, out):
        """Moves the blocks of data `src` into `out` (from new_output: the
        elements no block fills must already be zero)."""
        self.copy.execute(src, out)

    def run(self, t, out=None):
        if out is None:
            out = self.new_output(t.bs.dtype, t.bs.data.device)
        self.execute(t.bs.data, out)
        return U1Tensor._from_data(self.legs, self.flux, out, self.layout)

    def destroy(self):
        self.copy.destroy()
