"""Synthetic U(1) block-sparse workloads (PEPS absorption, batched pairs).

A U(1) leg is a list of charge sectors (charge, size) laid end to end; a
tensor with leg signs s_l stores the block of charges (q_0, ..., q_{r-1})
iff sum_l s_l q_l == 0. Blocks are listed in lexicographic charge order with
leg 0 fastest; entries are N(0,1) from a seeded numpy Generator, optionally
scaled per block.

The workloads:
  c1  rank-4 x rank-4, legs of dim 16 = sectors {4,8,4} at charges -1,0,1,
      signs (+,+,-,-), contract A legs {2,3} with B legs {0,1}.
  peps(D, chi) the CTMRG-style absorption E(chi,D,D,chi) . a, with the
      double-layer site a = A . conj(A) over the physical leg of a rank-5
      site A(d=2, D, D, D, D); charges -2..2 (or -3..3) with binomial-shaped
      sector sizes.
  batched  many independent rank-4 x rank-4 contractions over two legs with
      random uneven sectors.
  enlarged_corners(D, chi) the corner product of a CTMRG step in the
      fused-leg view: block-diagonal matrices over the fused (chi, D, D) leg.
"""
from math import comb

import numpy as np


def binomial_sectors(dim, n_charges):
    """Sector sizes proportional to binomial(n_charges-1, k), floor 1,
    adjusted symmetrically (centre first) to sum to dim."""
    n = n_charges - 1
    w = np.array([comb(n, k) for k in range(n_charges)], dtype=np.float64)
    x = dim * w / w.sum()
    s = np.maximum(1, np.floor(x)).astype(np.int64)
    c = n_charges // 2
    order = sorted(range(n_charges), key=lambda k: (abs(k - c), k))
    rem = x - np.floor(x)
    while s.sum() > dim:
        # remove from the largest sector closest to the centre
        k = max(range(n_charges), key=lambda q: (s[q], -abs(q - c)))
        if s[k] <= 1:
            break
        if k != c and s.sum() - dim >= 2:
            s[k] -= 1
            s[n_charges - 1 - k] -= 1
        else:
            s[c] -= 1
    while s.sum() < dim:
        k = max(order, key=lambda q: rem[q])
        if k != c and dim - s.sum() >= 2:
            s[k] += 1
            s[n_charges - 1 - k] += 1
        else:
            s[c] += 1
        rem[k] = -1
        rem[n_charges - 1 - k] = -1
    assert s.sum() == dim and np.all(s >= 1)
    return s


def leg(charges, sizes):
    """A leg: list of (charge, begin, size)."""
    out, b = [], 0
    for q, n in zip(charges, sizes):
        out.append((int(q), b, int(n)))
        b += int(n)
    return out


def leg_dim(lg):
    return sum(n for _, _, n in lg)


def u1_structure(legs, signs, flux=0):
    """Begin/end points (block-major) of every charge-conserving block,
    lexicographic in charges with leg 0 fastest."""
    r = len(legs)
    begins, ends = [], []
    idx = [0] * r
    if r == 0:
        return np.zeros((0, 0), np.int64), np.zeros((0, 0), np.int64)
    while True:
        q = sum(signs[l] * legs[l][idx[l]][0] for l in range(r))
        if q == flux:
            begins.append([legs[l][idx[l]][1] for l in range(r)])
            ends.append([legs[l][idx[l]][1] + legs[l][idx[l]][2] - 1 for l in range(r)])
        l = 0
        while l < r:
            idx[l] += 1
            if idx[l] < len(legs[l]):
                break
            idx[l] = 0
            l += 1
        if l == r:
            break
    return np.array(begins, np.int64).reshape(-1, r), np.array(ends, np.int64).reshape(-1, r)


def random_data(begin, end, rng, log10_scale_range=None):
    """N(0,1) entries per block, optionally times 10^U(lo, hi) per block."""
    sizes = np.prod(end - begin + 1, axis=1) if len(begin) else np.zeros(0, np.int64)
    data = rng.standard_normal(int(sizes.sum()))
    if log10_scale_range is not None:
        lo, hi = log10_scale_range
        off = 0
        for n in sizes:
            data[off:off + n] *= 10.0 ** rng.uniform(lo, hi)
            off += n
    return data


class Tensor:
    """Host-side block-sparse tensor description: dims, block-major
    begin/end, data (compact column-major blocks in order)."""

    def __init__(self, dims, begin, end, data):
        self.dims = np.asarray(dims, np.int64)
        self.begin = begin
        self.end = end
        self.data = data

    @property
    def n_blocks(self):
        return self.begin.shape[0]


def u1_tensor(legs, signs, rng, flux=0, log10_scale_range=None):
    b, e = u1_structure(legs, signs, flux)
    return Tensor([leg_dim(l) for l in legs], b, e, random_data(b, e, rng, log10_scale_range))


def c1(seed=0):
    """Config 1: two rank-4 U(1) tensors, legs 16 = {4,8,4}, signs (+,+,-,-);
    contract A{2,3} with B{0,1}."""
    rng = np.random.default_rng(seed)
    lg = leg([-1, 0, 1], [4, 8, 4])
    signs = [1, 1, -1, -1]
    a = u1_tensor([lg] * 4, signs, rng)
    b = u1_tensor([lg] * 4, signs, rng)
    return a, b, [[2, 0], [3, 1]]


def peps_legs(D, chi):
    n_q = 5 if chi <= 128 else 7
    qs = list(range(-(n_q // 2), n_q // 2 + 1))
    return (leg(qs, binomial_sectors(chi, n_q)), leg(qs, binomial_sectors(D, n_q)),
            leg([-1, 1], [1, 1]))


def peps(D, chi, seed=0, log10_scale_range=None):
    """PEPS absorption step: returns (E, A, Aconj, legs_AA, legs_Ea).

    A(p,u,l,d,r) signs (+,+,+,-,-); conj(A) carries the same blocks and, for
    real data, the same entries. a = A . conj(A) over p gives
    a(u,l,d,r,u',l',d',r'); E(chi,l,l',chi) signs (+,-,+,-); E . a contracts
    E legs 1,2 with a legs 1 (l) and 5 (l')."""
    rng = np.random.default_rng(seed)
    lchi, lD, lp = peps_legs(D, chi)
    A = u1_tensor([lp, lD, lD, lD, lD], [1, 1, 1, -1, -1], rng,
                  log10_scale_range=log10_scale_range)
    Ac = Tensor(A.dims, A.begin, A.end, A.data.copy())
    E = u1_tensor([lchi, lD, lD, lchi], [1, -1, 1, -1], rng,
                  log10_scale_range=log10_scale_range)
    return E, A, Ac, [[0, 0]], [[1, 1], [2, 5]]


def dirichlet_leg(rng, dim_lo=8, dim_hi=96, n_lo=3, n_hi=7):
    dim = int(rng.integers(dim_lo, dim_hi + 1))
    n = int(rng.integers(n_lo, n_hi + 1))
    n = min(n, dim)
    w = rng.dirichlet(np.ones(n))
    s = np.maximum(1, np.floor(w * dim)).astype(np.int64)
    while s.sum() > dim:
        s[np.argmax(s)] -= 1
    while s.sum() < dim:
        s[np.argmax(w * dim - s)] += 1
    qs = [q - n // 2 for q in range(n)]
    return leg(qs, s)


def batched(n_pairs=512, seed=0, dim_lo=8, dim_hi=96):
    """Config 4: independent rank-4 x rank-4 contractions over two legs.
    Returns a list of (A, B, legs)."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n_pairs):
        L = [dirichlet_leg(rng, dim_lo, dim_hi) for _ in range(6)]
        a = u1_tensor([L[0], L[1], L[2], L[3]], [1, 1, -1, -1], rng)
        b = u1_tensor([L[2], L[3], L[4], L[5]], [1, 1, -1, -1], rng)
        out.append((a, b, [[2, 0], [3, 1]]))
    return out


def densify(t):
    """Dense column-major image (sum over covering blocks)."""
    dense = np.zeros(int(np.prod(t.dims)), np.float64)
    view = dense.reshape(tuple(t.dims[::-1]))
    off = 0
    for b, e in zip(t.begin, t.end):
        shape = tuple((e - b + 1)[::-1])
        n = int(np.prod(shape))
        sl = tuple(slice(int(x), int(y) + 1) for x, y in zip(b[::-1], e[::-1]))
        view[sl] += t.data[off:off + n].reshape(shape)
        off += n
    return dense


def fused_leg(legs, signs):
    """Sectors of the fusion of U(1) legs (charge = sum of sign * charge),
    ascending in charge: a list of (charge, begin, size)."""
    sizes = {0: 1}
    for lg, sg in zip(legs, signs):
        nxt = {}
        for q0, n0 in sizes.items():
            for q, _, n in lg:
                nxt[q0 + sg * q] = nxt.get(q0 + sg * q, 0) + n0 * n
        sizes = nxt
    qs = sorted(sizes)
    return leg(qs, [sizes[q] for q in qs])


def enlarged_corners(D=10, chi=256, seed=0):
    """The corner product of a CTMRG step in the fused-leg view: two
    enlarged corners C1, C2, each a matrix over the fused leg (chi, D ket,
    D bra) of dimension chi * D^2, block-diagonal in the fused U(1) charge
    (one block per charge sector), contracted over their shared fused leg
    (C1 leg 1 with C2 leg 0). Charges -3..3 (chi > 128) or -2..2 on each
    leg. Entries N(0,1). Returns (C1, C2, legs)."""
    rng = np.random.default_rng(seed)
    lchi, lD, _ = peps_legs(D, chi)
    f = fused_leg([lchi, lD, lD], [1, 1, -1])
    dim = leg_dim(f)
    begin = np.array([[b, b] for _, b, _ in f], np.int64)
    end = np.array([[b + n - 1, b + n - 1] for _, b, n in f], np.int64)
    c1 = Tensor([dim, dim], begin, end, random_data(begin, end, rng))
    c2 = Tensor([dim, dim], begin.copy(), end.copy(), random_data(begin, end, rng))
    return c1, c2, [[1, 0]]


def ctmrg_step(D, chi, seed=0):
    """Tensors of one CTMRG edge update: absorb the double-layer site into
    the edge, then renormalise the enlarged edge with two projectors.

    Returns (E, A, Ac, P, Q, legs) with legs a list of the four
    contractions' leg pairs:
      a  = A . conj(A) over p                 -> a(u,l,d,r,u',l',d',r')
      E1 = E . a over (l, l')                 -> E1(chi,chi',u,d,r,u',d',r')
      F  = E1 . P over (chi, u, u')           -> F(chi',d,r,d',r',x)
      G  = F . Q over (chi', d, d')           -> G(r,r',x,y)
    and G is then normalised by its FP64 norm. P(chi, u, u', x) and
    Q(chi', d, d', y) are U(1) projectors onto a new leg of dimension chi."""
    rng = np.random.default_rng(seed + 1000)
    E, A, Ac, l_aa, l_ea = peps(D, chi, seed=seed)
    lchi, lD, _ = peps_legs(D, chi)
    P = u1_tensor([lchi, lD, lD, lchi], [1, 1, -1, -1], rng)
    Q = u1_tensor([lchi, lD, lD, lchi], [1, 1, -1, -1], rng)
    legs = [l_aa, l_ea, [[0, 0], [2, 1], [5, 2]], [[0, 0], [1, 1], [3, 2]]]
    return E, A, Ac, P, Q, legs
