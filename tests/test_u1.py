"""The charge-aware front end (torque_amd.u1): block structure from U(1)
charges, blocks addressed by charge tuple, and contractions through the
merged-output plans checked against dense FP64 products of the dense images
(which are independent of the block bookkeeping: a wrong charge key, leg
order or block offset changes the dense result)."""
import numpy as np
import pytest
import torch

from torque_amd import TorqueError, U1Leg, U1Tensor
from torque_amd import workloads as W
from torque_amd.u1 import structure


def leg(qs, ns, sign=1):
    return U1Leg(qs, ns, sign)


def test_structure_matches_brute_force():
    # every sector tuple enumerated by hand: kept iff its signed charges sum
    # to the flux, in the order leg 0 fastest
    legs = [leg([-1, 0, 2], [2, 3, 1]), leg([0, 1], [2, 2], -1), leg([-2, 1], [1, 3])]
    for flux in (-2, 0, 1, 3):
        keys, begin, end = structure(legs, flux)
        want = []
        for i2 in range(2):
            for i1 in range(2):
                for i0 in range(3):
                    idx = (i0, i1, i2)
                    q = sum(l.sign * l.charges[i] for l, i in zip(legs, idx))
                    if q == flux:
                        want.append(idx)
        assert keys == [tuple(l.charges[i] for l, i in zip(legs, w)) for w in want]
        for k, (b, e) in enumerate(zip(begin, end)):
            for l, (lg, i) in enumerate(zip(legs, want[k])):
                assert b[l] == lg.begins[i] and e[l] - b[l] + 1 == lg.sizes[i]


def test_structure_matches_workload_generator():
    lg = W.leg([-1, 0, 1], [4, 8, 4])
    b, e = W.u1_structure([lg] * 4, [1, 1, -1, -1])
    l = leg([-1, 0, 1], [4, 8, 4])
    keys, begin, end = structure([l, l, l.dual(), l.dual()])
    assert len(keys) == 19 and np.array_equal(begin, b) and np.array_equal(end, e)


def test_leg_validation():
    with pytest.raises(TorqueError):
        U1Leg([0, 0], [1, 2])
    with pytest.raises(TorqueError):
        U1Leg([0, 1], [1, 0])
    with pytest.raises(TorqueError):
        U1Leg([0, 1], [1, 2], sign=2)
    with pytest.raises(TorqueError):
        U1Leg([0, 1], [1])
    l = U1Leg([3, -1], [2, 5], -1)
    assert l.dim == 7 and l.begins == (0, 2) and l.charge_at(2) == -1
    assert l.dual() == U1Leg([3, -1], [2, 5], 1) and l.dual() != l


def dense_contract(a, b, pairs):
    """Dense FP64 product on the host over the paired legs, free legs of a
    then of b."""
    ca = [p[0] for p in pairs]
    cb = [p[1] for p in pairs]
    return np.tensordot(a, b, axes=(ca, cb))


@pytest.mark.gpu
@pytest.mark.parametrize("seed", range(3))
def test_contract_rank4(cuda, seed):
    g = torch.Generator(device=cuda).manual_seed(100 + seed)
    rng = np.random.default_rng(seed)
    legs = [leg([-1, 0, 1], [int(x) for x in rng.integers(1, 6, 3)]) for _ in range(6)]
    # A(l0, l1, l2*, l3*), B(l2, l3, l4*, l5*): contract A2-B0, A3-B1
    A = U1Tensor.random([legs[0], legs[1], legs[2].dual(), legs[3].dual()], 0, g, device=cuda)
    B = U1Tensor.random([legs[2], legs[3], legs[4].dual(), legs[5].dual()], 1, g, device=cuda)
    C = A.contract(B, [(2, 0), (3, 1)])
    assert C.flux == 1 and C.rank == 4
    assert C.legs == [A.legs[0], A.legs[1], B.legs[2], B.legs[3]]
    # every output block conserves the result's flux, and its key matches
    # the sectors of its begin points
    for k, b in zip(C.keys, C.bs.begin_points):
        assert sum(l.sign * q for l, q in zip(C.legs, k)) == 1
        assert all(l.charge_at(x) == q for l, x, q in zip(C.legs, b, k))
    want = dense_contract(A.to_dense().cpu().numpy(), B.to_dense().cpu().numpy(),
                          [(2, 0), (3, 1)])
    got = C.to_dense().cpu().numpy()
    assert np.linalg.norm(got - want) <= 1e-12 * np.linalg.norm(want)
    assert abs(C.norm() - np.linalg.norm(want)) <= 1e-12 * np.linalg.norm(want)
    # a block read by key equals the dense image over its sector ranges
    k = C.keys[len(C.keys) // 2]
    sl = tuple(slice(l.begins[l.charges.index(q)], l.begins[l.charges.index(q)] +
                     l.sizes[l.charges.index(q)]) for l, q in zip(C.legs, k))
    assert np.array_equal(C.block(k).cpu().numpy(), got[sl])


@pytest.mark.gpu
def test_transpose_and_chain(cuda):
    g = torch.Generator(device=cuda).manual_seed(7)
    a, b, c = leg([-2, 0, 1], [3, 4, 2]), leg([0, 1], [5, 3]), leg([-1, 1], [2, 6])
    X = U1Tensor.random([a, b, c.dual()], 0, g, device=cuda)
    Y = U1Tensor.random([c, a.dual(), b], -1, g, device=cuda)
    XT = X.transpose([2, 0, 1])
    assert XT.legs == [c.dual(), a, b]
    assert np.array_equal(XT.to_dense().cpu().numpy(),
                          np.transpose(X.to_dense().cpu().numpy(), (2, 0, 1)))
    # X(a, b, c*) . Y(c, a*, b) over c and a: result (b, b), flux -1
    Z = X.contract(Y, [(2, 0), (0, 1)])
    assert Z.legs == [b, b] and Z.flux == -1
    want = dense_contract(X.to_dense().cpu().numpy(), Y.to_dense().cpu().numpy(),
                          [(2, 0), (0, 1)])
    assert np.linalg.norm(Z.to_dense().cpu().numpy() - want) <= 1e-12 * np.linalg.norm(want)


@pytest.mark.gpu
def test_incompatible_legs_refused(cuda):
    g = torch.Generator(device=cuda).manual_seed(3)
    a = leg([0, 1], [2, 3])
    X = U1Tensor.random([a, a.dual()], 0, g, device=cuda)
    Y = U1Tensor.random([a, a.dual()], 0, g, device=cuda)
    with pytest.raises(TorqueError, match="opposite"):
        X.contract(Y, [(0, 0)])  # both legs point the same way
    Z = U1Tensor.random([leg([0, 1], [2, 4]), a.dual()], 0, g, device=cuda)
    with pytest.raises(TorqueError, match="different sectors"):
        X.contract(Z, [(1, 0)])
