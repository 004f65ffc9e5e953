"""Leg fusion of the charge-aware front end (torque_amd.u1.fuse_legs,
U1Tensor.fuse / split) and the device box copy under it (tq_copy_plan_create
/ tq_copy_execute).

What pins it: the fused sectors against the workload generator's sector
count (an independent convolution of the legs' sector sizes); fusion of legs
with one sector each against numpy's column-major reshape of the dense
tensor; the box list as a bijection onto the fused buffer; fuse then split as
the identity; and a fused block-diagonal contraction, split back, against the
unfused contraction and a dense FP64 tensordot. The box copy is checked
element for element against a numpy restatement of the box formula on random
boxes."""
import numpy as np
import pytest
import torch

from torque_amd import CopyPlan, TorqueError, U1Leg, U1Tensor, default_handle
from torque_amd import workloads as W
from torque_amd.tensor import BOX_DTYPE
from torque_amd.u1 import _fusion_boxes, fuse_legs, structure


def u1leg(lg, sign=1):
    """U1Leg of a workload-generator leg [(charge, begin, size), ...]."""
    return U1Leg([q for q, _, _ in lg], [n for _, _, n in lg], sign)


def emulate(boxes, src, dst):
    """dst[dst_off + sum i_l d_l] = src[src_off + sum i_l s_l] over each box."""
    for b in boxes:
        r = int(b["rank"])
        dims = [int(x) for x in b["dims"][:r]]
        if r == 0:
            dst[b["dst_off"]] = src[b["src_off"]]
            continue
        if min(dims) == 0:
            continue
        idx = np.indices(dims[::-1]).reshape(r, -1)[::-1]
        so = b["src_off"] + (b["src_strides"][:r, None] * idx).sum(0)
        do = b["dst_off"] + (b["dst_strides"][:r, None] * idx).sum(0)
        dst[do] = src[so]


def test_fused_sectors_match_generator():
    lchi, lD, _ = W.peps_legs(10, 256)
    chi, D = u1leg(lchi), u1leg(lD)
    f, place = fuse_legs([chi, D, D.dual()])
    want = W.fused_leg([lchi, lD, lD], [1, 1, -1])
    assert list(f.charges) == [q for q, _, _ in want]
    assert list(f.sizes) == [n for _, _, n in want]
    assert f.sign == 1 and len(place) == 7 ** 3
    # the dual legs fuse into the dual of the fused leg, sector for sector
    g, place2 = fuse_legs([chi.dual(), D.dual(), D])
    assert g == f.dual() and place2 == place


def test_fuse_legs_places_tile_each_sector():
    rng = np.random.default_rng(5)
    for _ in range(20):
        legs = []
        for _ in range(rng.integers(1, 4)):
            qs = sorted(rng.choice(np.arange(-3, 4), rng.integers(1, 4), replace=False))
            legs.append(U1Leg(qs, rng.integers(1, 4, len(qs)), int(rng.choice([-1, 1]))))
        f, place = fuse_legs(legs)
        s = legs[0].sign
        for q, n in zip(f.charges, f.sizes):
            runs = sorted((o, idx) for idx, (qq, o) in place.items() if qq == q)
            off = 0
            prev = None
            for o, idx in runs:
                assert o == off  # the tuples of a sector sit back to back
                assert s * sum(l.sign * l.charges[i] for l, i in zip(legs, idx)) == q
                if prev is not None:  # in the order leg 0 fastest
                    assert idx[::-1] > prev[::-1]
                prev = idx
                off += int(np.prod([l.sizes[i] for l, i in zip(legs, idx)]))
            assert off == n


def cpu_tensor(legs, flux, seed):
    t = U1Tensor(legs, flux, device="cpu")
    n = int(t.bs.block_n_elem.sum())
    t.bs.data[:n] = torch.from_numpy(np.random.default_rng(seed).standard_normal(n))
    return t


def cpu_fuse(t, groups):
    """Fused tensor of t (CPU), its data moved by emulating the box list."""
    fused, places = [], []
    for g in groups:
        if len(g) == 1:
            leg = t.legs[g[0]]
            fused.append(leg)
            places.append({(i,): (q, 0) for i, q in enumerate(leg.charges)})
        else:
            leg, place = fuse_legs([t.legs[l] for l in g])
            fused.append(leg)
            places.append(place)
    fus = U1Tensor(fused, t.flux, device="cpu")
    boxes, covered, kept = _fusion_boxes(t, fus, groups, places)
    assert kept.all()
    out = np.full(fus.bs.data.numel(), np.nan)
    emulate(boxes, t.bs.data.numpy(), out)
    return fus, boxes, covered, out


def test_fusion_boxes_fill_fused_blocks_once():
    a, b, c = U1Leg([-1, 0, 1], [2, 3, 1]), U1Leg([0, 2], [2, 2], -1), U1Leg([-1, 1], [3, 2])
    t = cpu_tensor([a, b, c, a.dual(), c.dual()], 0, 1)
    for groups in ([[0, 1, 2], [3, 4]], [[0, 3], [1, 2, 4]], [[2], [0, 4, 1], [3]]):
        fus, boxes, covered, out = cpu_fuse(t, groups)
        # every conserving block of the unfused tensor lands somewhere, and
        # together the boxes write each fused element exactly once
        n = int(fus.bs.block_n_elem.sum())
        hits = np.zeros(n, np.int64)
        for bx in boxes:
            one = np.zeros(n, np.int64)
            emulate(np.array([bx]), np.ones(int(t.bs.block_n_elem.sum()), np.int64), one)
            hits += one
        assert np.array_equal(hits, np.ones(n, np.int64))
        assert np.array_equal(covered, fus.bs.block_n_elem)
        vals = t.bs.data.numpy()[: int(t.bs.block_n_elem.sum())]
        assert np.array_equal(np.sort(out[:n]), np.sort(vals))


def test_one_sector_fusion_is_column_major_reshape():
    legs = [U1Leg([1], [3]), U1Leg([0], [4]), U1Leg([2], [2], -1), U1Leg([1], [5], -1),
            U1Leg([0], [2])]
    t = cpu_tensor(legs, 1 + 0 - 2 - 1 + 0, 2)
    dense = t.to_dense().numpy()
    fus, _, _, out = cpu_fuse(t, [[0, 1], [2, 3, 4]])
    fus.bs.data[:] = torch.from_numpy(out)
    want = dense.reshape((3 * 4, 2 * 5 * 2), order="F")
    assert np.array_equal(fus.to_dense().numpy(), want)


# -- on the GPU --------------------------------------------------------------

def random_boxes(rng, n_in, n_out, count):
    """Random boxes inside both buffers with disjoint destinations (each box
    owns a slab of out), strides in any order, ranks 0..TQ_MAX_RANK."""
    boxes = np.zeros(count, BOX_DTYPE)
    slab = n_out // count
    for k in range(count):
        r = int(rng.integers(0, 13))
        dims = [int(x) for x in rng.integers(1, 4, r)] if r else []
        if r and rng.random() < 0.1:
            dims[int(rng.integers(r))] = 0  # an empty box
        n = int(np.prod(dims)) if dims else 1
        while n > slab:
            j = int(np.argmax(dims))
            dims[j] -= 1
            n = int(np.prod(dims))
        # destination: a permuted compact layout inside the slab
        order = rng.permutation(r)
        dstr = np.zeros(r, np.int64)
        acc = 1
        for l in order:
            dstr[l] = acc
            acc *= max(dims[l], 1)
        sstr = rng.integers(0, 7, r)
        span = int(sum((d - 1) * s for d, s in zip(dims, sstr))) if n else 0
        boxes[k]["rank"] = r
        boxes[k]["dims"][:r] = dims
        boxes[k]["src_strides"][:r] = sstr
        boxes[k]["dst_strides"][:r] = dstr
        boxes[k]["src_off"] = int(rng.integers(0, n_in - span))
        boxes[k]["dst_off"] = k * slab + int(rng.integers(0, slab - max(n, 1) + 1))
    return boxes


@pytest.mark.gpu
@pytest.mark.parametrize("dtype", [torch.float64, torch.float32, torch.float16])
def test_copy_plan_random_boxes(cuda, dtype):
    rng = np.random.default_rng(11)
    h = default_handle(cuda.index)
    n_in, n_out = 50000, 400000
    boxes = random_boxes(rng, n_in, n_out, 300)
    src = torch.randn(n_in, dtype=torch.float64, device=cuda).to(dtype)
    dst = torch.full((n_out,), 7.0, dtype=dtype, device=cuda)
    plan = CopyPlan(h, dtype, boxes, n_in, n_out)
    plan.execute(src, dst)
    want = np.full(n_out, 7.0, dtype=np.float64)
    emulate(boxes, src.double().cpu().numpy(), want)
    assert np.array_equal(dst.double().cpu().numpy(), want)
    plan.destroy()


@pytest.mark.gpu
def test_copy_plan_refuses_boxes_outside_buffers(cuda):
    h = default_handle(cuda.index)
    box = np.zeros(1, BOX_DTYPE)
    box["rank"] = 2
    box["dims"][0, :2] = [4, 5]
    box["src_strides"][0, :2] = [1, 4]
    box["dst_strides"][0, :2] = [1, 10]
    CopyPlan(h, torch.float64, box, 20, 46).destroy()  # last element 3 + 40 = 43
    with pytest.raises(TorqueError, match="outside"):
        CopyPlan(h, torch.float64, box, 20, 43)
    with pytest.raises(TorqueError, match="outside"):
        CopyPlan(h, torch.float64, box, 19, 46)
    box["src_strides"][0, 1] = -4
    with pytest.raises(TorqueError, match="negative"):
        CopyPlan(h, torch.float64, box, 20, 46)


def corner_legs(chi_sizes, d_sizes, qs):
    chi, D = U1Leg(qs, chi_sizes), U1Leg(qs, d_sizes)
    # an enlarged corner (chi, D ket, D bra | chi, D ket, D bra)
    return [chi, D, D.dual(), chi.dual(), D.dual(), D]


@pytest.mark.gpu
@pytest.mark.parametrize("groups", [[[0, 1, 2], [3, 4, 5]], [[0, 3], [1, 2, 4, 5]],
                                    [[5], [0, 2, 4], [1, 3]]])
def test_fuse_split_round_trip(cuda, groups):
    g = torch.Generator(device=cuda).manual_seed(3)
    legs = corner_legs([2, 3, 4, 3, 1], [1, 2, 1, 1, 1], [-2, -1, 0, 1, 2])
    t = U1Tensor.random(legs, 0, g, device=cuda)
    f = t.fuse(groups)
    assert f.rank == len(groups)
    assert abs(f.norm() - t.norm()) <= 1e-12 * t.norm()
    back = f.split()
    order = [l for gr in groups for l in gr]
    assert back.legs == [t.legs[l] for l in order]
    assert sorted(back.keys) == sorted(tuple(k[l] for l in order) for k in t.keys)
    want = np.transpose(t.to_dense().cpu().numpy(), order)
    assert np.array_equal(back.to_dense().cpu().numpy(), want)


@pytest.mark.gpu
def test_one_sector_fusion_on_device(cuda):
    g = torch.Generator(device=cuda).manual_seed(4)
    legs = [U1Leg([1], [3]), U1Leg([0], [4]), U1Leg([2], [2], -1), U1Leg([1], [5], -1),
            U1Leg([0], [2])]
    t = U1Tensor.random(legs, -2, g, device=cuda)
    f = t.fuse([[0, 1], [2, 3, 4]])
    want = t.to_dense().cpu().numpy().reshape((12, 20), order="F")
    assert np.array_equal(f.to_dense().cpu().numpy(), want)


@pytest.mark.gpu
@pytest.mark.parametrize("precision,tol", [("native", 1e-12), ("bf16x3", 1e-6)])
def test_fused_corner_product(cuda, precision, tol):
    g = torch.Generator(device=cuda).manual_seed(9)
    legs = corner_legs([2, 5, 6, 5, 2], [1, 1, 2, 1, 1], [-2, -1, 0, 1, 2])
    C1 = U1Tensor.random(legs, 0, g, device=cuda)
    C2 = U1Tensor.random(legs, 0, g, device=cuda)
    F1 = C1.fuse([[0, 1, 2], [3, 4, 5]])
    F2 = C2.fuse([[0, 1, 2], [3, 4, 5]])
    # block diagonal in the fused charge: one GEMM per fused sector
    assert all(k[0] == k[1] for k in F1.keys)
    F = F1.contract(F2, [(1, 0)], precision=precision)
    assert F.n_blocks == len(F1.legs[0].charges)
    C = F.split()
    assert C.legs == [C1.legs[0], C1.legs[1], C1.legs[2], C2.legs[3], C2.legs[4], C2.legs[5]]
    direct = C1.contract(C2, [(3, 0), (4, 1), (5, 2)])
    want = np.tensordot(C1.to_dense().cpu().numpy(), C2.to_dense().cpu().numpy(),
                        axes=([3, 4, 5], [0, 1, 2]))
    got = C.to_dense().cpu().numpy()
    assert np.linalg.norm(got - want) <= tol * np.linalg.norm(want)
    if precision == "native":
        dgot = direct.to_dense().cpu().numpy()
        assert np.linalg.norm(dgot - want) <= 1e-12 * np.linalg.norm(want)


@pytest.mark.gpu
def test_fused_legs_must_match_part_for_part(cuda):
    g = torch.Generator(device=cuda).manual_seed(5)
    a, b = U1Leg([0, 1], [2, 3]), U1Leg([0, 1], [3, 2])
    X = U1Tensor.random([a, b, a.dual(), b.dual()], 0, g, device=cuda).fuse([[0, 1], [2, 3]])
    Y = U1Tensor.random([b, a, b.dual(), a.dual()], 0, g, device=cuda).fuse([[0, 1], [2, 3]])
    # X's fused leg 1 is (a*, b*); Y's leg 0 is (b, a): same fused sector
    # sizes by symmetry here, but the sector tuples sit in other places
    with pytest.raises(TorqueError):
        X.contract(Y, [(1, 0)])
    Z = U1Tensor.random([a, b, a.dual(), b.dual()], 0, g, device=cuda).fuse([[0, 1], [2, 3]])
    ok = X.contract(Z, [(1, 0)])
    want = np.tensordot(X.split().to_dense().cpu().numpy(), Z.split().to_dense().cpu().numpy(),
                        axes=([2, 3], [0, 1]))
    got = ok.split().to_dense().cpu().numpy()
    assert np.linalg.norm(got - want) <= 1e-12 * np.linalg.norm(want)


@pytest.mark.gpu
def test_fused_corner_product_full_size(cuda):
    # the D=10, chi=256 enlarged corners as rank-6 U(1) tensors (9,331 blocks,
    # 66.35 M values each): the fused chain (fuse, 19 block-diagonal GEMMs,
    # split) against the unfused merged contraction over three legs (287,503
    # pairs), and the split-BF16 fused product against FP64 block by block
    lchi, lD, _ = W.peps_legs(10, 256)
    chi, D = u1leg(lchi), u1leg(lD)
    legs = [chi, D, D.dual(), chi.dual(), D.dual(), D]
    g = torch.Generator(device=cuda).manual_seed(21)
    C1 = U1Tensor.random(legs, 0, g, device=cuda)
    C2 = U1Tensor.random(legs, 0, g, device=cuda)
    assert C1.n_blocks == 9331 and int(C1.bs.block_n_elem.sum()) == 66349632
    F1 = C1.fuse([[0, 1, 2], [3, 4, 5]])
    F2 = C2.fuse([[0, 1, 2], [3, 4, 5]])
    assert [int(n) for n in F1.legs[0].sizes] == [n for _, _, n in
                                                 W.fused_leg([lchi, lD, lD], [1, 1, -1])]
    F = F1.contract(F2, [(1, 0)])
    C = F.split()
    direct = C1.contract(C2, [(3, 0), (4, 1), (5, 2)])
    assert sorted(direct.keys) == sorted(C.keys)
    # the direct blocks moved into the chain's block order by one copy launch
    boxes = np.zeros(C.n_blocks, BOX_DTYPE)
    for i, k in enumerate(C.keys):
        j = direct._index[k]
        assert np.array_equal(direct.bs.blocks_dimension[j], C.bs.blocks_dimension[i])
        boxes[i]["src_off"] = direct.bs.block_offsets[j]
        boxes[i]["dst_off"] = C.bs.block_offsets[i]
        boxes[i]["rank"] = 1
        boxes[i]["dims"][0] = C.bs.block_n_elem[i]
        boxes[i]["src_strides"][0] = 1
        boxes[i]["dst_strides"][0] = 1
    n = int(C.bs.block_n_elem.sum())
    moved = torch.empty(n, dtype=torch.float64, device=cuda)
    plan = CopyPlan(default_handle(cuda.index), torch.float64, boxes,
                    int(direct.bs.block_n_elem.sum()), n)
    plan.execute(direct.bs.data, moved)
    plan.destroy()
    ref = moved.norm()
    assert float((moved - C.bs.data[:n]).norm() / ref) <= 1e-12
    # split-BF16 on the fused product: every block to 1e-6 of FP64
    Fb = F1.contract(F2, [(1, 0)], precision="bf16x3")
    assert Fb.keys == F.keys
    for k in F.keys:
        x, y = F.block(k), Fb.block(k)
        assert float((x - y).norm() / x.norm()) <= 1e-6, k
