"""The benchmark workloads at their full sizes, on the GPU, against the oracle.

* PEPS absorption at D=8, chi=128 (charges -2..2, entries N(0,1) scaled per
  block by 10^U(-3,0)): a = A . conj(A) over the physical leg (12,800 pairs)
  and E . a over (l, l') (44,872 pairs, 716,363,520 output elements), in
  FP64, the FP32 tier and split-BF16, each against the oracle over every
  output block: overall and per-block relative Frobenius error.
* 512 independent rank-4 x rank-4 U(1) contractions in one batch plan,
  captured in a HIP graph: every replay equals the eager execute bit for
  bit, and a sample of contractions matches the oracle, in FP64 and
  split-BF16.
* The D=10, chi=256 enlarged-corner product in the fused-leg view (19
  block-diagonal GEMMs of 4 .. 3600): every block is checked, the blocks up
  to 1368 against the oracle in full and the larger ones on 64 sampled
  output columns against an FP64 product of the same operands.
* The 200,000,000-element rank-6 FP64 permutation: a bitwise round trip and
  sampled elements against the index formula.

Tolerances: FP64 1e-12, the FP32 tier 1e-5, split-BF16 1e-6 (relative
Frobenius, overall and per block).
"""
import os

import numpy as np
import pytest
import torch

from helpers import O, W
from torque_amd import BatchPlan, BlockSparseTensor, default_handle

pytestmark = pytest.mark.gpu

TOL = {"native": 1e-12, "fp32": 1e-5, "bf16x3": 1e-6}
THREADS = min(16, os.cpu_count() or 1)


def block_errors(got, ref, lengths):
    """(overall, largest per-block) relative Frobenius error of got against
    ref, both device tensors split into consecutive blocks of `lengths`."""
    lens = torch.as_tensor(np.asarray(lengths, np.int64), device=got.device)
    d2 = torch.segment_reduce((got - ref).square_(), "sum", lengths=lens)
    r2 = torch.segment_reduce(ref.square(), "sum", lengths=lens)
    overall = float(torch.sqrt(d2.sum() / r2.sum()))
    nz = r2 > 0
    worst = float(torch.sqrt((d2[nz] / r2[nz]).max())) if bool(nz.any()) else 0.0
    return overall, worst


def test_peps_absorption_d8_chi128_three_tiers(cuda):
    E, A, Ac, l1, l2 = W.peps(8, 128, seed=0, log10_scale_range=(-3.0, 0.0))
    TA = O.BS(A.dims, A.begin, A.end)
    ref_aa, p1 = O.contract(TA, A.data, TA, Ac.data, l1, threads=THREADS)
    aa_dims = np.concatenate([A.dims[1:], Ac.dims[1:]])
    Ta = O.BS(aa_dims, p1["out_begin"], p1["out_end"])
    TE = O.BS(E.dims, E.begin, E.end)
    ref_c, p2 = O.contract(TE, E.data, Ta, ref_aa, l2, threads=THREADS)
    assert p1["m"].size == 12800 and p2["m"].size == 44872
    assert ref_c.size == 716363520
    gA = BlockSparseTensor(A.dims, A.data, A.begin, A.end, device=cuda)
    gAc = BlockSparseTensor(Ac.dims, Ac.data, Ac.begin, Ac.end, device=cuda)
    gE = BlockSparseTensor(E.dims, E.data, E.begin, E.end, device=cuda)
    ga = BlockSparseTensor(aa_dims, ref_aa, p1["out_begin"], p1["out_end"], device=cuda)
    ref_aa_d = torch.from_numpy(ref_aa).to(cuda)
    ref_c_d = torch.from_numpy(ref_c).to(cuda)
    del ref_c
    len1 = p1["m"] * p1["n"]
    len2 = p2["m"] * p2["n"]
    for prec in ("native", "fp32", "bf16x3"):
        q1 = gA.plan_contract(gAc, l1, prec)
        aa = torch.empty(q1.out_n_elem, dtype=torch.float64, device=cuda)
        q1.execute(gA.data, gAc.data, aa)
        q2 = gE.plan_contract(ga, l2, prec)
        assert np.array_equal(q2.out_begin, p2["out_begin"])
        assert np.array_equal(q2.out_offsets, p2["out_offsets"])
        out = torch.empty(q2.out_n_elem, dtype=torch.float64, device=cuda)
        q2.execute(gE.data, ga.data, out)
        torch.cuda.synchronize()
        e1, w1 = block_errors(aa, ref_aa_d, len1)
        e2, w2 = block_errors(out, ref_c_d, len2)
        print("D=8 chi=128 %s: A.A* rel err %.2e (block max %.2e), E.a %.2e (block max %.2e)"
              % (prec, e1, w1, e2, w2), flush=True)
        assert e1 <= TOL[prec] and w1 <= TOL[prec], prec
        assert e2 <= TOL[prec] and w2 <= TOL[prec], prec
        del out, aa
        q1.destroy()
        q2.destroy()


def test_batched_512_graph_and_oracle(cuda):
    pairs = W.batched(512, seed=0)
    a_all = torch.tensor(np.concatenate([a.data for a, _, _ in pairs]), device=cuda)
    b_all = torch.tensor(np.concatenate([b.data for _, b, _ in pairs]), device=cuda)
    a_off = np.cumsum([0] + [a.data.size for a, _, _ in pairs])[:-1]
    b_off = np.cumsum([0] + [b.data.size for _, b, _ in pairs])[:-1]
    n_chk = 24
    refs = []
    for a, b, legs in pairs[:n_chk]:
        ref, _ = O.contract(O.BS(a.dims, a.begin, a.end), a.data,
                            O.BS(b.dims, b.begin, b.end), b.data, legs, threads=THREADS)
        refs.append(ref)
    for prec in ("native", "bf16x3"):
        plans = []
        for i, (a, b, legs) in enumerate(pairs):
            A = BlockSparseTensor(a.dims, a_all[a_off[i]:a_off[i] + a.data.size], a.begin,
                                  a.end, device=cuda)
            B = BlockSparseTensor(b.dims, b_all[b_off[i]:b_off[i] + b.data.size], b.begin,
                                  b.end, device=cuda)
            plans.append(A.plan_contract(B, legs, prec))
        c_off = np.cumsum([0] + [p.out_n_elem for p in plans])[:-1]
        batch = BatchPlan(default_handle(), plans, a_off, b_off, c_off)
        assert batch.n_pairs > 100000
        eager = torch.empty(batch.out_n_elem, dtype=torch.float64, device=cuda)
        batch.execute(a_all, b_all, eager)
        c = torch.zeros_like(eager)
        batch.capture(a_all, b_all, c)
        for r in range(3):
            c.zero_()
            torch.cuda.synchronize()
            batch.graph_launch()
            torch.cuda.synchronize()
            assert torch.equal(c, eager), (prec, r)
            tmp = torch.empty_like(eager)  # eager work between replays
            batch.execute(a_all, b_all, tmp)
            torch.cuda.synchronize()
            del tmp
        worst = 0.0
        for i in range(n_chk):
            got = c[c_off[i]:c_off[i] + plans[i].out_n_elem].cpu().numpy()
            e = np.linalg.norm(got - refs[i]) / np.linalg.norm(refs[i])
            worst = max(worst, e)
        print("batched 512 %s: %d pairs, worst of %d contractions vs oracle %.2e"
              % (prec, batch.n_pairs, n_chk, worst), flush=True)
        assert worst <= TOL[prec], prec
        batch.destroy()
        for p in plans:
            p.destroy()
        del eager, c
        torch.cuda.empty_cache()


def test_corner_product_d10_chi256_every_block(cuda):
    c1, c2, legs = W.enlarged_corners(10, 256, seed=0)
    sizes = (c1.end[:, 0] - c1.begin[:, 0] + 1).astype(np.int64)
    offs = np.concatenate([[0], np.cumsum(sizes * sizes)])
    small = np.nonzero(sizes <= 1368)[0]
    large = np.nonzero(sizes > 1368)[0]
    assert sizes.max() == 3600 and large.size == 7

    def sub(t):
        data = np.concatenate([t.data[offs[i]:offs[i + 1]] for i in small])
        return O.BS(t.dims, t.begin[small], t.end[small]), data

    TA, dA = sub(c1)
    TB, dB = sub(c2)
    ref_small, _ = O.contract(TA, dA, TB, dB, legs, threads=THREADS)
    rng = np.random.default_rng(5)
    cols = {}
    ref_cols = {}
    for i in large:
        n = int(sizes[i])
        a = c1.data[offs[i]:offs[i + 1]].reshape(n, n, order="F")
        b = c2.data[offs[i]:offs[i + 1]].reshape(n, n, order="F")
        cols[i] = np.sort(rng.choice(n, 64, replace=False))
        ref_cols[i] = a @ b[:, cols[i]]
    A = BlockSparseTensor(c1.dims, c1.data, c1.begin, c1.end, device=cuda)
    B = BlockSparseTensor(c2.dims, c2.data, c2.begin, c2.end, device=cuda)
    for prec in ("native", "fp32", "bf16x3"):
        p = A.plan_contract(B, legs, prec)
        assert p.n_pairs == sizes.size
        out = torch.empty(p.out_n_elem, dtype=torch.float64, device=cuda)
        p.execute(A.data, B.data, out)
        torch.cuda.synchronize()
        host = out.cpu().numpy()
        got = np.concatenate([host[p.out_offsets[i]:p.out_offsets[i] + sizes[i] ** 2]
                              for i in small])
        errs = {}
        off = 0
        for i in small:
            n2 = int(sizes[i]) ** 2
            r = ref_small[off:off + n2]
            errs[int(sizes[i])] = max(errs.get(int(sizes[i]), 0.0),
                                      float(np.linalg.norm(got[off:off + n2] - r)
                                            / np.linalg.norm(r)))
            off += n2
        for i in large:
            n = int(sizes[i])
            blk = host[p.out_offsets[i]:p.out_offsets[i] + n * n].reshape(n, n, order="F")
            g = blk[:, cols[i]]
            errs[n] = max(errs.get(n, 0.0), float(np.linalg.norm(g - ref_cols[i])
                                                   / np.linalg.norm(ref_cols[i])))
        print("corner product %s: per-block rel err %s" % (
            prec, {k: "%.2e" % v for k, v in sorted(errs.items())}), flush=True)
        assert max(errs.values()) <= TOL[prec], (prec, errs)
        p.destroy()
        del out


def test_permute_rank6_200m(cuda):
    import ctypes

    from torque_amd import lib
    dims = (10, 20, 25, 40, 50, 20)
    n = int(np.prod(dims))
    h = default_handle(cuda.index)
    x = torch.randn(n, dtype=torch.float64, device=cuda)
    y = torch.empty_like(x)
    z = torch.empty_like(x)
    rng = np.random.default_rng(3)
    strides = np.cumprod((1,) + dims[:-1])
    for perm in ((5, 4, 3, 2, 1, 0), (3, 5, 0, 4, 2, 1)):
        d = np.ascontiguousarray(dims, np.int64)
        p = np.ascontiguousarray(perm, np.int32)
        d2 = np.ascontiguousarray(np.asarray(dims)[list(perm)], np.int64)
        inv = np.ascontiguousarray(np.argsort(perm), np.int32)
        for src, dst, dd, pp in ((x, y, d, p), (y, z, d2, inv)):
            h.check(lib().tq_permute(
                h.ptr, 1, 6, dd.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), None,
                pp.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                ctypes.c_void_p(src.data_ptr()), ctypes.c_void_p(dst.data_ptr())))
        torch.cuda.synchronize()
        assert torch.equal(x, z), perm
        # output element j (column-major over dims[perm]) is input element
        # sum_k j_k * stride[perm[k]]
        j = rng.integers(0, n, 1 << 20)
        rem = j.copy()
        src_idx = np.zeros_like(j)
        for k in range(6):
            jk = rem % d2[k]
            rem //= d2[k]
            src_idx += jk * strides[perm[k]]
        jt = torch.from_numpy(j).to(cuda)
        st = torch.from_numpy(src_idx).to(cuda)
        assert torch.equal(y[jt], x[st]), perm
    del x, y, z
    torch.cuda.empty_cache()
