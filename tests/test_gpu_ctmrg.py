"""A CTMRG edge update as one HIP graph (tq_capture_begin / tq_capture_end):
the double layer a = A . conj(A), the absorption E . a, the two projector
contractions and the FP64 normalisation of the result on the device
(tq_normalize_f64_noalloc, the norm never leaves the GPU). Every replay of
the graph equals the eager sequence bit for bit, and the result matches the
oracle's chain of contractions (dense image, FP64 to 1e-12) normalised by
its norm."""
import numpy as np
import pytest
import torch

from helpers import O, W, rel_fro
from torque_amd import BlockSparseTensor, default_handle, normalize_f64, normalize_f64_noalloc

pytestmark = pytest.mark.gpu


class Chain:
    def __init__(self, cuda, D, chi, precision="native"):
        E, A, Ac, P, Q, legs = W.ctmrg_step(D, chi)
        self.host = (E, A, Ac, P, Q, legs)
        t = {n: BlockSparseTensor(x.dims, x.data, x.begin, x.end, device=cuda)
             for n, x in zip("E A Ac P Q".split(), (E, A, Ac, P, Q))}
        self.t = t
        self.plans, self.bufs, outs = [], [], []
        lhs = [t["A"], t["E"], None, None]
        rhs = [t["Ac"], None, t["P"], t["Q"]]
        prev = None
        for i in range(4):
            a = lhs[i] if lhs[i] is not None else prev
            b = rhs[i] if rhs[i] is not None else prev
            p = a.plan_contract(b, legs[i], precision, merged=True)
            buf = torch.empty(max(p.out_n_elem, 1), dtype=torch.float64, device=cuda)
            self.plans.append((p, a, b))
            self.bufs.append(buf)
            prev = p.output_tensor(buf)
            outs.append(prev)
        self.out = outs[-1]
        self.norm = torch.empty(1, dtype=torch.float64, device=cuda)

    def step(self, noalloc):
        for (p, a, b), buf in zip(self.plans, self.bufs):
            p.execute(a.data, b.data, buf)
        g = self.bufs[-1][:self.plans[-1][0].out_n_elem]
        if noalloc:
            normalize_f64_noalloc(g, self.norm)
        else:
            normalize_f64(g, self.norm)

    def poison(self):
        for buf in self.bufs:
            buf.fill_(float("nan"))
        self.norm.fill_(float("nan"))


def oracle_chain(E, A, Ac, P, Q, legs):
    """The oracle's (unmerged) chain; returns the dense result and its dims."""
    TA = O.BS(A.dims, A.begin, A.end)
    aa, p1 = O.contract(TA, A.data, TA, Ac.data, legs[0], threads=8)
    ta = O.BS(np.concatenate([A.dims[1:], Ac.dims[1:]]), p1["out_begin"], p1["out_end"])
    e1, p2 = O.contract(O.BS(E.dims, E.begin, E.end), E.data, ta, aa, legs[1], threads=8)
    d1 = np.array([E.dims[0], E.dims[3]] + list(ta.dims[[0, 2, 3, 4, 6, 7]]))
    te1 = O.BS(d1, p2["out_begin"], p2["out_end"])
    f, p3 = O.contract(te1, e1, O.BS(P.dims, P.begin, P.end), P.data, legs[2], threads=8)
    tf = O.BS(np.array(list(d1[[1, 3, 4, 6, 7]]) + [P.dims[3]]), p3["out_begin"],
              p3["out_end"])
    g, p4 = O.contract(tf, f, O.BS(Q.dims, Q.begin, Q.end), Q.data, legs[3], threads=8)
    dg = np.array(list(tf.dims[[2, 4, 5]]) + [Q.dims[3]])
    return O.densify(O.BS(dg, p4["out_begin"], p4["out_end"]), g), dg


def test_ctmrg_step_graph(cuda):
    ch = Chain(cuda, 6, 16)
    h = default_handle(cuda.index)
    ch.step(noalloc=False)  # eager, with the allocating normalisation
    torch.cuda.synchronize()
    eager = [b.clone() for b in ch.bufs]
    eager_norm = float(ch.norm[0])
    with h.capture() as cap:
        ch.step(noalloc=True)
    graph = cap.graph
    for r in range(3):
        ch.poison()
        torch.cuda.synchronize()
        graph.launch()
        torch.cuda.synchronize()
        for i, (b, e) in enumerate(zip(ch.bufs, eager)):
            assert torch.equal(b, e), (r, i)
        assert float(ch.norm[0]) == eager_norm, r
        ch.step(noalloc=False)  # eager work between replays
        torch.cuda.synchronize()
    # against the oracle: dense result normalised by its FP64 norm
    dense_ref, dg = oracle_chain(*ch.host)
    ref_norm = np.linalg.norm(dense_ref)
    out = ch.out
    assert np.array_equal(out.dimension, dg)
    got = O.densify(O.BS(dg, out.begin_points, out.end_points), ch.bufs[-1].cpu().numpy())
    assert abs(eager_norm - ref_norm) <= 1e-12 * ref_norm
    assert rel_fro(got, dense_ref / ref_norm) <= 1e-12
    assert abs(np.linalg.norm(got) - 1.0) <= 1e-14
