"""tq_normalize_f64_noalloc (the entry point that allocates nothing) against
tq_normalize_f64, through the C-ABI: two calls in a row on one GPU, the
second on a smaller input, must give the same norm and the same normalised
values bit for bit, and the norm must match numpy's."""
import ctypes

import numpy as np
import pytest
import torch

from torque_amd import default_handle, lib

pytestmark = pytest.mark.gpu


def _call(name, h, x, norm):
    fn = getattr(lib(), name)
    h.check(fn(h.ptr, x.numel(), ctypes.c_void_p(x.data_ptr()),
               ctypes.c_void_p(norm.data_ptr())))


def test_noalloc_normalize_matches_allocating_entry_point(cuda):
    h = default_handle(cuda.index)
    rng = np.random.default_rng(21)
    # the first call covers the whole scratch (1024 partial sums), the second
    # fewer: a stale partial of the first call would change the second norm
    for n in (3_000_017, 40_001):
        x = rng.standard_normal(n) * 10.0 ** rng.uniform(-3, 3)
        a = torch.tensor(x, device=cuda)
        b = a.clone()
        na = torch.full((1,), -1.0, dtype=torch.float64, device=cuda)
        nb = torch.full((1,), -1.0, dtype=torch.float64, device=cuda)
        _call("tq_normalize_f64", h, a, na)
        _call("tq_normalize_f64_noalloc", h, b, nb)
        torch.cuda.synchronize()
        assert float(na[0]) == float(nb[0]), n
        assert abs(float(nb[0]) - np.linalg.norm(x)) <= 1e-12 * np.linalg.norm(x), n
        assert torch.equal(a, b), n
    # an empty input reports norm 0
    e = torch.zeros(4, dtype=torch.float64, device=cuda)
    ne = torch.full((1,), -1.0, dtype=torch.float64, device=cuda)
    _call("tq_normalize_f64_noalloc", h, e[:0], ne)
    torch.cuda.synchronize()
    assert float(ne[0]) == 0.0
