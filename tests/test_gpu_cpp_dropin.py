"""Runs the C++ drop-in test program (the reference's GPU test sections
written against torque::gpu::BlockSparseTensor / DenseTensor)."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

BIN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "test_dropin")


def test_cpp_dropin_reference_sections(cuda, tmp_path):
    # the program also reads tmp_path/in.tqbs (a block-tensor file written
    # here), contracts it and writes tmp_path/out.tqbs, checked here against
    # the oracle
    import numpy as np

    from helpers import O, W, rel_fro
    from torque_amd import BlockSparseTensor
    assert os.path.exists(BIN), "tests/cpp/test_dropin is not built (torque_amd/build.py)"
    a, _, _ = W.c1(4)
    BlockSparseTensor(a.dims, a.data, a.begin, a.end, device=cuda).save(tmp_path / "in.tqbs")
    r = subprocess.run([BIN, str(tmp_path)], stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                       text=True, timeout=100)
    print(r.stdout)
    assert r.returncode == 0, r.stdout
    assert "0 failed" in r.stdout
    out = BlockSparseTensor.load(tmp_path / "out.tqbs", device="cpu")
    A = O.BS(a.dims, a.begin, a.end)
    ref_c, ref = O.contract(A, a.data, A, a.data, [[2, 0], [3, 1]])
    assert np.array_equal(out.begin_points, ref["out_begin"])
    assert np.array_equal(out.end_points, ref["out_end"])
    assert rel_fro(out.data.numpy(), ref_c) < 1e-12
