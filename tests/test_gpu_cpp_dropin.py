"""Runs the C++ drop-in test program (the reference's GPU test sections
written against torque::gpu::BlockSparseTensor / DenseTensor)."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

BIN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "test_dropin")


def test_cpp_dropin_reference_sections(cuda):
    assert os.path.exists(BIN), "tests/cpp/test_dropin is not built (torque_amd/build.py)"
    r = subprocess.run([BIN], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                       timeout=100)
    print(r.stdout)
    assert r.returncode == 0, r.stdout
    assert "0 failed" in r.stdout
