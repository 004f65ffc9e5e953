"""The C-ABI library loads and exports every entry point include/*.h
declares (no compute call: this runs without a GPU)."""
import ctypes
import glob
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b(tq_[a-z0-9_]+)\s*\(", text):
            names.add(m.group(1))
    return names


def test_header_declares_entry_points():
    names = declared_symbols()
    for must in ("tq_create", "tq_bs_contract_plan", "tq_bs_contract_execute",
                 "tq_permute", "tq_bs_permute", "tq_dense_contract", "tq_dot_f64"):
        assert must in names


def test_library_exports_all_declared_symbols():
    from torque_amd import _lib
    lib = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in sorted(declared_symbols()) if not hasattr(lib, n)]
    assert not missing, missing


def test_status_strings_without_gpu():
    from torque_amd import lib
    L = lib()
    assert L.tq_status_string(2).decode() == \
        "The dimensions from two tensors to be contracted do not match"
    assert L.tq_status_string(0).decode() == "success"
