"""torque_amd -- MI355X (gfx950) block-sparse tensor contraction.

The compute path is libtorque_hip.so (HIP kernels behind the C-ABI in
include/torque_hip.h); this package is its Python binding and a mirror of
the reference's tensor classes. C++ users include
torque_amd/include/torque/gpu/*.h instead.
"""
from ._lib import TorqueError, lib, LIB_PATH  # noqa: F401
from .tensor import (BatchPlan, BlockSparseTensor, ContractionPlan, DenseTensor, Handle,  # noqa: F401
                     SparseTensor,
                     default_handle, dot_f64, gemm_configs, index_table,
                     nest_sum, normalize_f64, normalize_f64_noalloc, nrm2_f64, scale_f64)
from .u1 import U1Leg, U1Tensor  # noqa: F401

__all__ = ["BatchPlan", "BlockSparseTensor", "ContractionPlan", "DenseTensor", "Handle",
           "SparseTensor", "U1Leg", "U1Tensor",
           "TorqueError", "default_handle", "dot_f64", "gemm_configs",
           "index_table", "nest_sum", "normalize_f64", "normalize_f64_noalloc", "nrm2_f64", "scale_f64", "lib"]
