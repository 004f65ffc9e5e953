"""torque_amd -- MI355X (gfx950) block-sparse tensor contraction.

The compute path is libtorque_hip.so (HIP kernels behind the C-ABI in
include/torque_hip.h); this package is its Python binding and a mirror of
the reference's tensor classes. C++ users include
torque_amd/include/torque/gpu/*.h instead.
"""
from ._lib import TorqueError, lib, LIB_PATH  # noqa: F401
from .tensor import (BatchPlan, BlockSparseTensor, ContractionPlan, CopyPlan,  # noqa: F401
                     DenseTensor, Handle, SparseTensor,
                     default_handle, dot_f64, gemm_configs, index_table,
                     nest_sum, normalize_f64, normalize_f64_noalloc, nrm2_f64, scale_f64)
from .u1 import FusionPlan, U1Leg, U1Tensor, fuse_legs  # noqa: F401

__all__ = ["BatchPlan", "BlockSparseTensor", "ContractionPlan", "CopyPlan", "DenseTensor",
           "FusionPlan", "Handle", "SparseTensor", "U1Leg", "U1Tensor", "fuse_legs",
           "TorqueError", "default_handle", "dot_f64", "gemm_configs",
           "index_table", "nest_sum", "normalize_f64", "normalize_f64_noalloc", "nrm2_f64", "scale_f64", "lib"]
