"""ctypes binding of libtorque_hip.so (the C-ABI declared in
include/torque_hip.h). Loading fails loudly when the library has not been
built: there is no CPU fallback for any entry point."""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libtorque_hip.so")

TQ_SUCCESS = 0
TQ_ERROR_INVALID_VALUE = 1
TQ_ERROR_DIMENSION_MISMATCH = 2
TQ_ERROR_RANK_MISMATCH = 3
TQ_ERROR_OUT_OF_RANGE = 4
TQ_ERROR_ALLOCATION = 5
TQ_ERROR_HIP = 6
TQ_ERROR_UNSUPPORTED = 7
TQ_ERROR_NOT_INITIALIZED = 8

TQ_F32 = 0
TQ_F64 = 1
TQ_F16 = 2

TQ_PREC_NATIVE = 0
TQ_PREC_F32 = 1
TQ_PREC_BF16X3 = 2

PRECISIONS = {"native": TQ_PREC_NATIVE, "fp64": TQ_PREC_NATIVE,
              "fp32": TQ_PREC_F32, "bf16x3": TQ_PREC_BF16X3}

c_i64_p = ctypes.POINTER(ctypes.c_int64)
c_i32_p = ctypes.POINTER(ctypes.c_int32)
c_f32_p = ctypes.POINTER(ctypes.c_float)


class BsDesc(ctypes.Structure):
    _fields_ = [("rank", ctypes.c_int32), ("n_blocks", ctypes.c_int64),
                ("dims", c_i64_p), ("begin", c_i64_p), ("end", c_i64_p),
                ("strides", c_i64_p), ("offsets", c_i64_p)]


class SpDesc(ctypes.Structure):
    _fields_ = [("rank", ctypes.c_int32), ("dims", c_i64_p), ("table", c_i64_p),
                ("nnz", ctypes.c_int64), ("indices", ctypes.c_void_p),
                ("values", ctypes.c_void_p)]


class PlanInfo(ctypes.Structure):
    _fields_ = [("out_rank", ctypes.c_int32), ("n_legs", ctypes.c_int32),
                ("n_pairs", ctypes.c_int64), ("out_n_elem", ctypes.c_int64),
                ("flops", ctypes.c_double), ("bytes", ctypes.c_double),
                ("a_layout", ctypes.c_int32), ("b_layout", ctypes.c_int32),
                ("n_tiles", ctypes.c_int64), ("n_segments", ctypes.c_int64),
                ("n_matched", ctypes.c_int64)]


TQ_MAX_RANK = 12


class Box(ctypes.Structure):
    """tq_box: one box of a block copy (tq_copy_plan_create)."""
    _fields_ = [("src_off", ctypes.c_int64), ("dst_off", ctypes.c_int64),
                ("rank", ctypes.c_int32), ("reserved", ctypes.c_int32),
                ("dims", ctypes.c_int64 * TQ_MAX_RANK),
                ("src_strides", ctypes.c_int64 * TQ_MAX_RANK),
                ("dst_strides", ctypes.c_int64 * TQ_MAX_RANK)]


class TorqueError(RuntimeError):
    """Raised for a non-zero tq_status; the message keeps the reference's
    "Error: " prefix (src/error.h:8-14)."""

    def __init__(self, status, message):
        super().__init__("Error: " + message)
        self.status = status


_lib = None


def _proto(lib):
    S = ctypes.c_int
    V = ctypes.c_void_p
    H = ctypes.c_void_p
    sig = {
        "tq_create": (S, [ctypes.POINTER(V), ctypes.c_int]),
        "tq_destroy": (S, [H]),
        "tq_set_stream": (S, [H, V]),
        "tq_synchronize": (S, [H]),
        "tq_status_string": (ctypes.c_char_p, [S]),
        "tq_last_error": (ctypes.c_char_p, [H]),
        "tq_malloc": (S, [H, ctypes.POINTER(V), ctypes.c_size_t]),
        "tq_free": (S, [H, V]),
        "tq_memcpy": (S, [H, V, V, ctypes.c_size_t, ctypes.c_int]),
        "tq_memset_zero": (S, [H, V, ctypes.c_size_t]),
        "tq_permute": (S, [H, S, ctypes.c_int32, c_i64_p, c_i64_p, c_i32_p, V, V]),
        "tq_bs_permute": (S, [H, S, ctypes.POINTER(BsDesc), c_i32_p, V, V, c_i64_p]),
        "tq_bs_slice": (S, [H, S, ctypes.POINTER(BsDesc), c_i64_p, V, V, c_i64_p, c_i64_p]),
        "tq_copy_plan_create": (S, [H, S, ctypes.c_int64, ctypes.POINTER(Box), ctypes.c_int64,
                                    ctypes.c_int64, ctypes.POINTER(V)]),
        "tq_copy_execute": (S, [V, V, V]),
        "tq_copy_plan_destroy": (S, [V]),
        "tq_bs_contract_plan": (S, [H, ctypes.POINTER(BsDesc), ctypes.POINTER(BsDesc),
                                    ctypes.c_int32, c_i32_p, S, S, ctypes.POINTER(V)]),
        "tq_bs_contract_plan_merged": (S, [H, ctypes.POINTER(BsDesc), ctypes.POINTER(BsDesc),
                                           ctypes.c_int32, c_i32_p, S, S, ctypes.POINTER(V)]),
        "tq_plan_get_info": (S, [V, ctypes.POINTER(PlanInfo)]),
        "tq_plan_get_output": (S, [V, c_i64_p, c_i64_p, c_i64_p, c_i64_p, c_i64_p]),
        "tq_plan_get_pairs": (S, [V, c_i64_p, c_i64_p, c_i64_p, c_i64_p, c_i64_p,
                                  c_i64_p, c_i64_p]),
        "tq_plan_get_worklist": (S, [V, c_i64_p]),
        "tq_bs_contract_execute": (S, [V, V, V, V]),
        "tq_bs_contract_pack": (S, [V, V, V]),
        "tq_bs_contract_multiply": (S, [V, V, V, V]),
        "tq_cost_features": (S, [c_i64_p, c_i64_p, c_i64_p, ctypes.c_int64, ctypes.c_int,
                                 ctypes.c_int32, ctypes.c_int32, c_f32_p]),
        "tq_plan_features": (S, [V, ctypes.c_int32, ctypes.c_int32, c_f32_p]),
        "tq_costmodel_n_params": (ctypes.c_int64, []),
        "tq_costmodel_create": (S, [H, ctypes.c_uint64, ctypes.POINTER(ctypes.c_void_p)]),
        "tq_costmodel_destroy": (S, [V]),
        "tq_costmodel_get_params": (S, [V, c_f32_p]),
        "tq_costmodel_set_params": (S, [V, c_f32_p]),
        "tq_costmodel_get_scaling": (S, [V, c_f32_p, c_f32_p, c_f32_p]),
        "tq_costmodel_set_scaling": (S, [V, c_f32_p, c_f32_p, c_f32_p]),
        "tq_costmodel_gradient": (S, [V, c_f32_p, c_f32_p, ctypes.c_int64, c_f32_p, c_f32_p]),
        "tq_costmodel_fit": (S, [V, c_f32_p, c_f32_p, ctypes.c_int64, ctypes.c_int32,
                                 ctypes.c_float, c_f32_p]),
        "tq_costmodel_predict": (S, [V, c_f32_p, ctypes.c_int64, ctypes.c_int32, c_f32_p]),
        "tq_plan_choose_config": (S, [V, V, c_i32_p]),
        "tq_plan_set_config": (S, [V, ctypes.c_int32]),
        "tq_plan_set_class_config": (S, [V, ctypes.c_int32, ctypes.c_int32]),
        "tq_plan_get_config": (S, [V, c_i32_p]),
        "tq_plan_set_class_split": (S, [V, ctypes.c_int32, ctypes.c_int32]),
        "tq_plan_get_split": (S, [V, c_i32_p]),
        "tq_batch_create": (S, [H, ctypes.c_int32, ctypes.POINTER(V), c_i64_p, c_i64_p,
                                c_i64_p, ctypes.POINTER(V)]),
        "tq_plan_capture": (S, [V, V, V, V]),
        "tq_plan_graph_launch": (S, [V]),
        "tq_capture_begin": (S, [H]),
        "tq_capture_end": (S, [H, ctypes.POINTER(V)]),
        "tq_graph_launch": (S, [V, H]),
        "tq_graph_destroy": (S, [V]),
        "tq_plan_set_small_k": (S, [V, ctypes.c_int32]),
        "tq_plan_set_split_operands": (S, [V, ctypes.c_int32]),
        "tq_plan_split_range_check": (S, [V, c_i32_p]),
        "tq_plan_autotune": (S, [V, V, V, V, ctypes.c_int32, c_i32_p,
                                 ctypes.POINTER(ctypes.c_float)]),
        "tq_plan_destroy": (S, [V]),
        "tq_gemm_config_count": (ctypes.c_int32, [S]),
        "tq_gemm_config_describe": (S, [S, ctypes.c_int32, c_i32_p, c_i32_p, c_i32_p,
                                        c_i32_p, c_i32_p]),
        "tq_dense_contract": (S, [H, S, ctypes.c_int32, c_i64_p, ctypes.c_int32, c_i64_p,
                                  ctypes.c_int32, c_i32_p, V, V, V]),
        "tq_sparse_contract": (S, [H, S, ctypes.POINTER(SpDesc), ctypes.POINTER(SpDesc),
                                   ctypes.c_int32, c_i32_p, c_i64_p, c_i64_p,
                                   ctypes.POINTER(V), ctypes.POINTER(V)]),
        "tq_sparse_permute_indices": (S, [H, ctypes.c_int32, c_i64_p, c_i64_p, c_i32_p,
                                          ctypes.c_int64, V, V]),
        "tq_dot_f64": (S, [H, ctypes.c_int64, V, V, ctypes.POINTER(ctypes.c_double)]),
        "tq_nrm2_f64": (S, [H, ctypes.c_int64, V, ctypes.POINTER(ctypes.c_double)]),
        "tq_scale_f64": (S, [H, ctypes.c_int64, V, ctypes.c_double]),
        "tq_normalize_f64": (S, [H, ctypes.c_int64, V, V]),
        "tq_normalize_f64_noalloc": (S, [H, ctypes.c_int64, V, V]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    return lib


def lib():
    """The loaded library; raises if libtorque_hip.so is missing."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                "libtorque_hip.so is not built (run torque_amd/build.py); "
                "the HIP path has no fallback")
        _lib = _proto(ctypes.CDLL(LIB_PATH))
    return _lib


def check(status, handle=None):
    if status != TQ_SUCCESS:
        L = lib()
        msg = L.tq_status_string(status).decode()
        if handle is not None:
            extra = L.tq_last_error(handle).decode()
            if extra and extra != msg:
                msg = extra if extra.startswith(msg) else msg + " (" + extra + ")"
        raise TorqueError(status, msg)
