"""ctypes wrapper of oracle/liboracle.so -- the CPU restatement of the
reference algorithm. TEST INFRASTRUCTURE ONLY: imported by tests/, by the
smoke check and by the benchmark's CPU-baseline leg, never by product code.
"""
import ctypes
import os

import numpy as np

_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "liboracle.so")
_lib = None
_i64p = ctypes.POINTER(ctypes.c_int64)
_i32p = ctypes.POINTER(ctypes.c_int32)
_f64p = ctypes.POINTER(ctypes.c_double)


class _Bs(ctypes.Structure):
    _fields_ = [("rank", ctypes.c_int32), ("n_blocks", ctypes.c_int64),
                ("dims", _i64p), ("begin", _i64p), ("end", _i64p),
                ("strides", _i64p), ("offsets", _i64p)]


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_PATH):
            raise ImportError("oracle/liboracle.so is not built (make -C oracle)")
        L = ctypes.CDLL(_PATH)
        B = ctypes.POINTER(_Bs)
        L.orc_bs_count_pairs.restype = ctypes.c_int64
        L.orc_bs_count_pairs.argtypes = [B, B, ctypes.c_int32, _i32p]
        L.orc_bs_plan.restype = ctypes.c_int64
        L.orc_bs_plan.argtypes = [B, B, ctypes.c_int32, _i32p] + [_i64p] * 10
        L.orc_bs_contract_f64.restype = ctypes.c_int32
        L.orc_bs_contract_f64.argtypes = [B, _f64p, B, _f64p, ctypes.c_int32, _i32p,
                                          _f64p, ctypes.c_int32]
        L.orc_bs_permute_f64.restype = ctypes.c_int32
        L.orc_bs_permute_f64.argtypes = [B, _f64p, _i32p, _f64p]
        L.orc_dense_permute_f64.restype = ctypes.c_int32
        L.orc_dense_permute_f64.argtypes = [ctypes.c_int32, _i64p, _i64p, _i32p, _f64p, _f64p]
        L.orc_dense_contract_f64.restype = ctypes.c_int32
        L.orc_dense_contract_f64.argtypes = [ctypes.c_int32, _i64p, _f64p, ctypes.c_int32,
                                             _i64p, _f64p, ctypes.c_int32, _i32p, _f64p]
        L.orc_bs_densify_f64.restype = ctypes.c_int32
        L.orc_bs_densify_f64.argtypes = [B, _f64p, _f64p]
        L.orc_sparse_contract_f64.restype = ctypes.c_int64
        L.orc_sparse_contract_f64.argtypes = [ctypes.c_int32, _i64p, _i64p, ctypes.c_int64,
                                              _i64p, _f64p, ctypes.c_int32, _i64p, _i64p,
                                              ctypes.c_int64, _i64p, _f64p, ctypes.c_int32,
                                              _i32p, _i64p, _f64p]
        L.orc_sparse_permute_indices.restype = ctypes.c_int32
        L.orc_sparse_permute_indices.argtypes = [ctypes.c_int32, _i64p, _i64p, _i32p,
                                                 ctypes.c_int64, _i64p, _i64p]
        L.orc_index_table.argtypes = [ctypes.c_int32, _i64p, _i64p]
        L.orc_index_to_indices.argtypes = [ctypes.c_int64, ctypes.c_int32, _i64p, _i64p, _i64p]
        L.orc_nest_sum.argtypes = [ctypes.c_int64, _i64p, ctypes.c_int64, _i64p]
        L.orc_sort_into_matrix_f64.argtypes = [_f64p, ctypes.c_int32, _i64p, _i64p,
                                               ctypes.c_int32, _i32p, _f64p]
        _lib = L
    return _lib


def _p(a, t):
    return a.ctypes.data_as(t) if a is not None and a.size else None


class BS:
    """Metadata of a block-sparse tensor for the oracle: block-major arrays."""

    def __init__(self, dims, begin, end, strides=None, offsets=None):
        self.dims = np.ascontiguousarray(dims, dtype=np.int64).reshape(-1)
        r = self.dims.size
        self.begin = np.ascontiguousarray(begin, dtype=np.int64).reshape(-1, r)
        self.end = np.ascontiguousarray(end, dtype=np.int64).reshape(-1, r)
        bd = self.end - self.begin + 1
        if strides is None:
            strides = np.cumprod(np.concatenate([np.ones((bd.shape[0], 1), np.int64),
                                                 bd[:, :-1]], axis=1), axis=1) if r else \
                np.zeros((bd.shape[0], 0), np.int64)
        self.strides = np.ascontiguousarray(strides, dtype=np.int64).reshape(-1, r)
        n_elem = np.prod(bd, axis=1).astype(np.int64) if bd.shape[0] else np.zeros(0, np.int64)
        if offsets is None:
            offsets = np.concatenate([[0], np.cumsum(n_elem)[:-1]]).astype(np.int64) \
                if n_elem.size else np.zeros(0, np.int64)
        self.offsets = np.ascontiguousarray(offsets, dtype=np.int64)
        self.n_elem = n_elem
        self.c = _Bs(r, self.begin.shape[0], _p(self.dims, _i64p), _p(self.begin, _i64p),
                     _p(self.end, _i64p), _p(self.strides, _i64p), _p(self.offsets, _i64p))

    @classmethod
    def of(cls, t):
        """From a torque_amd.BlockSparseTensor (metadata only)."""
        return cls(t.dimension, t.begin_points, t.end_points, t.index_tables, t.block_offsets)


def _legs(legs):
    return np.ascontiguousarray(np.asarray(legs, dtype=np.int32).reshape(-1, 2))


def plan(a, b, legs):
    L = lib()
    lg = _legs(legs)
    n = L.orc_bs_count_pairs(ctypes.byref(a.c), ctypes.byref(b.c), lg.shape[0], _p(lg, _i32p))
    if n < 0:
        raise ValueError("The dimensions from two tensors to be contracted do not match")
    ro = a.dims.size + b.dims.size - 2 * lg.shape[0]
    out = {k: np.zeros(n, np.int64) for k in
           ("a_block", "b_block", "a_offset", "b_offset", "m", "n", "k", "out_offsets")}
    ob = np.zeros((n, ro), np.int64)
    oe = np.zeros((n, ro), np.int64)
    L.orc_bs_plan(ctypes.byref(a.c), ctypes.byref(b.c), lg.shape[0], _p(lg, _i32p),
                  _p(out["a_block"], _i64p), _p(out["b_block"], _i64p), _p(ob, _i64p),
                  _p(oe, _i64p), _p(out["a_offset"], _i64p), _p(out["b_offset"], _i64p),
                  _p(out["m"], _i64p), _p(out["n"], _i64p), _p(out["k"], _i64p),
                  _p(out["out_offsets"], _i64p))
    out["out_begin"], out["out_end"] = ob, oe
    return out


def contract(a, a_data, b, b_data, legs, threads=1):
    p = plan(a, b, legs)
    ro = p["out_begin"].shape[1] if p["out_begin"].ndim == 2 else 0
    n_out = int((p["m"] * p["n"]).sum()) if ro > 0 else 1
    c = np.zeros(max(n_out, 1), np.float64)
    ad = np.ascontiguousarray(a_data, np.float64)
    bd = np.ascontiguousarray(b_data, np.float64)
    lg = _legs(legs)
    rc = lib().orc_bs_contract_f64(ctypes.byref(a.c), _p(ad, _f64p), ctypes.byref(b.c),
                                   _p(bd, _f64p), lg.shape[0], _p(lg, _i32p),
                                   _p(c, _f64p), int(threads))
    if rc != 0:
        raise ValueError("oracle contract failed")
    return c, p


def bs_permute(t, data, perm):
    out = np.zeros(max(int(t.n_elem.sum()), 1), np.float64)
    pm = np.ascontiguousarray(perm, np.int32)
    d = np.ascontiguousarray(data, np.float64)
    lib().orc_bs_permute_f64(ctypes.byref(t.c), _p(d, _f64p), _p(pm, _i32p), _p(out, _f64p))
    return out


def dense_permute(dims, data, perm, strides=None):
    dims = np.ascontiguousarray(dims, np.int64)
    pm = np.ascontiguousarray(perm, np.int32)
    d = np.ascontiguousarray(data, np.float64)
    out = np.zeros(max(int(np.prod(dims)), 1), np.float64)
    st = None if strides is None else np.ascontiguousarray(strides, np.int64)
    lib().orc_dense_permute_f64(dims.size, _p(dims, _i64p), _p(st, _i64p), _p(pm, _i32p),
                                _p(d, _f64p), _p(out, _f64p))
    return out


def dense_contract(dims_a, a, dims_b, b, legs):
    da = np.ascontiguousarray(dims_a, np.int64)
    db = np.ascontiguousarray(dims_b, np.int64)
    lg = _legs(legs)
    al = set(lg[:, 0].tolist())
    bl = set(lg[:, 1].tolist())
    od = [da[l] for l in range(da.size) if l not in al] + [db[l] for l in range(db.size) if l not in bl]
    out = np.zeros(max(int(np.prod(od)) if od else 1, 1), np.float64)
    A = np.ascontiguousarray(a, np.float64)
    B = np.ascontiguousarray(b, np.float64)
    rc = lib().orc_dense_contract_f64(da.size, _p(da, _i64p), _p(A, _f64p), db.size,
                                      _p(db, _i64p), _p(B, _f64p), lg.shape[0],
                                      _p(lg, _i32p), _p(out, _f64p))
    if rc != 0:
        raise ValueError("The dimensions from two tensors to be contracted do not match")
    return out, od


def densify(t, data):
    n = int(np.prod(t.dims)) if t.dims.size else 1
    out = np.zeros(max(n, 1), np.float64)
    d = np.ascontiguousarray(data, np.float64)
    lib().orc_bs_densify_f64(ctypes.byref(t.c), _p(d, _f64p), _p(out, _f64p))
    return out


def sparse_contract(dims_a, idx_a, val_a, dims_b, idx_b, val_b, legs, table_a=None,
                    table_b=None):
    """Element-sparse contraction; tables default to the compact ones.
    Returns (sorted unique output indices, values)."""
    L = lib()
    da = np.ascontiguousarray(dims_a, np.int64)
    db = np.ascontiguousarray(dims_b, np.int64)
    ta = None if table_a is None else np.ascontiguousarray(table_a, np.int64)
    tb = None if table_b is None else np.ascontiguousarray(table_b, np.int64)
    ia = np.ascontiguousarray(idx_a, np.int64)
    ib = np.ascontiguousarray(idx_b, np.int64)
    va = np.ascontiguousarray(val_a, np.float64)
    vb = np.ascontiguousarray(val_b, np.float64)
    lg = _legs(legs)
    args = [da.size, _p(da, _i64p), _p(ta, _i64p), ia.size, _p(ia, _i64p), _p(va, _f64p),
            db.size, _p(db, _i64p), _p(tb, _i64p), ib.size, _p(ib, _i64p), _p(vb, _f64p),
            lg.shape[0], _p(lg, _i32p)]
    n = L.orc_sparse_contract_f64(*args, None, None)
    if n < 0:
        raise ValueError("The dimensions from two tensors to be contracted do not match")
    oi = np.zeros(max(n, 1), np.int64)
    ov = np.zeros(max(n, 1), np.float64)
    L.orc_sparse_contract_f64(*args, _p(oi, _i64p), _p(ov, _f64p))
    return oi[:n], ov[:n]


def sparse_permute_indices(dims, indices, perm, table=None):
    """Indices of SparseTensor::hard_transpose(perm)."""
    d = np.ascontiguousarray(dims, np.int64)
    t = None if table is None else np.ascontiguousarray(table, np.int64)
    p = np.ascontiguousarray(perm, np.int32)
    i = np.ascontiguousarray(indices, np.int64)
    o = np.zeros(max(i.size, 1), np.int64)
    lib().orc_sparse_permute_indices(d.size, _p(d, _i64p), _p(t, _i64p), _p(p, _i32p), i.size,
                                     _p(i, _i64p), _p(o, _i64p))
    return o[:i.size]


def index_table(dims):
    d = np.ascontiguousarray(dims, np.int64)
    t = np.zeros(d.size, np.int64)
    lib().orc_index_table(d.size, _p(d, _i64p), _p(t, _i64p))
    return t


def index_to_indices(i, table, sort_index=None):
    t = np.ascontiguousarray(table, np.int64)
    s = None if sort_index is None else np.ascontiguousarray(sort_index, np.int64)
    out = np.zeros(t.size, np.int64)
    lib().orc_index_to_indices(int(i), t.size, _p(t, _i64p), _p(s, _i64p), _p(out, _i64p))
    return out


def nest_sum(v, initial=0):
    a = np.ascontiguousarray(v, np.int64)
    out = np.zeros(a.size, np.int64)
    lib().orc_nest_sum(a.size, _p(a, _i64p), int(initial), _p(out, _i64p))
    return out


def sort_into_matrix(block, dims, table, col_legs):
    d = np.ascontiguousarray(dims, np.int64)
    t = np.ascontiguousarray(table, np.int64)
    c = np.ascontiguousarray(col_legs, np.int32)
    b = np.ascontiguousarray(block, np.float64)
    out = np.zeros(max(int(np.prod(d)), 1), np.float64)
    lib().orc_sort_into_matrix_f64(_p(b, _f64p), d.size, _p(d, _i64p), _p(t, _i64p),
                                   c.size, _p(c, _i32p), _p(out, _f64p))
    return out
