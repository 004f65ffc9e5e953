"""Python mirror of the reference's GPU tensor classes over the C-ABI.

`BlockSparseTensor` follows torque::gpu::BlockSparseTensor<T>
(src/gpu/block_sparse.cuh:142-1181) and `DenseTensor` follows
torque::gpu::DenseTensor<T> (src/gpu/dense.cuh:52-604): same member names,
same method names and argument meaning, same "Error: " messages. Device data
lives in a flat torch tensor (torch provides the allocation and the stream);
every contraction and permutation runs in libtorque_hip.so and raises when
the tensor is not on a GPU -- there is no host fallback.

Metadata arrays are numpy int64 and block-major: `begin_points[b, l]` is leg
l of block b (the reference stores one umat column per block).
"""
import ctypes
import hashlib

import numpy as np
import torch

from . import _lib
from ._lib import BsDesc, PlanInfo, TorqueError, check, lib

_DT = {torch.float64: _lib.TQ_F64, torch.float32: _lib.TQ_F32, torch.float16: _lib.TQ_F16}
_DT_OF = {v: k for k, v in _DT.items()}
_DUMP_MAGIC = b"TQBS"


def _i64p(a):
    return a.ctypes.data_as(_lib.c_i64_p) if a is not None and a.size else None


def _i32p(a):
    return a.ctypes.data_as(_lib.c_i32_p) if a is not None and a.size else None


def index_table(dims):
    """Column-major strides, util::generate_index_table (src/util/space.h:10-24)."""
    t = np.ones(len(dims), dtype=np.int64)
    acc = 1
    for i, d in enumerate(dims):
        t[i] = acc
        acc *= int(d)
    return t


def nest_sum(v, initial=0):
    """Exclusive prefix sum, util::nest_sum (src/util/space.h:76-88)."""
    v = np.asarray(v, dtype=np.int64).reshape(-1)
    out = np.empty_like(v)
    if v.size:
        out[0] = initial
        np.cumsum(v[:-1], out=out[1:])
        out[1:] += initial
    return out


class Handle:
    """A library handle on one device; work is issued on torch's current
    stream of that device (replaces the reference's cublasHandle_t)."""

    def __init__(self, device=0):
        self.device = int(device)
        h = ctypes.c_void_p()
        check(lib().tq_create(ctypes.byref(h), self.device))
        self.ptr = h
        self.plans = {}

    def bind_stream(self):
        if getattr(self, "capturing", False):
            return  # the handle records on its own capture stream
        s = torch.cuda.current_stream(self.device).cuda_stream
        check(lib().tq_set_stream(self.ptr, ctypes.c_void_p(s)), self.ptr)

    def capture(self):
        """Context manager recording the handle's work into one HIP graph
        (tq_capture_begin / tq_capture_end); the Graph is set on the returned
        holder's `graph` attribute when the block exits."""
        return _Capture(self)

    def check(self, status):
        check(status, self.ptr)

    def synchronize(self):
        self.check(lib().tq_synchronize(self.ptr))

    def __del__(self):
        try:
            for p in list(self.plans.values()):
                p.destroy()
            if self.ptr:
                lib().tq_destroy(self.ptr)
        except Exception:
            pass


class Graph:
    """A recorded sequence of library calls (tq_graph_t)."""

    def __init__(self, handle, ptr):
        self.h = handle
        self.ptr = ptr

    def launch(self):
        self.h.bind_stream()
        self.h.check(lib().tq_graph_launch(self.ptr, self.h.ptr))

    def __del__(self):
        try:
            if self.ptr:
                lib().tq_graph_destroy(self.ptr)
                self.ptr = None
        except Exception:
            pass


class _Capture:
    def __init__(self, handle):
        self.h = handle
        self.graph = None

    def __enter__(self):
        self.h.bind_stream()
        torch.cuda.synchronize(self.h.device)
        self.h.check(lib().tq_capture_begin(self.h.ptr))
        self.h.capturing = True
        return self

    def __exit__(self, exc_type, exc, tb):
        self.h.capturing = False
        g = ctypes.c_void_p()
        st = lib().tq_capture_end(self.h.ptr, ctypes.byref(g))
        if exc_type is None:
            self.h.check(st)
            self.graph = Graph(self.h, g)
        elif st == 0:
            lib().tq_graph_destroy(g)
        return False


_handles = {}


def default_handle(device=None):
    if device is None:
        device = torch.cuda.current_device()
    if device not in _handles:
        _handles[device] = Handle(device)
    h = _handles[device]
    h.bind_stream()
    return h


def _extent(t):
    """Elements a kernel may read from the tensor's data buffer: one past
    the highest element address of any block."""
    if t.rank == 0:
        return 1
    if t.n_blocks == 0:
        return 0
    last = t.block_offsets + np.sum((t.end_points - t.begin_points) * t.index_tables, axis=1)
    return int(last.max()) + 1


def _require_gpu(t):
    if not t.is_cuda:
        raise TorqueError(_lib.TQ_ERROR_UNSUPPORTED,
                          "tensor data must be on a GPU (no host fallback)")


class BlockSparseTensor:
    """Block-sparse tensor; blocks may overlap and query() sums them."""

    def __init__(self, dimension, source_data=None, begin_points=None,
                 end_points=None, index_tables=None, dtype=torch.float64,
                 device=None):
        dim = np.asarray(dimension, dtype=np.int64).reshape(-1)
        self.rank = int(dim.size)
        self.dimension = dim
        self.dtype = dtype
        self.device = torch.device(device) if device is not None else (
            torch.device("cuda") if torch.cuda.is_available() else torch.device("cpu"))
        if begin_points is None:
            # explicit BlockSparseTensor(dimension): one zeroed element
            # (src/gpu/block_sparse.cuh:180-186)
            self.begin_points = np.zeros((0, self.rank), dtype=np.int64)
            self.end_points = np.zeros((0, self.rank), dtype=np.int64)
            self.index_tables = np.zeros((0, self.rank), dtype=np.int64)
            self.data = torch.zeros(1, dtype=dtype, device=self.device)
            self._meta_changed()
            return
        b = np.asarray(begin_points, dtype=np.int64).reshape(-1, self.rank)
        e = np.asarray(end_points, dtype=np.int64).reshape(-1, self.rank)
        self.begin_points, self.end_points = b.copy(), e.copy()
        bd = e - b + 1
        if index_tables is None:
            self.index_tables = np.stack([index_table(r) for r in bd]) if len(bd) else \
                np.zeros((0, self.rank), dtype=np.int64)
        else:
            self.index_tables = np.asarray(index_tables, dtype=np.int64).reshape(-1, self.rank).copy()
        n = int(np.prod(bd, axis=1).sum()) if len(bd) else 0
        if source_data is None:
            raise TorqueError(_lib.TQ_ERROR_INVALID_VALUE, "Source data not allocated!")
        src = torch.as_tensor(source_data).reshape(-1)
        if src.numel() < n:
            # the kernels read every block element: a short buffer would be
            # read past its end
            raise TorqueError(_lib.TQ_ERROR_INVALID_VALUE,
                              "source data holds %d elements, the blocks need %d"
                              % (src.numel(), n))
        self.data = src[:max(n, 1)].to(device=self.device, dtype=dtype).clone()
        self._meta_changed()

    # -- binary block-tensor file (fixtures, cross-run comparison) --
    def save(self, path):
        """Writes the tensor as a binary block-tensor file: "TQBS", then
        little-endian u32 version (1), u32 dtype (0 FP32, 1 FP64, 2 FP16), u32
        rank, u64 n_blocks, u64 dims[rank], u64 begin, end and index tables
        [n_blocks][rank] (block-major), u64 n_elem, the n_elem data values.
        The C++ classes read and write the same format."""
        data = self.data.detach().to("cpu").contiguous()
        with open(path, "wb") as f:
            f.write(_DUMP_MAGIC)
            np.array([1, _DT[self.dtype], self.rank], "<u4").tofile(f)
            np.array([self.n_blocks], "<u8").tofile(f)
            for a in (self.dimension, self.begin_points, self.end_points, self.index_tables):
                np.ascontiguousarray(a, "<u8").tofile(f)
            np.array([data.numel()], "<u8").tofile(f)
            f.write(data.numpy().astype(data.numpy().dtype.newbyteorder("<")).tobytes())

    @classmethod
    def load(cls, path, device=None):
        """Reads a binary block-tensor file written by save (or by the C++
        classes) onto `device`."""
        with open(path, "rb") as f:
            if f.read(4) != _DUMP_MAGIC:
                raise TorqueError(_lib.TQ_ERROR_INVALID_VALUE, "not a block-tensor file")
            version, code, rank = (int(x) for x in np.fromfile(f, "<u4", 3))
            if version != 1 or code not in _DT_OF:
                raise TorqueError(_lib.TQ_ERROR_UNSUPPORTED, "block-tensor file version or dtype")
            nb = int(np.fromfile(f, "<u8", 1)[0])
            dims = np.fromfile(f, "<u8", rank).astype(np.int64)
            meta = [np.fromfile(f, "<u8", nb * rank).astype(np.int64).reshape(nb, rank)
                    for _ in range(3)]
            n = int(np.fromfile(f, "<u8", 1)[0])
            dtype = _DT_OF[code]
            np_dtype = {torch.float64: "<f8", torch.float32: "<f4", torch.float16: "<f2"}[dtype]
            data = np.fromfile(f, np_dtype, n)
            if data.size != n:
                raise TorqueError(_lib.TQ_ERROR_INVALID_VALUE, "truncated block-tensor file")
        if nb == 0:
            t = cls(dims, dtype=dtype, device=device)
            t.data = torch.from_numpy(data.copy()).to(t.device)
            return t
        return cls(dims, torch.from_numpy(data.copy()), meta[0], meta[1], meta[2], dtype=dtype,
                   device=device)

    # -- derived metadata (src/gpu/block_sparse.cuh:199-212) --
    @classmethod
    def _adopt(cls, dimension, begin, end, data, dtype):
        """A tensor over `data` as is (no copy), compact column-major blocks
        in the order given."""
        res = cls.__new__(cls)
        res.dimension = np.asarray(dimension, dtype=np.int64).reshape(-1).copy()
        res.rank = int(res.dimension.size)
        res.dtype, res.device = dtype, data.device
        res.begin_points = np.asarray(begin, np.int64).reshape(-1, res.rank).copy()
        res.end_points = np.asarray(end, np.int64).reshape(-1, res.rank).copy()
        bd = res.end_points - res.begin_points + 1
        res.index_tables = (np.concatenate([np.ones((len(bd), 1), np.int64),
                                            np.cumprod(bd[:, :-1], axis=1)], axis=1)
                            if len(bd) and res.rank else np.zeros((len(bd), res.rank), np.int64))
        res.data = data
        res._meta_changed()
        return res

    def _meta_changed(self):
        self.blocks_dimension = self.end_points - self.begin_points + 1
        self.block_n_elem = (np.prod(self.blocks_dimension, axis=1).astype(np.int64)
                             if len(self.blocks_dimension) else np.zeros(0, np.int64))
        self.block_offsets = nest_sum(self.block_n_elem)
        self._desc = None
        self._key = None

    @property
    def n_blocks(self):
        return int(self.begin_points.shape[0])

    def desc(self):
        if self._desc is None:
            self._arrs = [np.ascontiguousarray(a, dtype=np.int64) for a in (
                self.dimension, self.begin_points, self.end_points,
                self.index_tables, self.block_offsets)]
            d = BsDesc()
            d.rank = self.rank
            d.n_blocks = self.n_blocks
            d.dims, d.begin, d.end, d.strides, d.offsets = [_i64p(a) for a in self._arrs]
            self._desc = d
        return self._desc

    def structure_key(self):
        if self._key is None:
            h = hashlib.sha1()
            for a in (self.dimension, self.begin_points, self.end_points,
                      self.index_tables, self.block_offsets):
                h.update(np.ascontiguousarray(a).tobytes())
                h.update(b"|")
            self._key = h.hexdigest()
        return self._key

    # -- element access (src/gpu/block_sparse.cuh:448-552) --
    def _check_indices(self, indices):
        idx = np.asarray(indices, dtype=np.int64).reshape(-1)
        if idx.size != self.rank:
            raise TorqueError(_lib.TQ_ERROR_RANK_MISMATCH, "Rank does not match")
        if np.any(idx >= self.dimension) or np.any(idx < 0):
            raise TorqueError(_lib.TQ_ERROR_OUT_OF_RANGE, "Indices out of boundary")
        return idx

    def _covering(self, idx):
        if self.n_blocks == 0:
            return np.zeros(0, dtype=np.int64)
        ok = np.all((idx >= self.begin_points) & (idx <= self.end_points), axis=1)
        return np.nonzero(ok)[0]

    def _address(self, blk, idx):
        rel = idx - self.begin_points[blk]
        return int(self.block_offsets[blk] + np.sum(rel * self.index_tables[blk]))

    def query(self, indices):
        idx = self._check_indices(indices)
        if self.rank == 0:
            return float(self.data[0])
        cov = self._covering(idx)
        total = 0.0
        for blk in cov:
            total += float(self.data[self._address(blk, idx)])
        return total

    def modify(self, indices, number):
        idx = self._check_indices(indices)
        if self.rank == 0:
            self.data[0] = number
            return
        cov = self._covering(idx)
        if len(cov):
            self.data[self._address(cov[0], idx)] = number
            for blk in cov[1:]:
                self.data[self._address(blk, idx)] = 0
        elif number != 0:
            self.append_block([number], idx, idx, np.zeros(self.rank, dtype=np.int64))

    def to_number(self):
        assert self.rank == 0
        return float(self.data[0])

    def append_block(self, source_data, begin_point, end_point, index_table):
        """src/gpu/block_sparse.cuh:376-410"""
        b = np.asarray(begin_point, dtype=np.int64).reshape(1, self.rank)
        e = np.asarray(end_point, dtype=np.int64).reshape(1, self.rank)
        t = np.asarray(index_table, dtype=np.int64).reshape(1, self.rank)
        n = int(np.prod(e - b + 1))
        had_blocks = self.n_blocks > 0
        old_n = int(self.block_n_elem.sum())
        src = torch.as_tensor(source_data).reshape(-1)
        if src.numel() < n:
            raise TorqueError(_lib.TQ_ERROR_INVALID_VALUE,
                              "source data holds %d elements, the block needs %d"
                              % (src.numel(), n))
        src = src[:n].to(self.device, self.dtype)
        self.begin_points = np.concatenate([self.begin_points, b])
        self.end_points = np.concatenate([self.end_points, e])
        self.index_tables = np.concatenate([self.index_tables, t])
        keep = self.data[:old_n] if had_blocks else self.data[:0]
        self.data = torch.cat([keep, src]).contiguous()
        self._meta_changed()

    def soft_transpose(self, permutation):
        perm = np.asarray(permutation, dtype=np.int64)
        if perm.size != self.rank:
            raise TorqueError(_lib.TQ_ERROR_RANK_MISMATCH,
                              "The number of permutation does not match the rank of tensor")
        self.dimension = self.dimension[perm]
        self.begin_points = self.begin_points[:, perm]
        self.end_points = self.end_points[:, perm]
        self.index_tables = self.index_tables[:, perm]
        self._meta_changed()

    # -- GPU operations --
    def hard_transpose(self, permutation):
        """BlockSparseTensor::hard_transpose (src/gpu/block_sparse.cuh:1031-1122)."""
        perm = np.asarray(permutation, dtype=np.int32).reshape(-1)
        if perm.size != self.rank:
            raise TorqueError(_lib.TQ_ERROR_RANK_MISMATCH,
                              "The number of permutation does not match the rank of tensor")
        _require_gpu(self.data)
        h = default_handle(self.data.device.index)
        n = int(self.block_n_elem.sum())
        out = torch.empty(max(n, 1), dtype=self.dtype, device=self.data.device)
        offs = np.zeros(self.n_blocks, dtype=np.int64)
        h.check(lib().tq_bs_permute(h.ptr, _DT[self.dtype], ctypes.byref(self.desc()),
                                    _i32p(perm), ctypes.c_void_p(self.data.data_ptr()),
                                    ctypes.c_void_p(out.data_ptr()), _i64p(offs)))
        res = BlockSparseTensor.__new__(BlockSparseTensor)
        res.rank, res.dtype, res.device = self.rank, self.dtype, self.data.device
        res.dimension = self.dimension[perm]
        res.begin_points = self.begin_points[:, perm].copy()
        res.end_points = self.end_points[:, perm].copy()
        bd = res.end_points - res.begin_points + 1
        res.index_tables = (np.stack([index_table(r) for r in bd]) if len(bd)
                            else np.zeros((0, self.rank), np.int64))
        res.data = out
        res._meta_changed()
        return res

    def slice(self, divisor):
        """BlockSparseTensor::slice (src/gpu/block_sparse.cuh:1124-1176)
        through tq_bs_slice: every block split into prod(divisor) equal
        sub-blocks (sub-block j decoded leg 0 fastest over the divisor), each
        holding the compact copy of its region of the parent block."""
        div = np.ascontiguousarray(np.asarray(divisor, dtype=np.int64).reshape(-1))
        if div.size != self.rank:
            raise TorqueError(_lib.TQ_ERROR_RANK_MISMATCH, "Rank does not match")
        _require_gpu(self.data)
        h = default_handle(self.data.device.index)
        nb = self.n_blocks * int(np.prod(div))
        nbeg = np.zeros((nb, self.rank), np.int64)
        nend = np.zeros((nb, self.rank), np.int64)
        out = torch.empty(max(int(self.block_n_elem.sum()), 1), dtype=self.dtype,
                          device=self.data.device)
        h.check(lib().tq_bs_slice(h.ptr, _DT[self.dtype], ctypes.byref(self.desc()),
                                  _i64p(div), ctypes.c_void_p(self.data.data_ptr()),
                                  ctypes.c_void_p(out.data_ptr()), _i64p(nbeg), _i64p(nend)))
        return BlockSparseTensor(self.dimension, out, nbeg, nend, dtype=self.dtype,
                                 device=self.data.device)

    def plan_contract(self, tensor, contracting_indices, precision="native", handle=None,
                      merged=False):
        """Build (or reuse) the device plan of contract(tensor, legs); merged
        plans write one block per distinct output block (tq_bs_contract_plan_merged)."""
        legs = np.asarray(contracting_indices, dtype=np.int32).reshape(-1, 2)
        h = handle or default_handle(self.data.device.index if self.data.is_cuda else None)
        key = (self.structure_key(), tensor.structure_key(), legs.tobytes(),
               str(self.dtype), precision, bool(merged))
        plan = h.plans.get(key)
        if plan is None:
            plan = ContractionPlan(h, self, tensor, legs, precision, merged)
            if len(h.plans) > 64:
                # drop the cache's reference only: a caller may still hold the
                # plan, which frees itself when the last reference goes
                h.plans.pop(next(iter(h.plans)))
            h.plans[key] = plan
        return plan

    def contract(self, tensor, contracting_indices, precision="native", handle=None,
                 merged=False):
        """BlockSparseTensor::contract (src/gpu/block_sparse.cuh:558-1026):
        one output block per matched pair, ordered by A block then B block
        (merged=True: one block per distinct output block)."""
        if tensor.dtype != self.dtype:
            raise TorqueError(_lib.TQ_ERROR_UNSUPPORTED, "operands must share a dtype")
        _require_gpu(self.data)
        _require_gpu(tensor.data)
        plan = self.plan_contract(tensor, contracting_indices, precision, handle, merged)
        out = torch.empty(max(plan.out_n_elem, 1), dtype=self.dtype,
                          device=self.data.device)
        plan.execute(self.data, tensor.data, out)
        return plan.output_tensor(out)


BOX_DTYPE = np.dtype([("src_off", "<i8"), ("dst_off", "<i8"), ("rank", "<i4"),
                      ("reserved", "<i4"), ("dims", "<i8", _lib.TQ_MAX_RANK),
                      ("src_strides", "<i8", _lib.TQ_MAX_RANK),
                      ("dst_strides", "<i8", _lib.TQ_MAX_RANK)])
assert BOX_DTYPE.itemsize == ctypes.sizeof(_lib.Box)


class CopyPlan:
    """Owner of a tq_copy_plan_t: boxes moved from a buffer of in_elems
    elements into one of out_elems (tq_copy_plan_create). `boxes` is a
    structured array of BOX_DTYPE, one row per tq_box."""

    def __init__(self, handle, dtype, boxes, in_elems, out_elems):
        boxes = np.ascontiguousarray(boxes, dtype=BOX_DTYPE)
        self.h = handle
        self.dtype = dtype
        self.in_elems, self.out_elems = int(in_elems), int(out_elems)
        self.n_boxes = int(boxes.shape[0])
        p = ctypes.c_void_p()
        handle.check(lib().tq_copy_plan_create(
            handle.ptr, _DT[dtype], self.n_boxes,
            boxes.ctypes.data_as(ctypes.POINTER(_lib.Box)), self.in_elems, self.out_elems,
            ctypes.byref(p)))
        self.ptr = p

    def execute(self, src, dst):
        """Copies the boxes from src to dst on torch's current stream."""
        if src.dtype != self.dtype or dst.dtype != self.dtype:
            raise TorqueError(_lib.TQ_ERROR_UNSUPPORTED, "buffers must have the plan's dtype")
        if src.numel() < self.in_elems or dst.numel() < self.out_elems:
            raise TorqueError(_lib.TQ_ERROR_INVALID_VALUE,
                              "buffer shorter than the planned copy")
        _require_gpu(src)
        _require_gpu(dst)
        self.h.bind_stream()
        self.h.check(lib().tq_copy_execute(self.ptr, ctypes.c_void_p(src.data_ptr()),
                                           ctypes.c_void_p(dst.data_ptr())))

    def destroy(self):
        if getattr(self, "ptr", None):
            lib().tq_copy_plan_destroy(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


class ContractionPlan:
    """Owner of a tq_plan_t: device matching, index maps, pack jobs and the
    grouped-GEMM work list of one contraction structure."""

    def __init__(self, handle, a, b, legs, precision="native", merged=False):
        self.h = handle
        legs = np.ascontiguousarray(np.asarray(legs, dtype=np.int32).reshape(-1, 2))
        self.legs = legs
        self.dtype = a.dtype
        self.precision = precision
        self.merged = bool(merged)
        p = ctypes.c_void_p()
        entry = lib().tq_bs_contract_plan_merged if merged else lib().tq_bs_contract_plan
        handle.check(entry(
            handle.ptr, ctypes.byref(a.desc()), ctypes.byref(b.desc()),
            legs.shape[0], _i32p(legs), _DT[a.dtype], _lib.PRECISIONS[precision],
            ctypes.byref(p)))
        self.ptr = p
        info = PlanInfo()
        handle.check(lib().tq_plan_get_info(p, ctypes.byref(info)))
        self.out_rank = info.out_rank
        self.n_pairs = info.n_pairs
        self.n_matched = info.n_matched
        self.out_n_elem = info.out_n_elem
        self.flops = info.flops
        self.bytes = info.bytes
        self.a_layout, self.b_layout = info.a_layout, info.b_layout
        ro, npairs = self.out_rank, self.n_pairs
        self.out_dims = np.zeros(ro, np.int64)
        self.out_begin = np.zeros((npairs, ro), np.int64)
        self.out_end = np.zeros((npairs, ro), np.int64)
        self.out_strides = np.zeros((npairs, ro), np.int64)
        self.out_offsets = np.zeros(npairs, np.int64)
        handle.check(lib().tq_plan_get_output(
            p, _i64p(self.out_dims), _i64p(self.out_begin), _i64p(self.out_end),
            _i64p(self.out_strides), _i64p(self.out_offsets)))
        self.device = a.data.device
        self.a_extent = _extent(a)
        self.b_extent = _extent(b)

    def _check_sizes(self, a_data, b_data, out=None):
        # the kernels take raw pointers: buffers too short for the block
        # structure would be read (or written) past their end
        if a_data.numel() < self.a_extent or b_data.numel() < self.b_extent:
            raise TorqueError(_lib.TQ_ERROR_INVALID_VALUE,
                              "operand data shorter than the planned block structure")
        if out is not None and out.numel() < self.out_n_elem:
            raise TorqueError(_lib.TQ_ERROR_INVALID_VALUE,
                              "output buffer shorter than the planned output")
        for t in (a_data, b_data) + ((out,) if out is not None else ()):
            _require_gpu(t)

    def work(self):
        """(MFMA tiles, streamed 64 x 64 segments) of the current configuration."""
        info = PlanInfo()
        self.h.check(lib().tq_plan_get_info(self.ptr, ctypes.byref(info)))
        return info.n_tiles, info.n_segments

    def pairs(self):
        n = self.n_matched
        arrs = [np.zeros(n, np.int64) for _ in range(7)]
        self.h.check(lib().tq_plan_get_pairs(self.ptr, *[_i64p(x) for x in arrs]))
        keys = ["a_block", "b_block", "a_offset", "b_offset", "m", "n", "k"]
        return dict(zip(keys, arrs))

    def worklist(self):
        order = np.zeros(self.n_pairs, np.int64)
        self.h.check(lib().tq_plan_get_worklist(self.ptr, _i64p(order)))
        return order

    def autotune(self, a_data, b_data, out, reps=3):
        """Time every tile configuration of each MFMA tile class
        (tq_plan_autotune); returns ([class 0, class 1] chosen configuration,
        [class 0 mean ms per configuration, class 1 ...])."""
        self._check_sizes(a_data, b_data, out)
        self.h.bind_stream()
        n = lib().tq_gemm_config_count(_lib.PRECISIONS[self.precision])
        times = (ctypes.c_float * max(2 * n, 1))()
        best = (ctypes.c_int32 * 2)()
        self.h.check(lib().tq_plan_autotune(
            self.ptr, ctypes.c_void_p(a_data.data_ptr()), ctypes.c_void_p(b_data.data_ptr()),
            ctypes.c_void_p(out.data_ptr()), int(reps), best, times))
        self._refresh_layouts()
        return [best[0], best[1]], [[times[c * n + i] for i in range(n)] for c in range(2)]

    def config(self):
        """Current tile configuration of [class 0, class 1]."""
        cfg = (ctypes.c_int32 * 2)()
        self.h.check(lib().tq_plan_get_config(self.ptr, cfg))
        return [cfg[0], cfg[1]]

    def set_class_config(self, cls, cfg):
        self.h.check(lib().tq_plan_set_class_config(self.ptr, int(cls), int(cfg)))

    def split(self):
        """Split-K of [class 0, class 1] (tq_plan_get_split)."""
        sp = (ctypes.c_int32 * 2)()
        self.h.check(lib().tq_plan_get_split(self.ptr, sp))
        return [sp[0], sp[1]]

    def set_class_split(self, cls, split):
        self.h.check(lib().tq_plan_set_class_split(self.ptr, int(cls), int(split)))

    def set_small_k(self, max_k):
        self.h.check(lib().tq_plan_set_small_k(self.ptr, int(max_k)))

    def set_config(self, cfg):
        self.h.check(lib().tq_plan_set_config(self.ptr, int(cfg)))

    def set_split_operands(self, form):
        """Split-BF16 operand form: "auto", "panels" or "in_place"
        (tq_plan_set_split_operands); refreshes the reported layouts."""
        code = {"auto": 0, "panels": 1, "in_place": 2}[form]
        self.h.check(lib().tq_plan_set_split_operands(self.ptr, code))
        self._refresh_layouts()

    def split_range_check(self):
        """True when a split-BF16 execute since the last check split an
        operand value outside the tier's domain (tq_plan_split_range_check)."""
        v = (ctypes.c_int32 * 1)()
        self.h.check(lib().tq_plan_split_range_check(self.ptr, v))
        return bool(v[0])

    def _refresh_layouts(self):
        info = PlanInfo()
        self.h.check(lib().tq_plan_get_info(self.ptr, ctypes.byref(info)))
        self.a_layout, self.b_layout = info.a_layout, info.b_layout

    def execute(self, a_data, b_data, out):
        self._check_sizes(a_data, b_data, out)
        self.h.bind_stream()
        self.h.check(lib().tq_bs_contract_execute(
            self.ptr, ctypes.c_void_p(a_data.data_ptr()),
            ctypes.c_void_p(b_data.data_ptr()), ctypes.c_void_p(out.data_ptr())))

    def pack(self, a_data, b_data):
        self._check_sizes(a_data, b_data)
        self.h.bind_stream()
        self.h.check(lib().tq_bs_contract_pack(
            self.ptr, ctypes.c_void_p(a_data.data_ptr()), ctypes.c_void_p(b_data.data_ptr())))

    def multiply(self, a_data, b_data, out):
        self._check_sizes(a_data, b_data, out)
        self.h.bind_stream()
        self.h.check(lib().tq_bs_contract_multiply(
            self.ptr, ctypes.c_void_p(a_data.data_ptr()),
            ctypes.c_void_p(b_data.data_ptr()), ctypes.c_void_p(out.data_ptr())))

    def output_tensor(self, out):
        res = BlockSparseTensor.__new__(BlockSparseTensor)
        res.rank, res.dtype, res.device = self.out_rank, self.dtype, out.device
        res.dimension = self.out_dims.copy()
        if self.out_rank == 0:
            res.begin_points = np.zeros((0, 0), np.int64)
            res.end_points = np.zeros((0, 0), np.int64)
            res.index_tables = np.zeros((0, 0), np.int64)
        else:
            res.begin_points = self.out_begin.copy()
            res.end_points = self.out_end.copy()
            res.index_tables = self.out_strides.copy()
        res.data = out
        res._meta_changed()
        return res

    def capture(self, a_data, b_data, out):
        """Capture execute(a, b, out) into a HIP graph (tq_plan_capture)."""
        self._check_sizes(a_data, b_data, out)
        self.h.bind_stream()
        self.h.check(lib().tq_plan_capture(
            self.ptr, ctypes.c_void_p(a_data.data_ptr()), ctypes.c_void_p(b_data.data_ptr()),
            ctypes.c_void_p(out.data_ptr())))

    def graph_launch(self):
        self.h.bind_stream()
        self.h.check(lib().tq_plan_graph_launch(self.ptr))

    def destroy(self):
        """Frees the device plan now and drops it from the handle's cache."""
        cache = getattr(getattr(self, "h", None), "plans", None)
        if cache:
            for k in [k for k, v in cache.items() if v is self]:
                del cache[k]
        if getattr(self, "ptr", None):
            lib().tq_plan_destroy(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


class BatchPlan(ContractionPlan):
    """Several contraction plans joined into one work list
    (tq_batch_create). Plan i reads A at a_offsets[i] of the shared A buffer,
    B at b_offsets[i], and writes C at c_offsets[i]."""

    def __init__(self, handle, plans, a_offsets, b_offsets, c_offsets):
        self.h = handle
        self.dtype = plans[0].dtype
        self.precision = plans[0].precision
        arr = (ctypes.c_void_p * len(plans))(*[p.ptr for p in plans])
        ao = np.ascontiguousarray(a_offsets, np.int64)
        bo = np.ascontiguousarray(b_offsets, np.int64)
        co = np.ascontiguousarray(c_offsets, np.int64)
        ptr = ctypes.c_void_p()
        handle.check(lib().tq_batch_create(handle.ptr, len(plans), arr, _i64p(ao), _i64p(bo),
                                           _i64p(co), ctypes.byref(ptr)))
        self.ptr = ptr
        info = PlanInfo()
        handle.check(lib().tq_plan_get_info(ptr, ctypes.byref(info)))
        self.n_pairs = info.n_pairs
        self.n_matched = info.n_matched
        self.out_n_elem = info.out_n_elem
        self.flops = info.flops
        self.bytes = info.bytes
        self.out_rank = info.out_rank
        self.a_extent = max([int(o) + p.a_extent for o, p in zip(ao, plans)] or [0])
        self.b_extent = max([int(o) + p.b_extent for o, p in zip(bo, plans)] or [0])


def gemm_configs(precision="native", with_kind=False):
    """(bm, bn, bk) of every tile configuration, plus (wave_tile, min_waves)
    when with_kind (wave_tile 1: one wavefront per tile; min_waves: the
    wavefronts per SIMD its registers keep resident)."""
    L = lib()
    pr = _lib.PRECISIONS[precision]
    out = []
    for c in range(L.tq_gemm_config_count(pr)):
        bm, bn, bk, wt, mw = (ctypes.c_int32() for _ in range(5))
        check(L.tq_gemm_config_describe(pr, c, ctypes.byref(bm), ctypes.byref(bn),
                                        ctypes.byref(bk), ctypes.byref(wt), ctypes.byref(mw)))
        out.append((bm.value, bn.value, bk.value) + ((wt.value, mw.value) if with_kind else ()))
    return out


class DenseTensor:
    """Dense tensor (src/gpu/dense.cuh:52-604); column-major, leg 0 fastest."""

    def __init__(self, dimension, source_data=None, dtype=torch.float64, device=None):
        dim = np.asarray(dimension, dtype=np.int64).reshape(-1)
        self.rank = int(dim.size)
        self.dimension = dim
        self.index_table = index_table(dim)
        self.dtype = dtype
        dev = torch.device(device) if device is not None else (
            torch.device("cuda") if torch.cuda.is_available() else torch.device("cpu"))
        n = int(np.prod(dim)) if self.rank else 1
        if source_data is None:
            self.data = torch.zeros(n, dtype=dtype, device=dev)
        else:
            self.data = torch.as_tensor(source_data).reshape(-1)[:n].to(dev, dtype).clone()

    def query(self, indices):
        idx = np.asarray(indices, dtype=np.int64).reshape(-1)
        if idx.size != self.rank:
            raise TorqueError(_lib.TQ_ERROR_RANK_MISMATCH, "Rank does not match")
        if np.any(idx >= self.dimension) or np.any(idx < 0):
            raise TorqueError(_lib.TQ_ERROR_OUT_OF_RANGE, "Indices out of boundary")
        return float(self.data[int(np.sum(idx * self.index_table))])

    def modify(self, indices, number):
        idx = np.asarray(indices, dtype=np.int64).reshape(-1)
        if idx.size != self.rank:
            raise TorqueError(_lib.TQ_ERROR_RANK_MISMATCH, "Rank does not match")
        if np.any(idx >= self.dimension) or np.any(idx < 0):
            raise TorqueError(_lib.TQ_ERROR_OUT_OF_RANGE, "Indices out of boundary")
        self.data[int(np.sum(idx * self.index_table))] = number

    def initialize(self, source_data):
        n = self.data.numel()
        self.data.copy_(torch.as_tensor(source_data).reshape(-1)[:n].to(self.data))

    def to_number(self):
        assert self.rank == 0
        return float(self.data[0])

    def soft_transpose(self, permutation):
        perm = np.asarray(permutation, dtype=np.int64)
        if perm.size != self.rank:
            raise TorqueError(_lib.TQ_ERROR_RANK_MISMATCH,
                              "The number of permutation does not match the rank of tensor")
        self.index_table = self.index_table[perm]
        self.dimension = self.dimension[perm]

    def hard_transpose(self, permutation):
        """DenseTensor::hard_transpose (src/gpu/dense.cuh:561-599)."""
        perm = np.asarray(permutation, dtype=np.int32).reshape(-1)
        if perm.size != self.rank:
            raise TorqueError(_lib.TQ_ERROR_RANK_MISMATCH,
                              "The number of permutation does not match the rank of tensor")
        _require_gpu(self.data)
        h = default_handle(self.data.device.index)
        # undo a soft transpose: storage dims are the dims sorted by stride
        order = np.argsort(self.index_table, kind="stable")
        store_dims = np.ascontiguousarray(self.dimension[order], dtype=np.int64)
        inv = np.empty_like(order)
        inv[order] = np.arange(self.rank)
        # leg l of this tensor is storage leg inv[l]; out leg k = leg perm[k]
        sperm = np.ascontiguousarray(inv[perm], dtype=np.int32)
        out = torch.empty_like(self.data)
        h.check(lib().tq_permute(h.ptr, _DT[self.dtype], self.rank, _i64p(store_dims),
                                 None, _i32p(sperm), ctypes.c_void_p(self.data.data_ptr()),
                                 ctypes.c_void_p(out.data_ptr())))
        res = DenseTensor.__new__(DenseTensor)
        res.rank, res.dtype = self.rank, self.dtype
        res.dimension = self.dimension[perm.astype(np.int64)]
        res.index_table = index_table(res.dimension)
        res.data = out
        return res

    def contract(self, tensor, contracting_indices):
        """DenseTensor::contract (src/gpu/dense.cuh:439-555)."""
        legs = np.ascontiguousarray(np.asarray(contracting_indices, dtype=np.int32).reshape(-1, 2))
        for a_leg, b_leg in legs:
            if self.dimension[a_leg] != tensor.dimension[b_leg]:
                raise TorqueError(_lib.TQ_ERROR_DIMENSION_MISMATCH,
                                  "The dimensions from two tensors to be contracted do not match")
        _require_gpu(self.data)
        h = default_handle(self.data.device.index)
        a_free = [l for l in range(self.rank) if l not in legs[:, 0]]
        b_free = [l for l in range(tensor.rank) if l not in legs[:, 1]]
        new_dim = np.concatenate([self.dimension[a_free], tensor.dimension[b_free]]).astype(np.int64)
        n = int(np.prod(new_dim)) if new_dim.size else 1
        out = torch.empty(n, dtype=self.dtype, device=self.data.device)
        da = np.ascontiguousarray(self.dimension, np.int64)
        db = np.ascontiguousarray(tensor.dimension, np.int64)
        h.check(lib().tq_dense_contract(h.ptr, _DT[self.dtype], self.rank, _i64p(da),
                                        tensor.rank, _i64p(db), legs.shape[0], _i32p(legs),
                                        ctypes.c_void_p(self.data.data_ptr()),
                                        ctypes.c_void_p(tensor.data.data_ptr()),
                                        ctypes.c_void_p(out.data_ptr())))
        res = DenseTensor.__new__(DenseTensor)
        res.rank, res.dtype = int(new_dim.size), self.dtype
        res.dimension = new_dim
        res.index_table = index_table(new_dim)
        res.data = out
        return res


class SparseTensor:
    """Element-sparse tensor (src/gpu/sparse.cuh:84-441): `indices[i]` is the
    linear index of stored element i under `index_table`, `data[i]` its
    value. Indices and values live on the device (int64 and the dtype);
    contract and hard_transpose run in libtorque_hip.so."""

    def __init__(self, dimension, source_data=None, indices=None, index_table_=None,
                 dtype=torch.float64, device=None):
        dim = np.asarray(dimension, dtype=np.int64).reshape(-1)
        self.rank = int(dim.size)
        self.dimension = dim
        self.index_table = (index_table(dim) if index_table_ is None
                            else np.asarray(index_table_, dtype=np.int64).reshape(-1).copy())
        self.dtype = dtype
        dev = torch.device(device) if device is not None else (
            torch.device("cuda") if torch.cuda.is_available() else torch.device("cpu"))
        if source_data is None:
            # src/gpu/sparse.cuh:98-107: no elements, data = {0}
            self.indices = torch.zeros(0, dtype=torch.int64, device=dev)
            self.data = torch.zeros(1, dtype=dtype, device=dev)
        else:
            if self.rank > 0 and indices is None:
                raise TorqueError(_lib.TQ_ERROR_INVALID_VALUE, "Source data not allocated!")
            idx = torch.as_tensor(np.asarray(indices if indices is not None else [0],
                                             dtype=np.int64)).reshape(-1)
            n = idx.numel() if self.rank > 0 else 1
            self.indices = idx.to(dev) if self.rank > 0 else \
                torch.zeros(0, dtype=torch.int64, device=dev)
            self.data = torch.as_tensor(source_data).reshape(-1)[:n].to(dev, dtype).clone()

    @property
    def nnz(self):
        return int(self.indices.numel())

    def _desc(self):
        dims = np.ascontiguousarray(self.dimension, np.int64)
        table = np.ascontiguousarray(self.index_table, np.int64)
        self._keep = (dims, table)
        return _lib.SpDesc(self.rank, _i64p(dims), _i64p(table), self.nnz,
                           ctypes.c_void_p(self.indices.data_ptr() if self.nnz else 0),
                           ctypes.c_void_p(self.data.data_ptr() if self.nnz else 0))

    def initialize(self, source_data, indices):
        """src/gpu/sparse.cuh:174-187."""
        idx = torch.as_tensor(np.asarray(indices, dtype=np.int64)).reshape(-1)
        self.indices = idx.to(self.data.device)
        self.data = torch.as_tensor(source_data).reshape(-1)[:idx.numel()].to(
            self.data.device, self.dtype).clone()

    def _position(self, idx):
        if idx.size != self.rank:
            raise TorqueError(_lib.TQ_ERROR_RANK_MISMATCH, "Rank does not match")
        key = int(np.sum(idx * self.index_table))
        hit = torch.nonzero(self.indices == key).reshape(-1)
        return key, (int(hit[0]) if hit.numel() else -1)

    def query(self, indices):
        """src/gpu/sparse.cuh:240-266: the stored value, 0 if absent."""
        idx = np.asarray(indices, dtype=np.int64).reshape(-1)
        if self.rank == 0:
            if idx.size != 0:
                raise TorqueError(_lib.TQ_ERROR_RANK_MISMATCH, "Rank does not match")
            return float(self.data[0])
        _, pos = self._position(idx)
        return float(self.data[pos]) if pos >= 0 else 0.0

    def modify(self, indices, number):
        """src/gpu/sparse.cuh:191-236: overwrite, remove (number == 0, the
        last element moves into the hole) or append."""
        idx = np.asarray(indices, dtype=np.int64).reshape(-1)
        if self.rank == 0:
            if idx.size != 0:
                raise TorqueError(_lib.TQ_ERROR_RANK_MISMATCH, "Rank does not match")
            self.data[0] = number
            return
        key, pos = self._position(idx)
        if pos >= 0:
            if number == 0:
                last = self.nnz - 1
                self.data[pos] = self.data[last]
                self.indices[pos] = self.indices[last]
                self.indices = self.indices[:last].clone()
                self.data = self.data[:last].clone()
            else:
                self.data[pos] = number
        elif number != 0:
            # appended after the stored values (src/gpu/sparse.cuh:226-227
            # pushes onto a buffer that still holds the placeholder value of
            # a tensor built from its dimension alone)
            dev, old = self.data.device, self.nnz
            self.indices = torch.cat([self.indices, torch.tensor([key], dtype=torch.int64,
                                                                 device=dev)])
            self.data = torch.cat([self.data[:old],
                                   torch.tensor([number], dtype=self.dtype, device=dev)])

    def to_number(self):
        assert self.rank == 0
        return float(self.data[0])

    def soft_transpose(self, permutation):
        """src/gpu/sparse.cuh:380-392: permutes dimension and index table only."""
        perm = np.asarray(permutation, dtype=np.int64).reshape(-1)
        if perm.size != self.rank:
            raise TorqueError(_lib.TQ_ERROR_RANK_MISMATCH,
                              "The number of permutation does not match the rank of tensor")
        self.index_table = self.index_table[perm]
        self.dimension = self.dimension[perm]

    def hard_transpose(self, permutation):
        """src/gpu/sparse.cuh:398-432: same values in the same order, indices
        re-encoded under the compact table of the permuted dimension
        (tq_sparse_permute_indices; each stored index is decoded, where the
        reference decodes the storage position, src/gpu/sparse.cuh:420-422)."""
        perm = np.ascontiguousarray(np.asarray(permutation, dtype=np.int32).reshape(-1))
        if perm.size != self.rank:
            raise TorqueError(_lib.TQ_ERROR_RANK_MISMATCH,
                              "The number of permutation does not match the rank of tensor")
        _require_gpu(self.data)
        h = default_handle(self.data.device.index)
        dims = np.ascontiguousarray(self.dimension, np.int64)
        table = np.ascontiguousarray(self.index_table, np.int64)
        out = torch.empty_like(self.indices)
        h.check(lib().tq_sparse_permute_indices(
            h.ptr, self.rank, _i64p(dims), _i64p(table), _i32p(perm), self.nnz,
            ctypes.c_void_p(self.indices.data_ptr() if self.nnz else 0),
            ctypes.c_void_p(out.data_ptr() if self.nnz else 0)))
        res = SparseTensor.__new__(SparseTensor)
        res.rank, res.dtype = self.rank, self.dtype
        res.dimension = self.dimension[perm.astype(np.int64)]
        res.index_table = index_table(res.dimension)
        res.indices = out
        res.data = self.data.clone()
        return res

    def contract(self, tensor, contracting_indices):
        """SparseTensor::contract (src/gpu/sparse.cuh:270-375) through
        tq_sparse_contract: every output index that receives a product,
        ascending, with the summed value."""
        legs = np.ascontiguousarray(np.asarray(contracting_indices, dtype=np.int32).reshape(-1, 2))
        for a_leg, b_leg in legs:
            if self.dimension[a_leg] != tensor.dimension[b_leg]:
                raise TorqueError(_lib.TQ_ERROR_DIMENSION_MISMATCH,
                                  "The dimensions from two tensors to be contracted do not match")
        _require_gpu(self.data)
        _require_gpu(tensor.data)
        h = default_handle(self.data.device.index)
        out_rank = self.rank + tensor.rank - 2 * legs.shape[0]
        out_dims = np.zeros(max(out_rank, 1), np.int64)
        nnz = ctypes.c_int64()
        pi, pv = ctypes.c_void_p(), ctypes.c_void_p()
        da, db = self._desc(), tensor._desc()
        h.check(lib().tq_sparse_contract(h.ptr, _DT[self.dtype], ctypes.byref(da),
                                         ctypes.byref(db), legs.shape[0], _i32p(legs),
                                         _i64p(out_dims), ctypes.byref(nnz),
                                         ctypes.byref(pi), ctypes.byref(pv)))
        n = nnz.value
        dev = self.data.device
        res = SparseTensor.__new__(SparseTensor)
        res.rank, res.dtype = out_rank, self.dtype
        res.dimension = out_dims[:out_rank].copy()
        res.index_table = index_table(res.dimension)
        res.indices = torch.empty(n if out_rank > 0 else 0, dtype=torch.int64, device=dev)
        res.data = torch.empty(max(n, 1) if out_rank == 0 else n, dtype=self.dtype, device=dev)
        try:
            if out_rank > 0 and n > 0:
                h.check(lib().tq_memcpy(h.ptr, ctypes.c_void_p(res.indices.data_ptr()), pi,
                                        8 * n, 2))
            if res.data.numel():
                h.check(lib().tq_memcpy(h.ptr, ctypes.c_void_p(res.data.data_ptr()), pv,
                                        res.data.element_size() * res.data.numel(), 2))
            h.synchronize()
        finally:
            lib().tq_free(h.ptr, pi)
            lib().tq_free(h.ptr, pv)
        return res


def dot_f64(x, y):
    h = default_handle(x.device.index)
    r = ctypes.c_double()
    h.check(lib().tq_dot_f64(h.ptr, x.numel(), ctypes.c_void_p(x.data_ptr()),
                             ctypes.c_void_p(y.data_ptr()), ctypes.byref(r)))
    return r.value


def nrm2_f64(x):
    h = default_handle(x.device.index)
    r = ctypes.c_double()
    h.check(lib().tq_nrm2_f64(h.ptr, x.numel(), ctypes.c_void_p(x.data_ptr()),
                              ctypes.byref(r)))
    return r.value


def scale_f64(x, alpha):
    h = default_handle(x.device.index)
    h.check(lib().tq_scale_f64(h.ptr, x.numel(), ctypes.c_void_p(x.data_ptr()),
                               ctypes.c_double(alpha)))


def normalize_f64(x, norm=None):
    """x <- x / ||x|| on the device (tq_normalize_f64); the FP64 norm is
    written to the one-element device tensor `norm` (allocated when None),
    which is returned."""
    h = default_handle(x.device.index)
    if norm is None:
        norm = torch.empty(1, dtype=torch.float64, device=x.device)
    h.check(lib().tq_normalize_f64(h.ptr, x.numel(), ctypes.c_void_p(x.data_ptr()),
                                   ctypes.c_void_p(norm.data_ptr())))
    return norm


def normalize_f64_noalloc(x, norm):
    """x <- x / ||x|| with the norm written to the one-element device tensor
    `norm`, allocating nothing (tq_normalize_f64_noalloc): usable while a
    handle captures a graph."""
    h = default_handle(x.device.index)
    h.check(lib().tq_normalize_f64_noalloc(h.ptr, x.numel(), ctypes.c_void_p(x.data_ptr()),
                                           ctypes.c_void_p(norm.data_ptr())))
