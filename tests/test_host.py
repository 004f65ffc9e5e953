"""Host-side logic of the Python mirror (CPU torch data, no kernels):
element access semantics of the reference and its error behaviour."""
import numpy as np
import pytest
import torch

from torque_amd import BlockSparseTensor, DenseTensor, TorqueError, index_table, nest_sum


def test_init_modify_append_semantics():
    # test/gpu/block_sparse_test.cu:10-33
    t = BlockSparseTensor([4, 3], [1, 2, 3, 4], [[1, 1]], [[2, 2]], device="cpu")
    assert t.query([1, 1]) == 1
    assert t.query([2, 2]) == 4
    assert t.query([3, 2]) == 0
    t.modify([2, 2], 0)
    assert t.query([2, 2]) == 0
    t.modify([3, 2], 5)
    assert t.query([3, 2]) == 5
    t.append_block([1, 2, 3, 4], [0, 0], [1, 1], [1, 2])
    assert t.query([1, 1]) == 5  # overlapping blocks sum


def test_rank3_append_and_scalar():
    # test/gpu/block_sparse_test.cu:62-74, 93-102
    t = BlockSparseTensor([2, 2, 3], device="cpu")
    assert t.query([1, 1, 0]) == 0
    t.append_block(np.arange(1, 13), [0, 0, 0], [1, 1, 2], [1, 2, 4])
    assert t.query([0, 1, 2]) == 11
    s = BlockSparseTensor([], device="cpu")
    assert s.rank == 0 and s.query([]) == 0
    s.modify([], 1)
    assert s.query([]) == 1


def test_metadata_derivation():
    t = BlockSparseTensor([5, 6], np.zeros(22), [[0, 0], [1, 2]], [[1, 2], [4, 5]], device="cpu")
    assert t.block_n_elem.tolist() == [6, 16]
    assert t.block_offsets.tolist() == [0, 6]
    assert t.index_tables.tolist() == [[1, 2], [1, 4]]
    assert index_table([3, 4, 5]).tolist() == [1, 3, 12]
    assert nest_sum([2, 3, 4]).tolist() == [0, 2, 5]


def test_errors_match_reference_messages():
    t = BlockSparseTensor([4, 3], device="cpu")
    with pytest.raises(TorqueError, match="^Error: Rank does not match"):
        t.query([1])
    with pytest.raises(TorqueError, match="^Error: Indices out of boundary"):
        t.query([4, 0])
    with pytest.raises(TorqueError, match="number of permutation"):
        t.soft_transpose([0])
    d = DenseTensor([2, 2], device="cpu")
    with pytest.raises(TorqueError, match="do not match"):
        d.contract(DenseTensor([3], device="cpu"), [[0, 0]])


def test_no_host_fallback():
    t = BlockSparseTensor([2], [1, 2], [[0]], [[1]], device="cpu")
    with pytest.raises(TorqueError, match="GPU"):
        t.contract(t, [[0, 0]])
    with pytest.raises(TorqueError, match="GPU"):
        t.hard_transpose([0])


def test_dense_host_access():
    d = DenseTensor([4, 3], device="cpu")
    d.initialize(torch.arange(1, 13, dtype=torch.float64))
    assert d.query([1, 2]) == 10


def test_short_source_data_is_rejected():
    # the kernels read every element of every block: data shorter than the
    # blocks need is an error, in the constructor and in append_block
    with pytest.raises(TorqueError, match="the blocks need 22"):
        BlockSparseTensor([5, 6], np.zeros(20), [[0, 0], [1, 2]], [[1, 2], [4, 5]],
                          device="cpu")
    t = BlockSparseTensor([4, 3], [1, 2, 3, 4], [[1, 1]], [[2, 2]], device="cpu")
    with pytest.raises(TorqueError, match="the block needs 4"):
        t.append_block([1, 2, 3], [0, 0], [1, 1], [1, 2])


def test_block_tensor_file_round_trip(tmp_path):
    # the binary block-tensor file keeps metadata and values bit for bit, in
    # every dtype, including a tensor without blocks
    rng = np.random.default_rng(3)
    for dtype in (torch.float64, torch.float32, torch.float16):
        t = BlockSparseTensor([5, 6], rng.standard_normal(22), [[0, 0], [1, 2]],
                              [[1, 2], [4, 5]], index_tables=[[3, 1], [1, 4]], dtype=dtype,
                              device="cpu")
        t.save(tmp_path / "t.tqbs")
        u = BlockSparseTensor.load(tmp_path / "t.tqbs", device="cpu")
        assert u.dtype == dtype
        for k in ("dimension", "begin_points", "end_points", "index_tables", "block_offsets"):
            assert np.array_equal(getattr(u, k), getattr(t, k)), k
        assert torch.equal(u.data, t.data)
    e = BlockSparseTensor([3, 4], device="cpu")
    e.save(tmp_path / "e.tqbs")
    assert BlockSparseTensor.load(tmp_path / "e.tqbs", device="cpu").n_blocks == 0
    (tmp_path / "bad.tqbs").write_bytes(b"XXXX")
    with pytest.raises(TorqueError, match="block-tensor"):
        BlockSparseTensor.load(tmp_path / "bad.tqbs", device="cpu")
