// torque/gpu/block_sparse.h -- drop-in torque::gpu::BlockSparseTensor<T> over
// the C-ABI of libtorque_hip.so (MI355X, gfx950).
//
// Follows src/gpu/block_sparse.cuh:142-1181 of the reference: the same public
// members, constructors, append_block(s) / modify / query / to_number /
// contract / hard_transpose, the same "Error: ..." exceptions for invalid
// arguments and the same output layout of contract (one block per matched
// pair, A block major, B block ascending, column-major [A free, B free]).
// Differences a caller sees: contract takes a tq_handle_t where the
// reference takes a cublasHandle_t; device failures throw Error instead of
// exiting; T is float, double or torque::half (IEEE binary16 bits, the
// reference's half).
#pragma once

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

#include "torque/error.h"
#include "torque/half.h"
#include "torque/index.h"

namespace torque {
namespace gpu {

enum MemcpyKind { HostToDevice = 0, DeviceToHost = 1, DeviceToDevice = 2 };

template <typename T>
constexpr tq_dtype dtype_of() {
  static_assert(std::is_same<T, float>::value || std::is_same<T, double>::value ||
                    std::is_same<T, half>::value,
                "torque::gpu tensors hold float, double or torque::half");
  return std::is_same<T, double>::value ? TQ_F64
         : std::is_same<T, half>::value ? TQ_F16
                                        : TQ_F32;
}

template <typename T>
class BlockSparseTensor {
 public:
  uword rank = 0;
  uvec dimension;
  umat blocks_dimension;
  umat index_tables;
  umat begin_points;
  umat end_points;
  uvec block_n_elem;
  uvec block_offsets;

  // src/gpu/block_sparse.cuh:180-186: no blocks, one zeroed element
  explicit BlockSparseTensor(const uvec &dim) : rank(dim.n_elem()), dimension(dim) {
    allocate(1);
    check(tq_memset_zero(default_handle(), data, sizeof(T)), default_handle());
  }

  // src/gpu/block_sparse.cuh:189-234
  BlockSparseTensor(const T *source_data, const umat &begin, const umat &end,
                    const uvec &total_dimension, MemcpyKind kind = HostToDevice)
      : rank(total_dimension.n_elem()), dimension(total_dimension) {
    if (!source_data) throw Error("Source data not allocated!");
    set_blocks(begin, end, nullptr);
    const uword n = total();
    allocate(std::max<uword>(n, 1));
    copy_in(data, source_data, std::max<uword>(n, 1), kind);
  }

  // one block given by uvec begin/end points (the reference's callers pass
  // arma::uvec, which converts to a one-column umat)
  BlockSparseTensor(const T *source_data, const uvec &begin, const uvec &end,
                    const uvec &total_dimension, MemcpyKind kind = HostToDevice)
      : BlockSparseTensor(source_data, umat::column(begin), umat::column(end),
                          total_dimension, kind) {}

  // src/gpu/block_sparse.cuh:238-279 (explicit index tables)
  BlockSparseTensor(const T *source_data, const umat &begin, const umat &end,
                    const uvec &total_dimension, const umat &tables,
                    MemcpyKind kind = DeviceToDevice)
      : rank(total_dimension.n_elem()), dimension(total_dimension) {
    if (!source_data) throw Error("Source data not allocated!");
    set_blocks(begin, end, &tables);
    const uword n = total();
    allocate(std::max<uword>(n, 1));
    copy_in(data, source_data, std::max<uword>(n, 1), kind);
  }

  BlockSparseTensor(const BlockSparseTensor &o)
      : rank(o.rank), dimension(o.dimension), blocks_dimension(o.blocks_dimension),
        index_tables(o.index_tables), begin_points(o.begin_points),
        end_points(o.end_points), block_n_elem(o.block_n_elem),
        block_offsets(o.block_offsets) {
    const uword n = std::max<uword>(total(), 1);
    allocate(n);
    copy_in(data, o.data, n, DeviceToDevice);
  }

  BlockSparseTensor(BlockSparseTensor &&o) noexcept { swap(*this, o); }
  BlockSparseTensor &operator=(BlockSparseTensor o) noexcept {
    swap(*this, o);
    return *this;
  }
  friend void swap(BlockSparseTensor &a, BlockSparseTensor &b) noexcept {
    using std::swap;
    swap(a.data, b.data);
    swap(a.rank, b.rank);
    swap(a.dimension, b.dimension);
    swap(a.blocks_dimension, b.blocks_dimension);
    swap(a.index_tables, b.index_tables);
    swap(a.begin_points, b.begin_points);
    swap(a.end_points, b.end_points);
    swap(a.block_n_elem, b.block_n_elem);
    swap(a.block_offsets, b.block_offsets);
  }
  ~BlockSparseTensor() {
    if (data) (void)tq_free(default_handle(), data);
  }

  const T *device_data() const { return data; }
  T *device_data() { return data; }
  uword n_blocks() const { return begin_points.n_cols(); }

  T to_number() const {
    T v;
    check(tq_memcpy(default_handle(), &v, data, sizeof(T), DeviceToHost), default_handle());
    return v;
  }

  // src/gpu/block_sparse.cuh:376-410
  void append_block(const T *source_data, const uvec &begin, const uvec &end,
                    const uvec &index_table, MemcpyKind kind = HostToDevice) {
    umat b(rank, 1), e(rank, 1), t(rank, 1);
    for (uword l = 0; l < rank; ++l) {
      b(l, 0) = begin[l];
      e(l, 0) = end[l];
      t(l, 0) = index_table[l];
    }
    append_blocks(source_data, b, e, t, kind);
  }

  // src/gpu/block_sparse.cuh:412-446
  void append_blocks(const T *source_data, const umat &begin, const umat &end,
                     const umat &tables, MemcpyKind kind = HostToDevice) {
    const uword old_n = total();
    const bool had_blocks = n_blocks() > 0;
    uword add = 0;
    for (std::size_t j = 0; j < begin.n_cols(); ++j) {
      uword n = 1;
      for (uword l = 0; l < rank; ++l) n *= end(l, j) - begin(l, j) + 1;
      begin_points.append_col(begin.col(j));
      end_points.append_col(end.col(j));
      index_tables.append_col(tables.col(j));
      add += n;
    }
    derive();
    T *fresh = nullptr;
    check(tq_malloc(default_handle(), reinterpret_cast<void **>(&fresh),
                    sizeof(T) * std::max<uword>(old_n + add, 1)),
          default_handle());
    if (had_blocks && old_n) copy_in(fresh, data, old_n, DeviceToDevice);
    if (add) copy_in(fresh + old_n, source_data, add, kind);
    (void)tq_free(default_handle(), data);
    data = fresh;
  }

  // src/gpu/block_sparse.cuh:452-505
  void modify(const uvec &indices, const T number) {
    check_indices(indices);
    if (rank == 0) {
      copy_in(data, &number, 1, HostToDevice);
      return;
    }
    const std::vector<uword> cover = covering(indices);
    if (!cover.empty()) {
      copy_in(data + address(cover[0], indices), &number, 1, HostToDevice);
      const T zero = 0;
      for (std::size_t q = 1; q < cover.size(); ++q)
        copy_in(data + address(cover[q], indices), &zero, 1, HostToDevice);
    } else if (number != 0) {
      append_block(&number, indices, indices, uvec(rank, 0));
    }
  }

  // src/gpu/block_sparse.cuh:510-552: sums every covering block
  T query(const uvec &indices) const {
    check_indices(indices);
    if (rank == 0) return to_number();
    T acc = 0;
    for (uword b : covering(indices)) {
      T v;
      check(tq_memcpy(default_handle(), &v, data + address(b, indices), sizeof(T),
                      DeviceToHost),
            default_handle());
      acc += v;
    }
    return acc;
  }

  // src/gpu/block_sparse.cuh:558-1026 (device matching, permute-to-GEMM
  // layout, one grouped GEMM launch; the output blocks follow the reference).
  // Plans come from the handle's cache, so a repeated contraction of the
  // same block structure (a CTMRG step) skips planning.
  BlockSparseTensor contract(tq_handle_t handle, const BlockSparseTensor &tensor,
                             const umat &contracting_indices,
                             tq_precision precision = TQ_PREC_NATIVE) const {
    return contract_impl(handle, tensor, contracting_indices, precision, false);
  }

  // As contract, with one output block per distinct output block (the sum
  // of the pairs writing it) instead of one per matched pair
  // (tq_bs_contract_plan_merged); the dense image is the same.
  BlockSparseTensor contract_merged(tq_handle_t handle, const BlockSparseTensor &tensor,
                                    const umat &contracting_indices,
                                    tq_precision precision = TQ_PREC_NATIVE) const {
    return contract_impl(handle, tensor, contracting_indices, precision, true);
  }

 private:
  BlockSparseTensor contract_impl(tq_handle_t handle, const BlockSparseTensor &tensor,
                                  const umat &contracting_indices, tq_precision precision,
                                  bool merged) const {
    const std::size_t n_legs = contracting_indices.n_rows();
    std::vector<int32_t> legs(2 * n_legs);
    for (std::size_t k = 0; k < n_legs; ++k) {
      legs[2 * k] = (int32_t)contracting_indices(k, 0);
      legs[2 * k + 1] = (int32_t)contracting_indices(k, 1);
    }
    for (std::size_t k = 0; k < n_legs; ++k)
      if (legs[2 * k] >= (int32_t)rank || legs[2 * k + 1] >= (int32_t)tensor.rank ||
          dimension[legs[2 * k]] != tensor.dimension[legs[2 * k + 1]])
        throw Error("The dimensions from two tensors to be contracted do not match");
    Desc da(*this), db(tensor);
    tq_plan_t plan = nullptr;  // owned by the handle's cache
    check(tq_bs_contract_plan_cached(handle, &da.d, &db.d, (int32_t)n_legs, legs.data(),
                                     dtype_of<T>(), precision, merged ? 1 : 0, &plan),
          handle);
    tq_plan_info info;
    check(tq_plan_get_info(plan, &info), handle);
    uvec new_dim(info.out_rank);
    std::vector<int64_t> dims(info.out_rank);
    check(tq_plan_get_output(plan, dims.data(), nullptr, nullptr, nullptr, nullptr), handle);
    for (int l = 0; l < info.out_rank; ++l) new_dim[l] = (uword)dims[l];
    if (info.n_pairs == 0) return BlockSparseTensor(new_dim);
    const uword ro = info.out_rank, np = info.n_pairs;
    std::vector<int64_t> b(np * ro), e(np * ro), t(np * ro);
    check(tq_plan_get_output(plan, nullptr, b.data(), e.data(), t.data(), nullptr), handle);
    BlockSparseTensor out;
    out.rank = ro;
    out.dimension = new_dim;
    out.allocate(std::max<int64_t>(info.out_n_elem, 1));
    check(tq_bs_contract_execute(plan, data, tensor.data, out.data), handle);
    check(tq_synchronize(handle), handle);
    if (ro > 0) {
      umat bb(ro, np), ee(ro, np), tt(ro, np);
      for (uword p = 0; p < np; ++p)
        for (uword l = 0; l < ro; ++l) {
          bb(l, p) = (uword)b[p * ro + l];
          ee(l, p) = (uword)e[p * ro + l];
          tt(l, p) = (uword)t[p * ro + l];
        }
      out.begin_points = bb;
      out.end_points = ee;
      out.index_tables = tt;
      out.derive();
    }
    return out;
  }

 public:
  // src/gpu/block_sparse.cuh:1031-1122
  BlockSparseTensor hard_transpose(const uvec &permutation) const {
    if (permutation.n_elem() != rank)
      throw Error("The number of permutation does not match the rank of tensor");
    std::vector<int32_t> perm(rank);
    std::vector<bool> seen(rank, false);
    for (uword k = 0; k < rank; ++k) {
      if (permutation[k] >= rank || seen[permutation[k]])
        throw Error("The number of permutation does not match the rank of tensor");
      seen[permutation[k]] = true;
      perm[k] = (int32_t)permutation[k];
    }
    BlockSparseTensor out;
    out.rank = rank;
    out.dimension = uvec(rank);
    for (uword k = 0; k < rank; ++k) out.dimension[k] = dimension[perm[k]];
    const uword nb = n_blocks();
    umat bb(rank, nb), ee(rank, nb), tt(rank, nb);
    for (uword j = 0; j < nb; ++j) {
      uword acc = 1;
      for (uword k = 0; k < rank; ++k) {
        bb(k, j) = begin_points(perm[k], j);
        ee(k, j) = end_points(perm[k], j);
        tt(k, j) = acc;
        acc *= ee(k, j) - bb(k, j) + 1;
      }
    }
    out.begin_points = nb ? bb : umat();
    out.end_points = nb ? ee : umat();
    out.index_tables = nb ? tt : umat();
    out.derive();
    out.allocate(std::max<uword>(out.total(), 1));
    Desc d(*this);
    check(tq_bs_permute(default_handle(), dtype_of<T>(), &d.d, perm.data(), data, out.data,
                        nullptr),
          default_handle());
    check(tq_synchronize(default_handle()), default_handle());
    return out;
  }

  // src/gpu/block_sparse.cuh:1124-1176 through tq_bs_slice: every block
  // split into prod(divisor) sub-blocks, each holding the values of its
  // region (the reference keeps the old data buffer under the new blocks)
  BlockSparseTensor slice(const uvec &divisor) const {
    if (divisor.n_elem() != rank) throw Error("Rank does not match");
    uword n_sub = 1;
    for (uword l = 0; l < rank; ++l) n_sub *= divisor[l];
    const uword nb = n_blocks() * n_sub;
    std::vector<int64_t> div(divisor.vec().begin(), divisor.vec().end());
    std::vector<int64_t> bb(nb * rank), ee(nb * rank);
    BlockSparseTensor out;
    out.rank = rank;
    out.dimension = dimension;
    out.allocate(std::max<uword>(total(), 1));
    Desc d(*this);
    check(tq_bs_slice(default_handle(), dtype_of<T>(), &d.d, div.data(), data, out.data,
                      bb.data(), ee.data()),
          default_handle());
    umat b(rank, nb), e(rank, nb);
    for (uword j = 0; j < nb; ++j)
      for (uword l = 0; l < rank; ++l) {
        b(l, j) = (uword)bb[j * rank + l];
        e(l, j) = (uword)ee[j * rank + l];
      }
    out.set_blocks(b, e, nullptr);
    return out;
  }

 private:
  T *data = nullptr;

  BlockSparseTensor() = default;

  // Block-major metadata of this tensor as the C-ABI descriptor.
  struct Desc {
    tq_bs_desc d;
    std::vector<int64_t> dims, offs;
    explicit Desc(const BlockSparseTensor &t) {
      dims.assign(t.dimension.vec().begin(), t.dimension.vec().end());
      offs.assign(t.block_offsets.vec().begin(), t.block_offsets.vec().end());
      d.rank = (int32_t)t.rank;
      d.n_blocks = (int64_t)t.n_blocks();
      d.dims = dims.data();
      d.begin = reinterpret_cast<const int64_t *>(t.begin_points.data());
      d.end = reinterpret_cast<const int64_t *>(t.end_points.data());
      d.strides = reinterpret_cast<const int64_t *>(t.index_tables.data());
      d.offsets = offs.data();
    }
  };

  void allocate(uword n) {
    check(tq_malloc(default_handle(), reinterpret_cast<void **>(&data), sizeof(T) * n),
          default_handle());
  }
  static void copy_in(T *dst, const T *src, uword n, MemcpyKind kind) {
    check(tq_memcpy(default_handle(), dst, src, sizeof(T) * n, kind), default_handle());
  }
  uword total() const {
    uword s = 0;
    for (std::size_t b = 0; b < block_n_elem.n_elem(); ++b) s += block_n_elem[b];
    return s;
  }
  void set_blocks(const umat &begin, const umat &end, const umat *tables) {
    if (rank == 0) return;
    begin_points = begin;
    end_points = end;
    if (tables) {
      index_tables = *tables;
    } else {
      index_tables = umat(rank, begin.n_cols());
      for (std::size_t j = 0; j < begin.n_cols(); ++j) {
        uword acc = 1;
        for (uword l = 0; l < rank; ++l) {
          index_tables(l, j) = acc;
          acc *= end(l, j) - begin(l, j) + 1;
        }
      }
    }
    derive();
  }
  // blocks_dimension, block_n_elem, block_offsets from begin/end points
  void derive() {
    const std::size_t nb = begin_points.n_cols();
    blocks_dimension = umat(rank, nb);
    block_n_elem = uvec(nb);
    for (std::size_t j = 0; j < nb; ++j) {
      uword n = 1;
      for (uword l = 0; l < rank; ++l) {
        blocks_dimension(l, j) = end_points(l, j) - begin_points(l, j) + 1;
        n *= blocks_dimension(l, j);
      }
      block_n_elem[j] = n;
    }
    block_offsets = util::nest_sum(block_n_elem);
  }
  void check_indices(const uvec &idx) const {
    if (idx.n_elem() != rank) throw Error("Rank does not match");
    for (uword l = 0; l < rank; ++l)
      if (idx[l] >= dimension[l]) throw Error("Indices out of boundary");
  }
  std::vector<uword> covering(const uvec &idx) const {
    std::vector<uword> out;
    for (uword j = 0; j < n_blocks(); ++j) {
      bool in = true;
      for (uword l = 0; l < rank && in; ++l)
        in = idx[l] >= begin_points(l, j) && idx[l] <= end_points(l, j);
      if (in) out.push_back(j);
    }
    return out;
  }
  uword address(uword b, const uvec &idx) const {
    uword a = block_offsets[b];
    for (uword l = 0; l < rank; ++l) a += (idx[l] - begin_points(l, b)) * index_tables(l, b);
    return a;
  }
};

static_assert(sizeof(uword) == 8, "block-tensor files hold 64-bit metadata");

// Binary block-tensor file, the format torque_amd.BlockSparseTensor.save
// writes: "TQBS", little-endian u32 version (1), u32 dtype (tq_dtype), u32
// rank, u64 n_blocks, u64 dims[rank], u64 begin, end and index tables
// [n_blocks][rank] (block-major), u64 n_elem, then the n_elem values.
template <typename T>
void save(const BlockSparseTensor<T> &t, const std::string &path) {
  uword n = 0;
  for (uword j = 0; j < t.n_blocks(); ++j) n += t.block_n_elem[j];
  n = std::max<uword>(n, 1);
  std::vector<T> host(n);
  check(tq_memcpy(default_handle(), host.data(), t.device_data(), sizeof(T) * n, DeviceToHost),
        default_handle());
  std::FILE *f = std::fopen(path.c_str(), "wb");
  if (!f) throw Error("cannot write " + path);
  const uint32_t head[3] = {1, (uint32_t)dtype_of<T>(), (uint32_t)t.rank};
  const uint64_t nb = t.n_blocks(), ne = n;
  bool ok = std::fwrite("TQBS", 1, 4, f) == 4 && std::fwrite(head, 4, 3, f) == 3 &&
            std::fwrite(&nb, 8, 1, f) == 1 &&
            std::fwrite(t.dimension.data(), 8, t.rank, f) == t.rank;
  for (const umat *m : {&t.begin_points, &t.end_points, &t.index_tables})
    ok = ok && std::fwrite(m->data(), 8, nb * t.rank, f) == nb * t.rank;
  ok = ok && std::fwrite(&ne, 8, 1, f) == 1 && std::fwrite(host.data(), sizeof(T), n, f) == n;
  std::fclose(f);
  if (!ok) throw Error("cannot write " + path);
}

template <typename T>
BlockSparseTensor<T> load_block_sparse(const std::string &path) {
  std::FILE *f = std::fopen(path.c_str(), "rb");
  if (!f) throw Error("cannot read " + path);
  char magic[4] = {0, 0, 0, 0};
  uint32_t head[3] = {0, 0, 0};
  uint64_t nb = 0, ne = 0;
  bool ok = std::fread(magic, 1, 4, f) == 4 && std::memcmp(magic, "TQBS", 4) == 0 &&
            std::fread(head, 4, 3, f) == 3 && head[0] == 1 &&
            head[1] == (uint32_t)dtype_of<T>() && std::fread(&nb, 8, 1, f) == 1;
  const uword rank = ok ? head[2] : 0;
  uvec dims(rank);
  umat b(rank, nb), e(rank, nb), tb(rank, nb);
  ok = ok && std::fread(dims.data(), 8, rank, f) == rank;
  for (umat *m : {&b, &e, &tb})
    ok = ok && std::fread(m->data(), 8, nb * rank, f) == nb * rank;
  ok = ok && std::fread(&ne, 8, 1, f) == 1 && ne >= 1 && ne < (uint64_t(1) << 40);
  std::vector<T> host(ok ? ne : 0);
  ok = ok && std::fread(host.data(), sizeof(T), ne, f) == ne;
  std::fclose(f);
  if (!ok) throw Error("not a block-tensor file of this type: " + path);
  if (nb == 0) {
    BlockSparseTensor<T> t(dims);
    t.modify(uvec(rank, 0), host[0]);
    return t;
  }
  return BlockSparseTensor<T>(host.data(), b, e, dims, tb, HostToDevice);
}

}  // namespace gpu
}  // namespace torque
