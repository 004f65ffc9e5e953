// torque/gpu/dense.h -- drop-in torque::gpu::DenseTensor<T> over the C-ABI
// (src/gpu/dense.cuh:52-604 of the reference): column-major dense tensor with
// query / modify / initialize / to_number, hard_transpose (the cuTT call
// becomes tq_permute) and contract (permute + cuBLAS becomes
// tq_dense_contract).
#pragma once

#include <vector>

#include "torque/gpu/block_sparse.h"

namespace torque {
namespace gpu {

template <typename T>
class DenseTensor {
 public:
  uword rank = 0;
  uvec dimension;
  uvec index_table;

  // src/gpu/dense.cuh:66-80: zero-filled (one element for rank 0)
  explicit DenseTensor(const uvec &dim) : rank(dim.n_elem()), dimension(dim) {
    index_table = util::generate_index_table(dim);
    allocate();
    check(tq_memset_zero(default_handle(), data, sizeof(T) * count()), default_handle());
  }

  // src/gpu/dense.cuh:83-109
  DenseTensor(const T *source_data, const uvec &dim) : rank(dim.n_elem()), dimension(dim) {
    if (!source_data) throw Error("Source data not allocated!");
    index_table = util::generate_index_table(dim);
    allocate();
    check(tq_memcpy(default_handle(), data, source_data, sizeof(T) * count(), HostToDevice),
          default_handle());
  }

  DenseTensor(const DenseTensor &o)
      : rank(o.rank), dimension(o.dimension), index_table(o.index_table) {
    allocate();
    check(tq_memcpy(default_handle(), data, o.data, sizeof(T) * count(), DeviceToDevice),
          default_handle());
  }
  DenseTensor(DenseTensor &&o) noexcept { swap(*this, o); }
  DenseTensor &operator=(DenseTensor o) noexcept {
    swap(*this, o);
    return *this;
  }
  friend void swap(DenseTensor &a, DenseTensor &b) noexcept {
    using std::swap;
    swap(a.rank, b.rank);
    swap(a.dimension, b.dimension);
    swap(a.index_table, b.index_table);
    swap(a.data, b.data);
  }
  ~DenseTensor() {
    if (data) (void)tq_free(default_handle(), data);
  }

  const T *device_data() const { return data; }

  T to_number() const {
    T v;
    check(tq_memcpy(default_handle(), &v, data, sizeof(T), DeviceToHost), default_handle());
    return v;
  }

  void initialize(const T *source_data) {
    if (!source_data) throw Error("Source data not allocated!");
    check(tq_memcpy(default_handle(), data, source_data, sizeof(T) * count(), HostToDevice),
          default_handle());
  }

  void modify(const uvec &indices, const T number) {
    check(tq_memcpy(default_handle(), data + offset(indices), &number, sizeof(T),
                    HostToDevice),
          default_handle());
  }

  T query(const uvec &indices) const {
    T v;
    check(tq_memcpy(default_handle(), &v, data + offset(indices), sizeof(T), DeviceToHost),
          default_handle());
    return v;
  }

  // src/gpu/dense.cuh:561-599
  DenseTensor hard_transpose(const uvec &permutation) const {
    if (permutation.n_elem() != rank)
      throw Error("The number of permutation does not match the rank of tensor");
    std::vector<int64_t> dims(dimension.vec().begin(), dimension.vec().end());
    std::vector<int32_t> perm(rank);
    std::vector<bool> seen(rank, false);
    uvec nd(rank);
    // every leg exactly once, checked before any index or allocation uses it
    for (uword k = 0; k < rank; ++k) {
      if (permutation[k] >= rank || seen[permutation[k]])
        throw Error("The number of permutation does not match the rank of tensor");
      seen[permutation[k]] = true;
    }
    for (uword k = 0; k < rank; ++k) {
      perm[k] = (int32_t)permutation[k];
      nd[k] = dimension[permutation[k]];
    }
    DenseTensor out(nd, 0);
    check(tq_permute(default_handle(), dtype_of<T>(), (int32_t)rank, dims.data(), nullptr,
                     perm.data(), data, out.data),
          default_handle());
    check(tq_synchronize(default_handle()), default_handle());
    return out;
  }

  // src/gpu/dense.cuh:439-555
  DenseTensor contract(tq_handle_t handle, const DenseTensor &tensor,
                       const umat &contracting_indices) const {
    const std::size_t n_legs = contracting_indices.n_rows();
    std::vector<int32_t> legs(2 * n_legs);
    std::vector<bool> a_c(rank, false), b_c(tensor.rank, false);
    for (std::size_t k = 0; k < n_legs; ++k) {
      legs[2 * k] = (int32_t)contracting_indices(k, 0);
      legs[2 * k + 1] = (int32_t)contracting_indices(k, 1);
      if (contracting_indices(k, 0) >= rank || contracting_indices(k, 1) >= tensor.rank ||
          dimension[legs[2 * k]] != tensor.dimension[legs[2 * k + 1]])
        throw Error("The dimensions from two tensors to be contracted do not match");
      a_c[legs[2 * k]] = true;
      b_c[legs[2 * k + 1]] = true;
    }
    std::vector<uword> nd;
    for (uword l = 0; l < rank; ++l)
      if (!a_c[l]) nd.push_back(dimension[l]);
    for (uword l = 0; l < tensor.rank; ++l)
      if (!b_c[l]) nd.push_back(tensor.dimension[l]);
    DenseTensor out(uvec(nd), 0);
    std::vector<int64_t> da(dimension.vec().begin(), dimension.vec().end());
    std::vector<int64_t> db(tensor.dimension.vec().begin(), tensor.dimension.vec().end());
    check(tq_dense_contract(handle, dtype_of<T>(), (int32_t)rank, da.data(),
                            (int32_t)tensor.rank, db.data(), (int32_t)n_legs, legs.data(),
                            data, tensor.data, out.data),
          handle);
    return out;
  }

 private:
  T *data = nullptr;

  // uninitialised storage of the given dimensions (internal)
  DenseTensor(const uvec &dim, int) : rank(dim.n_elem()), dimension(dim) {
    index_table = util::generate_index_table(dim);
    allocate();
  }
  uword count() const {
    uword n = 1;
    for (uword l = 0; l < rank; ++l) n *= dimension[l];
    return n;
  }
  void allocate() {
    check(tq_malloc(default_handle(), reinterpret_cast<void **>(&data), sizeof(T) * count()),
          default_handle());
  }
  uword offset(const uvec &idx) const {
    if (idx.n_elem() != rank) throw Error("Rank does not match");
    uword o = 0;
    for (uword l = 0; l < rank; ++l) {
      if (idx[l] >= dimension[l]) throw Error("Indices out of boundary");
      o += idx[l] * index_table[l];
    }
    return o;
  }
};

}  // namespace gpu
}  // namespace torque
