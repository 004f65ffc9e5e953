// torque/gpu/sparse.h -- drop-in torque::gpu::SparseTensor<T> over the C-ABI
// of libtorque_hip.so (src/gpu/sparse.cuh:84-441 of the reference): element
// i of the tensor is (indices[i], data[i]), indices[i] being the linear index
// under index_table. The same public members, constructors and methods
// (initialize / modify / query / to_number / contract / soft_transpose /
// hard_transpose) and the same "Error: ..." messages. Differences a caller
// sees: contract takes a tq_handle_t where the reference takes a
// cublasHandle_t; indices are 64-bit (the reference narrows to int32); the
// device vectors are torque::gpu::DeviceVector in place of
// thrust::device_vector.
#pragma once

#include <algorithm>
#include <utility>
#include <vector>

#include "torque/gpu/block_sparse.h"

namespace torque {
namespace gpu {

// A device array with the few thrust::device_vector operations the tensor
// API uses: size(), element read, data().
template <typename E>
class DeviceVector {
 public:
  DeviceVector() = default;
  explicit DeviceVector(const std::vector<E> &host) { assign(host.data(), host.size()); }
  DeviceVector(const DeviceVector &o) {
    resize(o.n_);
    if (n_)
      check(tq_memcpy(default_handle(), p_, o.p_, sizeof(E) * n_, DeviceToDevice),
            default_handle());
  }
  DeviceVector(DeviceVector &&o) noexcept { swap(*this, o); }
  DeviceVector &operator=(DeviceVector o) noexcept {
    swap(*this, o);
    return *this;
  }
  friend void swap(DeviceVector &a, DeviceVector &b) noexcept {
    std::swap(a.p_, b.p_);
    std::swap(a.n_, b.n_);
  }
  ~DeviceVector() {
    if (p_) (void)tq_free(default_handle(), p_);
  }
  // adopts memory allocated by the library (tq_malloc or a C-ABI output)
  static DeviceVector adopt(E *p, std::size_t n) {
    DeviceVector v;
    v.p_ = p;
    v.n_ = n;
    return v;
  }

  std::size_t size() const { return n_; }
  bool empty() const { return n_ == 0; }
  E *data() { return p_; }
  const E *data() const { return p_; }
  E operator[](std::size_t i) const {
    E v;
    check(tq_memcpy(default_handle(), &v, p_ + i, sizeof(E), DeviceToHost), default_handle());
    check(tq_synchronize(default_handle()), default_handle());
    return v;
  }
  void set(std::size_t i, const E &v) {
    check(tq_memcpy(default_handle(), p_ + i, &v, sizeof(E), HostToDevice), default_handle());
    check(tq_synchronize(default_handle()), default_handle());
  }
  std::vector<E> to_host() const {
    std::vector<E> h(n_);
    if (n_) {
      check(tq_memcpy(default_handle(), h.data(), p_, sizeof(E) * n_, DeviceToHost),
            default_handle());
      check(tq_synchronize(default_handle()), default_handle());
    }
    return h;
  }
  void assign(const E *host, std::size_t n) {
    resize(n);
    if (n)
      check(tq_memcpy(default_handle(), p_, host, sizeof(E) * n, HostToDevice),
            default_handle());
  }
  // new length; the first min(old, n) elements are kept
  void resize(std::size_t n) {
    if (n == n_ && p_) return;
    E *q = nullptr;
    check(tq_malloc(default_handle(), reinterpret_cast<void **>(&q), sizeof(E) * std::max<std::size_t>(n, 1)),
          default_handle());
    const std::size_t keep = std::min(n, n_);
    if (keep)
      check(tq_memcpy(default_handle(), q, p_, sizeof(E) * keep, DeviceToDevice),
            default_handle());
    check(tq_synchronize(default_handle()), default_handle());
    if (p_) (void)tq_free(default_handle(), p_);
    p_ = q;
    n_ = n;
  }

 private:
  E *p_ = nullptr;
  std::size_t n_ = 0;
};

template <typename T>
class SparseTensor {
 public:
  uword rank = 0;
  uvec dimension;
  DeviceVector<int64_t> indices;
  uvec index_table;

  // src/gpu/sparse.cuh:98-107: no elements, data = {0}
  explicit SparseTensor(const uvec &dim) : rank(dim.n_elem()), dimension(dim) {
    index_table = util::generate_index_table(dim);
    const T zero = 0;
    data.assign(&zero, 1);
  }

  // src/gpu/sparse.cuh:110-145
  SparseTensor(const T *source_data, const uvec &idx, const uvec &table, const uvec &dim)
      : rank(dim.n_elem()), dimension(dim), index_table(table) {
    if (!source_data) throw Error("Source data not allocated!");
    if (rank > 0) {
      set_indices(idx);
      data.assign(source_data, idx.n_elem());
    } else {
      data.assign(source_data, 1);
    }
  }

  T to_number() const { return data[0]; }

  // src/gpu/sparse.cuh:174-187
  void initialize(const T *source_data, const uvec &idx) {
    if (!source_data) throw Error("Source data not allocated!");
    set_indices(idx);
    data.assign(source_data, idx.n_elem());
  }

  // src/gpu/sparse.cuh:191-236: overwrite, remove (number == 0: the last
  // element moves into the hole) or append
  void modify(const uvec &idx, const T number) {
    if (idx.n_elem() != rank) throw Error("Rank does not match");
    if (rank == 0) {
      data.set(0, number);
      return;
    }
    const int64_t key = linear(idx);
    std::vector<int64_t> h = indices.to_host();
    const auto it = std::find(h.begin(), h.end(), key);
    const std::size_t n = h.size();
    if (it != h.end()) {
      const std::size_t pos = (std::size_t)(it - h.begin());
      if (number == 0) {
        data.set(pos, data[n - 1]);
        indices.set(pos, h[n - 1]);
        indices.resize(n - 1);
        data.resize(n - 1);
      } else {
        data.set(pos, number);
      }
    } else if (number != 0) {
      indices.resize(n + 1);
      indices.set(n, key);
      data.resize(n + 1);
      data.set(n, number);
    }
  }

  // src/gpu/sparse.cuh:240-266
  T query(const uvec &idx) const {
    if (idx.n_elem() != rank) throw Error("Rank does not match");
    if (rank == 0) return data[0];
    const int64_t key = linear(idx);
    const std::vector<int64_t> h = indices.to_host();
    const auto it = std::find(h.begin(), h.end(), key);
    return it == h.end() ? T(0) : data[(std::size_t)(it - h.begin())];
  }

  // src/gpu/sparse.cuh:270-375 through tq_sparse_contract: every output
  // index that receives a product, ascending, with its summed value
  SparseTensor contract(tq_handle_t handle, const SparseTensor &tensor,
                        const umat &contracting_indices) const {
    const std::size_t n_legs = contracting_indices.n_rows();
    std::vector<int32_t> legs(2 * n_legs);
    for (std::size_t k = 0; k < n_legs; ++k) {
      legs[2 * k] = (int32_t)contracting_indices(k, 0);
      legs[2 * k + 1] = (int32_t)contracting_indices(k, 1);
      if (contracting_indices(k, 0) >= rank || contracting_indices(k, 1) >= tensor.rank ||
          dimension[legs[2 * k]] != tensor.dimension[legs[2 * k + 1]])
        throw Error("The dimensions from two tensors to be contracted do not match");
    }
    std::vector<int64_t> da, ta, db, tb;
    tq_sp_desc a = desc(da, ta), b = tensor.desc(db, tb);
    const uword out_rank = rank + tensor.rank - 2 * n_legs;
    std::vector<int64_t> od(std::max<uword>(out_rank, 1));
    int64_t nnz = 0;
    int64_t *pi = nullptr;
    void *pv = nullptr;
    check(tq_sparse_contract(handle, dtype_of<T>(), &a, &b, (int32_t)n_legs, legs.data(),
                             od.data(), &nnz, &pi, &pv),
          handle);
    std::vector<uword> nd(od.begin(), od.begin() + out_rank);
    SparseTensor out{uvec(nd), 0};
    if (out_rank > 0) {
      out.indices = DeviceVector<int64_t>::adopt(pi, (std::size_t)nnz);
    } else {
      (void)tq_free(handle, pi);
      nnz = 1;
    }
    out.data = DeviceVector<T>::adopt(reinterpret_cast<T *>(pv), (std::size_t)nnz);
    return out;
  }

  // src/gpu/sparse.cuh:380-392
  void soft_transpose(const uvec &permutation) {
    if (permutation.n_elem() != rank)
      throw Error("The number of permutation does not match the rank of tensor");
    uvec t(rank), d(rank);
    for (uword k = 0; k < rank; ++k) {
      t[k] = index_table[permutation[k]];
      d[k] = dimension[permutation[k]];
    }
    index_table = t;
    dimension = d;
  }

  // src/gpu/sparse.cuh:395-432 (each stored index is decoded; the reference
  // decodes the storage position, :420-422)
  SparseTensor hard_transpose(const uvec &permutation) const {
    if (permutation.n_elem() != rank)
      throw Error("The number of permutation does not match the rank of tensor");
    uvec nd(rank);
    std::vector<int32_t> perm(rank);
    for (uword k = 0; k < rank; ++k) {
      nd[k] = dimension[permutation[k]];
      perm[k] = (int32_t)permutation[k];
    }
    SparseTensor out{nd, 0};
    out.data = data;
    out.indices.resize(indices.size());
    std::vector<int64_t> d(dimension.vec().begin(), dimension.vec().end());
    std::vector<int64_t> t(index_table.vec().begin(), index_table.vec().end());
    check(tq_sparse_permute_indices(default_handle(), (int32_t)rank, d.data(), t.data(),
                                    perm.data(), (int64_t)indices.size(), indices.data(),
                                    out.indices.data()),
          default_handle());
    check(tq_synchronize(default_handle()), default_handle());
    return out;
  }

  const T *device_data() const { return data.data(); }

 protected:
  DeviceVector<T> data;

 private:
  // empty tensor of the given dimension (internal)
  SparseTensor(const uvec &dim, int) : rank(dim.n_elem()), dimension(dim) {
    index_table = util::generate_index_table(dim);
  }
  int64_t linear(const uvec &idx) const {
    int64_t o = 0;
    for (uword l = 0; l < rank; ++l) o += (int64_t)(idx[l] * index_table[l]);
    return o;
  }
  void set_indices(const uvec &idx) {
    std::vector<int64_t> h(idx.vec().begin(), idx.vec().end());
    indices.assign(h.data(), h.size());
  }
  tq_sp_desc desc(std::vector<int64_t> &d, std::vector<int64_t> &t) const {
    d.assign(dimension.vec().begin(), dimension.vec().end());
    t.assign(index_table.vec().begin(), index_table.vec().end());
    return tq_sp_desc{(int32_t)rank, d.data(), t.data(), (int64_t)indices.size(),
                      indices.data(), data.data()};
  }
};

}  // namespace gpu
}  // namespace torque
