// torque/index.h -- the small index containers the tensor classes use in
// place of Armadillo's uvec / umat (the reference's src/util/space.h types).
// Only what the tensor API needs: brace initialisation, element access,
// columns, and the index helpers of src/util/space.h:10-88.
#pragma once

#include <cstdint>
#include <initializer_list>
#include <numeric>
#include <stdexcept>
#include <vector>

namespace torque {

using uword = uint64_t;

// Column vector of unsigned indices (arma::uvec).
class uvec {
 public:
  uvec() = default;
  explicit uvec(std::size_t n, uword fill = 0) : v_(n, fill) {}
  uvec(std::initializer_list<uword> l) : v_(l) {}
  explicit uvec(std::vector<uword> v) : v_(std::move(v)) {}

  std::size_t n_elem() const { return v_.size(); }
  uword &operator()(std::size_t i) { return v_.at(i); }
  uword operator()(std::size_t i) const { return v_.at(i); }
  uword &operator[](std::size_t i) { return v_[i]; }
  uword operator[](std::size_t i) const { return v_[i]; }
  const uword *data() const { return v_.data(); }
  uword *data() { return v_.data(); }
  const std::vector<uword> &vec() const { return v_; }
  bool operator==(const uvec &o) const { return v_ == o.v_; }

 private:
  std::vector<uword> v_;
};

// Column-major matrix of unsigned indices (arma::umat). As with Armadillo's
// brace initialisers, umat({a, b}) and umat({{a, b}}) are one row; each inner
// list is a row. Block metadata keeps one column per block.
class umat {
 public:
  umat() = default;
  umat(std::size_t rows, std::size_t cols, uword fill = 0)
      : r_(rows), c_(cols), v_(rows * cols, fill) {}
  umat(std::initializer_list<uword> row) : r_(1), c_(row.size()), v_(row) {}
  umat(std::initializer_list<std::initializer_list<uword>> rows) {
    r_ = rows.size();
    c_ = r_ ? rows.begin()->size() : 0;
    v_.assign(r_ * c_, 0);
    std::size_t i = 0;
    for (const auto &row : rows) {
      if (row.size() != c_) throw std::invalid_argument("umat: ragged initialiser");
      std::size_t j = 0;
      for (uword x : row) (*this)(i, j++) = x;
      ++i;
    }
  }

  // one column (a uvec used where the reference passes a umat)
  static umat column(const uvec &c) {
    umat m;
    m.append_col(c);
    return m;
  }

  std::size_t n_rows() const { return r_; }
  std::size_t n_cols() const { return c_; }
  uword &operator()(std::size_t i, std::size_t j) { return v_.at(j * r_ + i); }
  uword operator()(std::size_t i, std::size_t j) const { return v_.at(j * r_ + i); }
  uvec col(std::size_t j) const {
    uvec out(r_);
    for (std::size_t i = 0; i < r_; ++i) out[i] = (*this)(i, j);
    return out;
  }
  void append_col(const uvec &c) {
    if (c_ == 0 && r_ == 0) r_ = c.n_elem();
    if (c.n_elem() != r_) throw std::invalid_argument("umat: column length");
    v_.insert(v_.end(), c.vec().begin(), c.vec().end());
    ++c_;
  }
  const uword *data() const { return v_.data(); }
  uword *data() { return v_.data(); }

 private:
  std::size_t r_ = 0, c_ = 0;
  std::vector<uword> v_;
};

namespace util {

// util::generate_index_table, src/util/space.h:10-24
inline uvec generate_index_table(const uvec &dimension) {
  uvec t(dimension.n_elem());
  uword acc = 1;
  for (std::size_t i = 0; i < dimension.n_elem(); ++i) {
    t[i] = acc;
    acc *= dimension[i];
  }
  return t;
}

// util::nest_sum, src/util/space.h:76-88
inline uvec nest_sum(const uvec &v, uword initial = 0) {
  uvec out(v.n_elem());
  uword acc = initial;
  for (std::size_t i = 0; i < v.n_elem(); ++i) {
    out[i] = acc;
    acc += v[i];
  }
  return out;
}

}  // namespace util
}  // namespace torque
