// The reference's GPU test sections (test/gpu/block_sparse_test.cu:10-232,
// test/gpu/dense_test.cu:9-103, test/gpu/sparse_test.cu:45-333) written
// against the drop-in C++ classes:
// the same tensors, calls and expected values, with tq_handle_t in place of
// cublasHandle_t. Prints one line per failed check and exits non-zero.
#include <cstdio>
#include <vector>

#include "torque/gpu/block_sparse.h"
#include "torque/gpu/dense.h"
#include "torque/gpu/sparse.h"

static int g_fail = 0, g_checks = 0;
#define CHECK(cond)                                                     \
  do {                                                                  \
    ++g_checks;                                                         \
    if (!(cond)) {                                                      \
      ++g_fail;                                                         \
      std::printf("FAILED %s:%d: %s\n", __FILE__, __LINE__, #cond);     \
    }                                                                   \
  } while (0)

using torque::umat;
using torque::uvec;
using torque::uword;
using BST = torque::gpu::BlockSparseTensor<float>;
using DT = torque::gpu::DenseTensor<float>;
using ST = torque::gpu::SparseTensor<float>;
using torque::half;
using torque::gpu::BlockSparseTensor;
using torque::gpu::load_block_sparse;

static void block_sparse_sections(tq_handle_t handle) {
  {  // block sparse matrix initialization
    std::vector<float> a = {1, 2, 3, 4};
    BST tensor(a.data(), uvec{1, 1}, uvec{2, 2}, uvec{4, 3});
    CHECK(tensor.query({1, 1}) == 1);
    CHECK(tensor.query({2, 2}) == 4);
    CHECK(tensor.query({3, 2}) == 0);
    tensor.modify({2, 2}, 0);
    CHECK(tensor.query({2, 2}) == 0);
    tensor.modify({3, 2}, 5);
    CHECK(tensor.query({3, 2}) == 5);
    tensor.append_block(a.data(), {0, 0}, {1, 1}, {1, 2});
    CHECK(tensor.query({1, 1}) == 5);
  }
  {  // block sparse matrix transposition
    std::vector<float> a = {1, 2, 3, 4};
    BST tensor(a.data(), uvec{2, 0}, uvec{3, 1}, uvec{4, 3});
    tensor = tensor.hard_transpose({1, 0});
    CHECK(tensor.query({0, 3}) == 2);
    CHECK(tensor.query({1, 2}) == 3);
    const auto original = tensor.hard_transpose({1, 0});
    CHECK(original.query({3, 0}) == 2);
    CHECK(original.query({2, 1}) == 3);
  }
  {  // block sparse tensor initialization
    BST tensor(uvec{2, 2, 3});
    CHECK(tensor.query({1, 1, 0}) == 0);
    std::vector<float> a = {1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12};
    tensor.append_block(a.data(), {0, 0, 0}, {1, 1, 2}, {1, 2, 4});
    CHECK(tensor.query({0, 1, 2}) == 11);
  }
  {  // block sparse tensor transposition (overlapping blocks)
    std::vector<float> a = {1, 2, 3, 4, 5, 6, 0, 0, 0, 0, 0, 0};
    std::vector<float> b = {7, 8, 9, 10, 11, 12};
    BST tensor(uvec{2, 2, 3});
    tensor.append_block(a.data(), {0, 0, 0}, {1, 1, 2}, {1, 2, 4});
    tensor.append_block(b.data(), {1, 0, 0}, {1, 1, 2}, {1, 1, 2});
    const auto t = tensor.hard_transpose({1, 0, 2});
    CHECK(t.query({1, 0, 0}) == 3);
    CHECK(t.query({1, 0, 1}) == 0);
    CHECK(t.query({1, 1, 2}) == 12);
  }
  {  // scalar
    BST tensor(uvec{});
    CHECK(tensor.rank == 0);
    CHECK(tensor.query({}) == 0);
    tensor.modify({}, 1);
    CHECK(tensor.query({}) == 1);
  }
  {  // vector contraction
    const std::vector<float> v1{0, 1}, v2{2, 3, 4};
    BST t(uvec{5});
    t.append_block(v1.data(), {0}, {1}, {1});
    t.append_block(v2.data(), {2}, {4}, {1});
    CHECK(t.contract(handle, t, umat({{0, 0}})).to_number() == 30);
  }
  {  // matrix multiplication against the dense tensor
    const std::vector<float> vec{0, 1, 0, 0, 2, 3, 0, 0, 0, 0, 4, 5, 0, 0, 6, 7};
    const std::vector<float> b1{0, 1, 2, 3}, b2{4, 5, 6, 7};
    const DT dense(vec.data(), {4, 4});
    BST sparse(uvec{4, 4});
    sparse.append_block(b1.data(), {0, 0}, {1, 1}, {1, 2});
    sparse.append_block(b2.data(), {2, 2}, {3, 3}, {1, 2});
    const auto d2 = dense.contract(handle, dense, umat({1, 0}));
    const auto s2 = sparse.contract(handle, sparse, umat({{1, 0}}));
    for (uword i = 0; i < 4; ++i)
      for (uword j = 0; j < 4; ++j) CHECK(d2.query({i, j}) == s2.query({i, j}));
  }
  {  // matrix inner product with overlapping blocks
    const std::vector<float> b1{0, 1, 2, 3}, b2{0, 7, 8, 9};
    BST t(uvec{3, 3});
    t.append_block(b1.data(), {0, 0}, {1, 1}, {1, 2});
    t.append_block(b2.data(), {1, 1}, {2, 2}, {1, 2});
    // accu(ref % ref.t()) for ref = {{0,1,0},{2,3,7},{0,8,9}}^T
    const float ref[3][3] = {{0, 2, 0}, {1, 3, 8}, {0, 7, 9}};
    float expect = 0;
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) expect += ref[i][j] * ref[j][i];
    CHECK(t.contract(handle, t, umat({{1, 0}, {0, 1}})).to_number() == expect);
  }
  {  // tensor-vector contraction
    const std::vector<float> d1{0, 1, 2, 3}, d2{4, 5, 6, 7}, v{1, 2};
    BST t(uvec{2, 2, 2});
    t.append_block(d1.data(), {0, 0, 0}, {1, 1, 0}, {1, 2, 4});
    t.append_block(d2.data(), {0, 0, 1}, {1, 1, 1}, {1, 2, 4});
    BST vt(uvec{2});
    vt.append_block(v.data(), {0}, {1}, {1});
    const auto c1 = t.contract(handle, vt, umat{{0, 0}});
    CHECK(c1.query({0, 0}) == 2);
    CHECK(c1.query({1, 0}) == 8);
    CHECK(c1.query({0, 1}) == 14);
    CHECK(c1.query({1, 1}) == 20);
    const auto c2 = t.contract(handle, vt, umat{{1, 0}});
    CHECK(c2.query({0, 0}) == 4);
    CHECK(c2.query({1, 0}) == 7);
    CHECK(c2.query({0, 1}) == 16);
    CHECK(c2.query({1, 1}) == 19);
    const auto c3 = t.contract(handle, vt, umat{{2, 0}});
    CHECK(c3.query({1, 0}) == 11);
  }
  {  // tensor-matrix contraction
    const std::vector<float> d1{0, 1, 2, 3}, d2{4, 5, 6, 7}, d3{8, 9, 10, 11}, m{1, 2, 3, 4};
    BST t(uvec{2, 2, 3});
    t.append_block(d3.data(), {0, 0, 2}, {1, 1, 2}, {1, 2, 4});
    t.append_block(d2.data(), {0, 0, 1}, {1, 1, 1}, {1, 2, 4});
    t.append_block(d1.data(), {0, 0, 0}, {1, 1, 0}, {1, 2, 4});
    BST mt(uvec{2, 2});
    mt.append_block(m.data(), {0, 0}, {1, 1}, {1, 2});
    const auto c = t.contract(handle, mt, umat{{0, 1}, {1, 0}});
    CHECK(c.query({0}) == 19);
    CHECK(c.query({1}) == 59);
    CHECK(c.query({2}) == 99);
  }
  {  // errors keep the reference's messages
    BST t(uvec{5}), u(uvec{4});
    bool thrown = false;
    try {
      (void)t.contract(handle, u, umat{{0, 0}});
    } catch (const Error &e) {
      thrown = std::string(e.what()) ==
               "Error: The dimensions from two tensors to be contracted do not match";
    }
    CHECK(thrown);
  }
}

static void slice_section() {
  // slice as the reference's benchmarks use it (block_sparse_benchmark.cu:73-83):
  // the sliced tensor answers every query as before
  std::vector<float> a(16);
  for (int i = 0; i < 16; ++i) a[i] = (float)(i + 1);
  BST t(a.data(), uvec{0, 0}, uvec{3, 3}, uvec{4, 4});
  const auto s = t.slice({2, 2});
  CHECK(s.n_blocks() == 4);
  for (uword i = 0; i < 4; ++i)
    for (uword j = 0; j < 4; ++j) CHECK(s.query({i, j}) == t.query({i, j}));
}

static void dense_sections() {
  {  // dense matrix initialization
    DT t(uvec{4, 3});
    CHECK(t.query({1, 1}) == 0);
    std::vector<float> a = {1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12};
    t.initialize(a.data());
    CHECK(t.query({1, 2}) == 10);
  }
  {  // dense matrix transposition
    DT t(uvec{4, 3});
    std::vector<float> a = {1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12};
    t.initialize(a.data());
    CHECK(t.query({1, 2}) == 10);
    t = t.hard_transpose({1, 0});
    CHECK(t.query({1, 2}) == 7);
    const auto back = t.hard_transpose({1, 0});
    for (uword i = 0; i < 3; ++i)
      for (uword j = 0; j < 4; ++j) CHECK(t.query({i, j}) == back.query({j, i}));
  }
  {  // dense tensor transposition
    std::vector<float> a = {1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12};
    DT t(a.data(), {2, 2, 3});
    CHECK(t.query({0, 0, 2}) == 9);
    CHECK(t.query({0, 1, 2}) == 11);
    t = t.hard_transpose({0, 2, 1});
    CHECK(t.query({0, 2, 0}) == 9);
    CHECK(t.query({0, 2, 1}) == 11);
    const auto t2 = t.hard_transpose({1, 0, 2});
    CHECK(t2.query({2, 0, 0}) == 9);
    CHECK(t2.query({2, 0, 1}) == 11);
  }
  {  // scalar
    DT t(uvec{});
    CHECK(t.rank == 0);
    CHECK(t.query({}) == 0);
    t.modify({}, 1);
    CHECK(t.query({}) == 1);
  }
}

static void sparse_sections(tq_handle_t handle) {
  using torque::util::generate_index_table;
  {  // sparse matrix initialization
    ST t(uvec{4, 3});
    CHECK(t.query({1, 1}) == 0);
    std::vector<float> a = {1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12};
    t.initialize(a.data(), {0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11});
    CHECK(t.query({1, 2}) == 10);
  }
  {  // sparse matrix transposition
    ST t(uvec{4, 3});
    std::vector<float> a = {1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12};
    t.initialize(a.data(), {0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11});
    CHECK(t.query({1, 2}) == 10);
    t.soft_transpose({1, 0});
    CHECK(t.query({1, 2}) == 7);
    const auto tt = t.hard_transpose({1, 0});
    for (uword i = 0; i < 3; ++i)
      for (uword j = 0; j < 4; ++j) CHECK(t.query({i, j}) == tt.query({j, i}));
  }
  {  // sparse tensor initialization
    ST t(uvec{2, 2, 3});
    CHECK(t.query({1, 1, 0}) == 0);
    std::vector<float> a = {1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12};
    t.initialize(a.data(), {0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11});
    CHECK(t.query({0, 1, 2}) == 11);
  }
  {  // sparse tensor transposition
    std::vector<float> a = {1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12};
    const uvec dim{2, 2, 3};
    ST t(a.data(), {0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11}, generate_index_table(dim), dim);
    CHECK(t.query({0, 0, 2}) == 9);
    CHECK(t.query({0, 1, 2}) == 11);
    t.soft_transpose({0, 2, 1});
    CHECK(t.query({0, 2, 0}) == 9);
    CHECK(t.query({0, 2, 1}) == 11);
    const auto t2 = t.hard_transpose({1, 0, 2});
    CHECK(t2.query({2, 0, 0}) == 9);
    CHECK(t2.query({2, 0, 1}) == 11);
  }
  {  // scalar
    ST t(uvec{});
    CHECK(t.rank == 0);
    CHECK(t.query({}) == 0);
    t.modify({}, 1);
    CHECK(t.query({}) == 1);
  }
  {  // vector contraction
    std::vector<float> v = {0, 1, 2, 3, 4};
    const uvec dim{5};
    const ST t(v.data(), {0, 1, 2, 3, 4}, generate_index_table(dim), dim);
    CHECK(t.contract(handle, t, umat({0, 0})).to_number() == 30);
  }
  {  // matrix multiplication
    std::vector<float> v = {0, 1, 2, 3, 4, 5, 6, 7, 8};
    const uvec dim{3, 3};
    const ST t(v.data(), {0, 1, 2, 3, 4, 5, 6, 7, 8}, generate_index_table(dim), dim);
    const auto sq = t.contract(handle, t, umat({1, 0}));
    const float ref[3][3] = {{15, 42, 69}, {18, 54, 90}, {21, 66, 111}};
    for (uword i = 0; i < 3; ++i)
      for (uword j = 0; j < 3; ++j) CHECK(sq.query({i, j}) == ref[i][j]);
  }
  {  // matrix inner product
    std::vector<float> v = {0, 1, 2, 3, 4, 5, 6, 7, 8};
    const uvec dim{3, 3};
    const ST t(v.data(), {0, 1, 2, 3, 4, 5, 6, 7, 8}, generate_index_table(dim), dim);
    float expect = 0;
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) expect += v[i + 3 * j] * v[j + 3 * i];
    CHECK(t.contract(handle, t, umat({{1, 0}, {0, 1}})).to_number() == expect);
  }
  {  // tensor-vector contraction
    std::vector<float> td = {0, 1, 2, 3, 4, 5, 6, 7};
    const uvec tdim{2, 2, 2};
    const ST t(td.data(), {0, 1, 2, 3, 4, 5, 6, 7}, generate_index_table(tdim), tdim);
    std::vector<float> vd = {1, 2};
    const ST v(vd.data(), {0, 1}, generate_index_table({2}), {2});
    const auto c = t.contract(handle, v, umat{{0, 0}});
    CHECK(c.query({0, 0}) == 2);
    CHECK(c.query({1, 0}) == 8);
    CHECK(c.query({0, 1}) == 14);
    CHECK(c.query({1, 1}) == 20);
    const auto c2 = t.contract(handle, v, umat{{1, 0}});
    CHECK(c2.query({0, 0}) == 4);
    CHECK(c2.query({1, 0}) == 7);
    CHECK(c2.query({0, 1}) == 16);
    CHECK(c2.query({1, 1}) == 19);
    const auto c3 = t.contract(handle, v, umat{{2, 0}});
    CHECK(c3.query({1, 0}) == 11);
  }
  {  // tensor-matrix contraction
    std::vector<float> td = {0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11};
    const uvec tdim{2, 2, 3};
    const ST t(td.data(), {0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11}, generate_index_table(tdim),
               tdim);
    std::vector<float> md = {1, 2, 3, 4};
    const uvec mdim{2, 2};
    const ST m(md.data(), {0, 1, 2, 3}, generate_index_table(mdim), mdim);
    const auto c = t.contract(handle, m, umat{{0, 1}, {1, 0}});
    CHECK(c.query({0}) == 19);
    CHECK(c.query({1}) == 59);
    CHECK(c.query({2}) == 99);
  }
}

// Not a reference section: contractions repeated through the handle's plan
// cache give the same values, and contract_merged writes one block per
// distinct output block with the same dense image.
static void cache_and_merged_section(tq_handle_t handle) {
  const std::vector<float> d1{0, 1, 2, 3}, d2{4, 5, 6, 7}, v{1, 2};
  BST t(uvec{2, 2, 2});
  t.append_block(d1.data(), {0, 0, 0}, {1, 1, 0}, {1, 2, 4});
  t.append_block(d2.data(), {0, 0, 1}, {1, 1, 1}, {1, 2, 4});
  BST vt(uvec{2});
  vt.append_block(v.data(), {0}, {1}, {1});
  for (int rep = 0; rep < 3; ++rep) {
    const auto c = t.contract(handle, vt, umat{{2, 0}});
    CHECK(c.n_blocks() == 2);
    CHECK(c.query({1, 0}) == 11);
    CHECK(c.query({0, 1}) == 2 * 1 + 6 * 2);
  }
  const auto m = t.contract_merged(handle, vt, umat{{2, 0}});
  CHECK(m.n_blocks() == 1);
  CHECK(m.query({0, 0}) == 8);
  CHECK(m.query({1, 0}) == 11);
  CHECK(m.query({0, 1}) == 14);
  CHECK(m.query({1, 1}) == 17);
  // a cached plan belongs to the handle: destroying it is refused
  std::vector<int64_t> dims{2, 2, 2}, beg{0, 0, 0}, end{1, 1, 1}, str{1, 2, 4}, off{0};
  std::vector<int64_t> vd{2}, vb{0}, ve{1}, vs{1};
  tq_bs_desc da{3, 1, dims.data(), beg.data(), end.data(), str.data(), off.data()};
  tq_bs_desc db{1, 1, vd.data(), vb.data(), ve.data(), vs.data(), off.data()};
  const int32_t legs[2] = {2, 0};
  tq_plan_t p1 = nullptr, p2 = nullptr;
  CHECK(tq_bs_contract_plan_cached(handle, &da, &db, 1, legs, TQ_F32, TQ_PREC_NATIVE, 0, &p1) ==
        TQ_SUCCESS);
  CHECK(tq_bs_contract_plan_cached(handle, &da, &db, 1, legs, TQ_F32, TQ_PREC_NATIVE, 0, &p2) ==
        TQ_SUCCESS);
  CHECK(p1 != nullptr && p1 == p2);
  CHECK(tq_plan_destroy(p1) == TQ_ERROR_INVALID_VALUE);
}

// Not a reference section: the reference's `half` instantiation as
// torque::half (FP16 MFMA, FP32 accumulation); small integers keep every
// product and sum exact.
static void half_section(tq_handle_t handle) {
  using HT = BlockSparseTensor<half>;
  const std::vector<half> d1{0, 1, 2, 3}, d2{4, 5, 6, 7}, v{1, 2};
  HT t(uvec{2, 2, 2});
  t.append_block(d1.data(), {0, 0, 0}, {1, 1, 0}, {1, 2, 4});
  t.append_block(d2.data(), {0, 0, 1}, {1, 1, 1}, {1, 2, 4});
  HT vt(uvec{2});
  vt.append_block(v.data(), {0}, {1}, {1});
  const auto c1 = t.contract(handle, vt, umat{{0, 0}});
  CHECK(float(c1.query({0, 0})) == 2);
  CHECK(float(c1.query({1, 0})) == 8);
  CHECK(float(c1.query({0, 1})) == 14);
  CHECK(float(c1.query({1, 1})) == 20);
  CHECK(float(t.contract(handle, vt, umat{{2, 0}}).query({1, 0})) == 11);
  CHECK(float(t.contract_merged(handle, vt, umat{{2, 0}}).query({1, 1})) == 17);
  const auto tt = t.hard_transpose({2, 0, 1});
  CHECK(float(tt.query({1, 0, 1})) == 6);
  CHECK(float(t.contract(handle, t, umat{{0, 0}, {1, 1}, {2, 2}}).to_number()) == 140);
  // one 48 x 40 x 72 GEMM block through the MFMA tiles (K > 8)
  const uword m = 48, k = 40, n = 72;
  std::vector<half> a(m * k), b(k * n);
  for (uword i = 0; i < m * k; ++i) a[i] = half((int)(i * 7 % 7) - 3 + (int)(i % 5));
  for (uword i = 0; i < k * n; ++i) b[i] = half((int)(i * 3 % 5) - 2);
  HT A(a.data(), uvec{0, 0}, uvec{m - 1, k - 1}, uvec{m, k});
  HT B(b.data(), uvec{0, 0}, uvec{k - 1, n - 1}, uvec{k, n});
  const auto C = A.contract(handle, B, umat{{1, 0}});
  bool exact = true;
  for (uword i = 0; i < m; ++i)
    for (uword j = 0; j < n; ++j) {
      float ref = 0;
      for (uword q = 0; q < k; ++q) ref += float(a[i + m * q]) * float(b[q + k * j]);
      exact = exact && float(C.query({i, j})) == ref;
    }
  CHECK(exact);
}

// Not a reference section: the binary block-tensor file. Reads dir/in.tqbs
// (written by the Python test), contracts it with itself over legs (2, 0),
// (3, 1) and writes the result to dir/out.tqbs for the Python test to check.
static void file_section(tq_handle_t handle, const std::string &dir) {
  if (dir.empty()) return;
  const auto t = load_block_sparse<double>(dir + "/in.tqbs");
  const auto c = t.contract(handle, t, umat({{2, 0}, {3, 1}}));
  torque::gpu::save(c, dir + "/out.tqbs");
  const auto back = load_block_sparse<double>(dir + "/out.tqbs");
  CHECK(back.n_blocks() == c.n_blocks());
  CHECK(back.query({1, 2, 3, 4}) == c.query({1, 2, 3, 4}));
}

int main(int argc, char **argv) {
  tq_handle_t handle = nullptr;
  if (tq_create(&handle, 0) != TQ_SUCCESS) {
    std::printf("FAILED: no device\n");
    return 2;
  }
  try {
    block_sparse_sections(handle);
    dense_sections();
    sparse_sections(handle);
    slice_section();
    cache_and_merged_section(handle);
    half_section(handle);
    file_section(handle, argc > 1 ? argv[1] : "");
  } catch (const std::exception &e) {
    std::printf("FAILED with exception: %s\n", e.what());
    ++g_fail;
  }
  tq_destroy(handle);
  std::printf("%d checks, %d failed\n", g_checks, g_fail);
  return g_fail ? 1 : 0;
}
