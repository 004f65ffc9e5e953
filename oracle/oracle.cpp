// oracle.cpp -- CPU restatement of Walter-Feng/Torque's block-sparse tensor
// algorithms, used ONLY as the checker by the tests, the smoke check and the
// benchmark's CPU-baseline leg. Plain C++17 + OpenMP, no Armadillo (which the
// reference needs and which is absent here). Each function cites the
// reference file:line it follows; see oracle.h for the interface.
#include "oracle.h"

#include <algorithm>
#include <cstring>
#include <map>
#include <numeric>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

using i64 = int64_t;

// Free legs of a tensor of rank r after removing the contracted legs, in
// ascending leg order (arma shed_rows keeps the remaining order).
std::vector<int> free_legs(int rank, const std::vector<int> &contracted) {
  std::vector<int> out;
  for (int l = 0; l < rank; ++l)
    if (std::find(contracted.begin(), contracted.end(), l) == contracted.end())
      out.push_back(l);
  return out;
}

// One matched pair of block_sparse::block_in_range
// (src/tensor/block_sparse.cpp:38-60 for the test, :62-148 for the points).
struct Pair {
  i64 a_block, b_block;
  std::vector<i64> lo, hi;  // contracted intersection per leg pair
};

struct Plan {
  int out_rank = 0;
  std::vector<int> a_legs, b_legs, a_free, b_free;
  std::vector<i64> out_dims;
  std::vector<Pair> pairs;
};

bool make_plan(const orc_bs *a, const orc_bs *b, int n_legs, const int32_t *legs,
               Plan &p) {
  p.a_legs.resize(n_legs);
  p.b_legs.resize(n_legs);
  for (int k = 0; k < n_legs; ++k) {
    p.a_legs[k] = legs[2 * k];
    p.b_legs[k] = legs[2 * k + 1];
    // dimension check, src/tensor/block_sparse.h:434-441
    if (a->dims[p.a_legs[k]] != b->dims[p.b_legs[k]]) return false;
  }
  p.a_free = free_legs(a->rank, p.a_legs);
  p.b_free = free_legs(b->rank, p.b_legs);
  // new dimension = A free ++ B free, src/tensor/block_sparse.h:445-451
  for (int l : p.a_free) p.out_dims.push_back(a->dims[l]);
  for (int l : p.b_free) p.out_dims.push_back(b->dims[l]);
  p.out_rank = (int)p.out_dims.size();
  // loop over A blocks, src/tensor/block_sparse.h:455-468, and the B scan of
  // block_in_range, src/tensor/block_sparse.cpp:38-60 (ascending B order,
  // arma::find at :62)
  for (i64 i = 0; i < a->n_blocks; ++i) {
    const i64 *ab = a->begin + i * a->rank;
    const i64 *ae = a->end + i * a->rank;
    for (i64 j = 0; j < b->n_blocks; ++j) {
      const i64 *bb = b->begin + j * b->rank;
      const i64 *be = b->end + j * b->rank;
      Pair pr;
      pr.a_block = i;
      pr.b_block = j;
      bool ok = true;
      for (int k = 0; k < n_legs && ok; ++k) {
        i64 lo = std::max(ab[p.a_legs[k]], bb[p.b_legs[k]]);
        i64 hi = std::min(ae[p.a_legs[k]], be[p.b_legs[k]]);
        ok = lo <= hi;
        pr.lo.push_back(lo);
        pr.hi.push_back(hi);
      }
      if (ok) p.pairs.push_back(std::move(pr));
    }
  }
  return true;
}

// Column-major decode of a linear index over dims (space.h:26-42 with the
// table of generate_index_table(dims)).
inline void decode(i64 idx, int n, const i64 *dims, i64 *digits) {
  for (int l = 0; l < n; ++l) {
    digits[l] = idx % dims[l];
    idx /= dims[l];
  }
}

struct PairGeom {
  i64 m, n, k;
  i64 a_off, b_off, c_off;
  std::vector<i64> a_dims, a_str;  // [free..., contracted...] of the A view
  std::vector<i64> b_dims, b_str;
  std::vector<i64> out_begin, out_end;
};

PairGeom geometry(const orc_bs *a, const orc_bs *b, const Plan &p,
                  const Pair &pr) {
  PairGeom g;
  const i64 *ab = a->begin + pr.a_block * a->rank;
  const i64 *ae = a->end + pr.a_block * a->rank;
  const i64 *as = a->strides + pr.a_block * a->rank;
  const i64 *bb = b->begin + pr.b_block * b->rank;
  const i64 *be = b->end + pr.b_block * b->rank;
  const i64 *bs = b->strides + pr.b_block * b->rank;
  const int nl = (int)p.a_legs.size();
  // sub-block offsets: block_offsets(i) + (sub_begin - begin)^T table_i,
  // src/tensor/block_sparse.h:516-525 and src/gpu/block_sparse.cuh:700-732
  g.a_off = a->offsets[pr.a_block];
  g.b_off = b->offsets[pr.b_block];
  for (int k = 0; k < nl; ++k) {
    g.a_off += (pr.lo[k] - ab[p.a_legs[k]]) * as[p.a_legs[k]];
    g.b_off += (pr.lo[k] - bb[p.b_legs[k]]) * bs[p.b_legs[k]];
  }
  g.m = 1;
  for (int l : p.a_free) {
    g.a_dims.push_back(ae[l] - ab[l] + 1);
    g.a_str.push_back(as[l]);
    g.m *= ae[l] - ab[l] + 1;
  }
  g.n = 1;
  for (int l : p.b_free) {
    g.b_dims.push_back(be[l] - bb[l] + 1);
    g.b_str.push_back(bs[l]);
    g.n *= be[l] - bb[l] + 1;
  }
  g.k = 1;
  for (int k = 0; k < nl; ++k) {
    const i64 d = pr.hi[k] - pr.lo[k] + 1;
    g.a_dims.push_back(d);
    g.a_str.push_back(as[p.a_legs[k]]);
    g.b_dims.push_back(d);
    g.b_str.push_back(bs[p.b_legs[k]]);
    g.k *= d;
  }
  // new begin/end points: A free ++ B free (src/tensor/block_sparse.cpp:66-119)
  for (int l : p.a_free) {
    g.out_begin.push_back(ab[l]);
    g.out_end.push_back(ae[l]);
  }
  for (int l : p.b_free) {
    g.out_begin.push_back(bb[l]);
    g.out_end.push_back(be[l]);
  }
  g.c_off = 0;
  return g;
}

// Gather a strided view with dims/strides (fastest leg first) into a compact
// column-major buffer: the element map of sort_into_matrix
// (src/tensor/block_sparse.h:48-76) once the legs are listed as
// [rows..., cols...].
void gather(const double *src, int n, const i64 *dims, const i64 *str,
            double *dst) {
  i64 total = 1;
  for (int l = 0; l < n; ++l) total *= dims[l];
  std::vector<i64> dig(n, 0);
  i64 off = 0;
  for (i64 e = 0; e < total; ++e) {
    dst[e] = src[off];
    // odometer increment
    for (int l = 0; l < n; ++l) {
      if (++dig[l] < dims[l]) {
        off += str[l];
        break;
      }
      off -= (dims[l] - 1) * str[l];
      dig[l] = 0;
    }
  }
}

// C(m x n) = A(m x k) * B(n x k)^T, all column-major compact
// (the arma product at src/tensor/block_sparse.h:541-542).
void gemm_nt(i64 m, i64 n, i64 k, const double *A, const double *B,
             double *C) {
  for (i64 j = 0; j < n; ++j) {
    double *c = C + j * m;
    for (i64 i = 0; i < m; ++i) c[i] = 0.0;
    for (i64 l = 0; l < k; ++l) {
      const double bv = B[j + n * l];
      const double *acol = A + l * m;
      for (i64 i = 0; i < m; ++i) c[i] += acol[i] * bv;
    }
  }
}

}  // namespace

extern "C" {

void orc_index_table(int32_t rank, const int64_t *dims, int64_t *table) {
  // src/util/space.h:10-24
  int64_t t = 1;
  for (int l = 0; l < rank; ++l) {
    table[l] = t;
    t *= dims[l];
  }
}

void orc_index_to_indices(int64_t i, int32_t rank, const int64_t *table,
                          const int64_t *sort_index, int64_t *indices) {
  // src/util/space.h:26-59: peel digits from the slowest stride down
  if (rank == 0) return;
  int64_t t = i;
  for (int j = rank - 1; j > 0; --j) {
    const int64_t l = sort_index ? sort_index[j] : j;
    indices[l] = t / table[l];
    t %= table[l];
  }
  indices[sort_index ? sort_index[0] : 0] = t;
}

void orc_nest_sum(int64_t n, const int64_t *in, int64_t init, int64_t *out) {
  // src/util/space.h:76-88
  int64_t t = init;
  for (int64_t i = 0; i < n; ++i) {
    out[i] = t;
    t += in[i];
  }
}

int64_t orc_bs_count_pairs(const orc_bs *a, const orc_bs *b, int32_t n_legs,
                           const int32_t *legs) {
  Plan p;
  if (!make_plan(a, b, n_legs, legs, p)) return -1;
  return (int64_t)p.pairs.size();
}

int64_t orc_bs_plan(const orc_bs *a, const orc_bs *b, int32_t n_legs,
                    const int32_t *legs, int64_t *a_block, int64_t *b_block,
                    int64_t *out_begin, int64_t *out_end, int64_t *a_offset,
                    int64_t *b_offset, int64_t *m, int64_t *n, int64_t *k,
                    int64_t *out_offsets) {
  Plan p;
  if (!make_plan(a, b, n_legs, legs, p)) return -1;
  const int ro = p.out_rank;
  int64_t off = 0;
  for (size_t q = 0; q < p.pairs.size(); ++q) {
    PairGeom g = geometry(a, b, p, p.pairs[q]);
    if (a_block) a_block[q] = p.pairs[q].a_block;
    if (b_block) b_block[q] = p.pairs[q].b_block;
    for (int l = 0; l < ro; ++l) {
      if (out_begin) out_begin[q * ro + l] = g.out_begin[l];
      if (out_end) out_end[q * ro + l] = g.out_end[l];
    }
    if (a_offset) a_offset[q] = g.a_off;
    if (b_offset) b_offset[q] = g.b_off;
    if (m) m[q] = g.m;
    if (n) n[q] = g.n;
    if (k) k[q] = g.k;
    // new_blocks_offsets = nest_sum(prod(dims)), src/gpu/block_sparse.cuh:667-671
    if (out_offsets) out_offsets[q] = off;
    off += g.m * g.n;
  }
  return (int64_t)p.pairs.size();
}

void orc_sort_into_matrix_f64(const double *block, int32_t rank,
                              const int64_t *dims, const int64_t *table,
                              int32_t n_col_legs, const int32_t *col_legs,
                              double *out) {
  // src/tensor/block_sparse.h:16-80: rows = non-column legs in order,
  // cols = col_legs in the order given
  std::vector<i64> d, s;
  for (int l = 0; l < rank; ++l) {
    bool is_col = false;
    for (int c = 0; c < n_col_legs; ++c) is_col |= col_legs[c] == l;
    if (!is_col) {
      d.push_back(dims[l]);
      s.push_back(table[l]);
    }
  }
  for (int c = 0; c < n_col_legs; ++c) {
    d.push_back(dims[col_legs[c]]);
    s.push_back(table[col_legs[c]]);
  }
  gather(block, rank, d.data(), s.data(), out);
}

int32_t orc_bs_contract_f64(const orc_bs *a, const double *a_data,
                            const orc_bs *b, const double *b_data,
                            int32_t n_legs, const int32_t *legs, double *c,
                            int32_t n_threads) {
  Plan p;
  if (!make_plan(a, b, n_legs, legs, p)) return -1;
  const i64 np = (i64)p.pairs.size();
  std::vector<PairGeom> geo(np);
  i64 off = 0;
  for (i64 q = 0; q < np; ++q) {
    geo[q] = geometry(a, b, p, p.pairs[q]);
    geo[q].c_off = off;
    off += geo[q].m * geo[q].n;
  }
  if (p.out_rank == 0) {
    // rank-0 branch, src/tensor/block_sparse.h:553-579 (D6: the reference
    // decodes with dims as a table; the intended decode is column-major over
    // the intersection dims, which is what the 1x1 GEMM below sums)
    double acc = 0.0;
    for (i64 q = 0; q < np; ++q) {
      const PairGeom &g = geo[q];
      std::vector<double> pa(g.k), pb(g.k);
      gather(a_data + g.a_off, (int)g.a_dims.size(), g.a_dims.data(),
             g.a_str.data(), pa.data());
      gather(b_data + g.b_off, (int)g.b_dims.size(), g.b_dims.data(),
             g.b_str.data(), pb.data());
      for (i64 l = 0; l < g.k; ++l) acc += pa[l] * pb[l];
    }
    c[0] = acc;
    return 0;
  }
#ifdef _OPENMP
  const int nt = n_threads > 0 ? n_threads : 1;
#pragma omp parallel for schedule(dynamic, 1) num_threads(nt)
#endif
  for (i64 q = 0; q < np; ++q) {
    // per pair: sort_into_matrix for A and B (src/tensor/block_sparse.h:
    // 527-539, D1 fixed), C = A * B^T (:541-542), append in pair order
    // (:548-551)
    const PairGeom &g = geo[q];
    std::vector<double> pa(g.m * g.k), pb(g.n * g.k);
    gather(a_data + g.a_off, (int)g.a_dims.size(), g.a_dims.data(),
           g.a_str.data(), pa.data());
    gather(b_data + g.b_off, (int)g.b_dims.size(), g.b_dims.data(),
           g.b_str.data(), pb.data());
    gemm_nt(g.m, g.n, g.k, pa.data(), pb.data(), c + g.c_off);
  }
  (void)n_threads;
  return 0;
}

int32_t orc_bs_permute_f64(const orc_bs *in, const double *data,
                           const int32_t *perm, double *out) {
  // src/tensor/block_sparse.h:608-651: new block dims = dims(perm); element
  // at new index n is old index o with o[perm[k]] = n[k]
  const int r = in->rank;
  i64 off_out = 0;
  std::vector<i64> d(r), s(r);
  for (i64 blk = 0; blk < in->n_blocks; ++blk) {
    const i64 *bg = in->begin + blk * r;
    const i64 *en = in->end + blk * r;
    const i64 *st = in->strides + blk * r;
    i64 n = 1;
    for (int k = 0; k < r; ++k) {
      d[k] = en[perm[k]] - bg[perm[k]] + 1;
      s[k] = st[perm[k]];
      n *= d[k];
    }
    gather(data + in->offsets[blk], r, d.data(), s.data(), out + off_out);
    off_out += n;
  }
  return 0;
}

int32_t orc_dense_permute_f64(int32_t rank, const int64_t *dims,
                              const int64_t *in_strides, const int32_t *perm,
                              const double *in, double *out) {
  // src/tensor/dense.h:325-366
  std::vector<i64> table(rank), d(rank), s(rank);
  orc_index_table(rank, dims, table.data());
  for (int k = 0; k < rank; ++k) {
    d[k] = dims[perm[k]];
    s[k] = in_strides ? in_strides[perm[k]] : table[perm[k]];
  }
  if (rank == 0) {
    out[0] = in[0];
    return 0;
  }
  gather(in, rank, d.data(), s.data(), out);
  return 0;
}

int32_t orc_dense_contract_f64(int32_t rank_a, const int64_t *dims_a,
                               const double *a, int32_t rank_b,
                               const int64_t *dims_b, const double *b,
                               int32_t n_legs, const int32_t *legs, double *c) {
  // src/tensor/dense.h:182-292: for every output element, sum over the
  // contracted index space (column-major over the leg pairs in order)
  std::vector<int> al(n_legs), bl(n_legs);
  std::vector<i64> cd(n_legs);
  for (int k = 0; k < n_legs; ++k) {
    al[k] = legs[2 * k];
    bl[k] = legs[2 * k + 1];
    if (dims_a[al[k]] != dims_b[bl[k]]) return -1;
    cd[k] = dims_a[al[k]];
  }
  std::vector<int> af = free_legs(rank_a, al), bf = free_legs(rank_b, bl);
  std::vector<i64> ta(rank_a), tb(rank_b);
  orc_index_table(rank_a, dims_a, ta.data());
  orc_index_table(rank_b, dims_b, tb.data());
  std::vector<i64> od;
  for (int l : af) od.push_back(dims_a[l]);
  for (int l : bf) od.push_back(dims_b[l]);
  i64 n_out = 1, n_con = 1;
  for (i64 v : od) n_out *= v;
  for (i64 v : cd) n_con *= v;
  std::vector<i64> odig(od.size()), cdig(n_legs);
  for (i64 o = 0; o < n_out; ++o) {
    decode(o, (int)od.size(), od.data(), odig.data());
    i64 base_a = 0, base_b = 0;
    for (size_t q = 0; q < af.size(); ++q) base_a += odig[q] * ta[af[q]];
    for (size_t q = 0; q < bf.size(); ++q)
      base_b += odig[af.size() + q] * tb[bf[q]];
    double acc = 0.0;
    for (i64 x = 0; x < n_con; ++x) {
      decode(x, n_legs, cd.data(), cdig.data());
      i64 ia = base_a, ib = base_b;
      for (int k = 0; k < n_legs; ++k) {
        ia += cdig[k] * ta[al[k]];
        ib += cdig[k] * tb[bl[k]];
      }
      acc += a[ia] * b[ib];
    }
    c[o] = acc;
  }
  return 0;
}

int32_t orc_bs_densify_f64(const orc_bs *t, const double *data, double *out) {
  // query() on every element: sum over all covering blocks
  // (src/tensor/block_sparse.h:373-409)
  const int r = t->rank;
  std::vector<i64> table(r);
  orc_index_table(r, t->dims, table.data());
  i64 total = 1;
  for (int l = 0; l < r; ++l) total *= t->dims[l];
  std::fill(out, out + total, 0.0);
  if (r == 0) {
    if (t->n_blocks == 0) out[0] = data ? data[0] : 0.0;
    return 0;
  }
  std::vector<i64> d(r), dig(r);
  for (i64 blk = 0; blk < t->n_blocks; ++blk) {
    const i64 *bg = t->begin + blk * r;
    const i64 *en = t->end + blk * r;
    const i64 *st = t->strides + blk * r;
    i64 n = 1;
    for (int l = 0; l < r; ++l) {
      d[l] = en[l] - bg[l] + 1;
      n *= d[l];
    }
    for (i64 e = 0; e < n; ++e) {
      decode(e, r, d.data(), dig.data());
      i64 src = t->offsets[blk], dst = 0;
      for (int l = 0; l < r; ++l) {
        src += dig[l] * st[l];
        dst += (dig[l] + bg[l]) * table[l];
      }
      out[dst] += data[src];
    }
  }
  return 0;
}

// util::index_to_indices with a sort index (src/util/space.h:44-61): the
// coordinates of linear index idx under a (possibly soft-transposed) table,
// peeled off from the largest stride down. Among equal strides the legs of
// extent 1 are peeled last, so they receive coordinate 0.
std::vector<int> table_order(int n, const i64 *dims, const i64 *table) {
  std::vector<int> order(n);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int x, int y) {
    if (table[x] != table[y]) return table[x] < table[y];
    return (dims[x] > 1) < (dims[y] > 1);
  });
  return order;
}

void decode_table(i64 idx, int n, const std::vector<int> &order, const i64 *table, i64 *coord) {
  for (int j = n - 1; j > 0; --j) {
    coord[order[j]] = idx / table[order[j]];
    idx %= table[order[j]];
  }
  if (n > 0) coord[order[0]] = idx;
}

std::vector<i64> table_or_compact(int rank, const int64_t *dims, const int64_t *table) {
  std::vector<i64> t(rank);
  i64 acc = 1;
  for (int l = 0; l < rank; ++l) {
    t[l] = table ? table[l] : acc;
    acc *= dims[l];
  }
  return t;
}

int64_t orc_sparse_contract_f64(int32_t rank_a, const int64_t *dims_a, const int64_t *table_a,
                                int64_t nnz_a, const int64_t *idx_a, const double *val_a,
                                int32_t rank_b, const int64_t *dims_b, const int64_t *table_b,
                                int64_t nnz_b, const int64_t *idx_b, const double *val_b,
                                int32_t n_legs, const int32_t *legs, int64_t *out_idx,
                                double *out_val) {
  // src/tensor/sparse.h:244-356, each tensor decoded with its own table (the
  // reference reads A's table for B at :285-286); A-major accumulation order
  std::vector<int> al(n_legs), bl(n_legs);
  for (int k = 0; k < n_legs; ++k) {
    al[k] = legs[2 * k];
    bl[k] = legs[2 * k + 1];
    if (dims_a[al[k]] != dims_b[bl[k]]) return -1;
  }
  std::vector<int> af = free_legs(rank_a, al), bf = free_legs(rank_b, bl);
  std::vector<i64> od;
  for (int l : af) od.push_back(dims_a[l]);
  for (int l : bf) od.push_back(dims_b[l]);
  std::vector<i64> ot(od.size());
  orc_index_table((int)od.size(), od.data(), ot.data());
  const std::vector<i64> ta = table_or_compact(rank_a, dims_a, table_a);
  const std::vector<i64> tb = table_or_compact(rank_b, dims_b, table_b);
  const std::vector<int> oa = table_order(rank_a, dims_a, ta.data());
  const std::vector<int> ob = table_order(rank_b, dims_b, tb.data());
  // B decoded once
  std::vector<i64> dball((size_t)nnz_b * rank_b);
  for (i64 j = 0; j < nnz_b; ++j)
    decode_table(idx_b[j], rank_b, ob, tb.data(), dball.data() + (size_t)j * rank_b);
  std::map<i64, double> acc;
  std::vector<i64> da(rank_a);
  for (i64 i = 0; i < nnz_a; ++i) {
    decode_table(idx_a[i], rank_a, oa, ta.data(), da.data());
    for (i64 j = 0; j < nnz_b; ++j) {
      const i64 *db = dball.data() + (size_t)j * rank_b;
      bool ok = true;
      for (int k = 0; k < n_legs && ok; ++k) ok = da[al[k]] == db[bl[k]];
      if (!ok) continue;
      i64 key = 0;
      size_t q = 0;
      for (int l : af) key += da[l] * ot[q++];
      for (int l : bf) key += db[l] * ot[q++];
      const double prod = val_a[i] * val_b[j];
      acc[key] += prod;
    }
  }
  if (od.empty()) {
    // full contraction: the single element {0: sum} (src/tensor/sparse.h:328-352)
    if (out_idx) {
      out_idx[0] = 0;
      out_val[0] = acc.empty() ? 0.0 : acc.begin()->second;
    }
    return 1;
  }
  if (out_idx) {
    i64 q = 0;
    for (auto &kv : acc) {
      out_idx[q] = kv.first;
      out_val[q] = kv.second;
      ++q;
    }
  }
  return (int64_t)acc.size();
}

int32_t orc_sparse_permute_indices(int32_t rank, const int64_t *dims, const int64_t *table,
                                   const int32_t *perm, int64_t nnz, const int64_t *in,
                                   int64_t *out) {
  // src/tensor/sparse.h:371-405 with each stored index decoded (the
  // reference decodes the storage position i, :391-393)
  const std::vector<i64> t = table_or_compact(rank, dims, table);
  const std::vector<int> order = table_order(rank, dims, t.data());
  std::vector<i64> nd(rank), nt(rank), c(rank);
  for (int k = 0; k < rank; ++k) nd[k] = dims[perm[k]];
  orc_index_table(rank, nd.data(), nt.data());
  for (i64 i = 0; i < nnz; ++i) {
    decode_table(in[i], rank, order, t.data(), c.data());
    i64 o = 0;
    for (int k = 0; k < rank; ++k) o += c[perm[k]] * nt[k];
    out[i] = o;
  }
  return 0;
}

}  // extern "C"
