/*
 * oracle.h -- CPU restatement of Walter-Feng/Torque's block-sparse tensor
 * algorithms. TEST INFRASTRUCTURE ONLY: the tests, the smoke check and the
 * benchmark's CPU-baseline leg call it as the checker; no product code links
 * it.
 *
 * Every function cites the reference file:line it restates. Where the
 * reference has a defect the intended semantics are used and the defect is
 * named (D1: wrong index table in CPU contract; D6: dims passed where a
 * stride table is expected in the rank-0 branch).
 *
 * Data model (src/tensor/block_sparse.h:114-145): metadata arrays are
 * block-major, entry [b*rank + l] is leg l of block b.
 */
#ifndef TORQUE_ORACLE_H
#define TORQUE_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_bs {
  int32_t rank;
  int64_t n_blocks;
  const int64_t *dims;
  const int64_t *begin;
  const int64_t *end;
  const int64_t *strides;
  const int64_t *offsets;
} orc_bs;

/* util::generate_index_table, src/util/space.h:10-24 */
void orc_index_table(int32_t rank, const int64_t *dims, int64_t *table);
/* util::index_to_indices (both overloads), src/util/space.h:26-59.
 * sort_index may be NULL (then the identity order is used). */
void orc_index_to_indices(int64_t i, int32_t rank, const int64_t *table,
                          const int64_t *sort_index, int64_t *indices);
/* util::nest_sum, src/util/space.h:76-88 */
void orc_nest_sum(int64_t n, const int64_t *in, int64_t init, int64_t *out);

/* Number of matched (A block, B block) pairs of contract(A, B, legs):
 * block_sparse::block_in_range over every A block,
 * src/tensor/block_sparse.cpp:6-175. Returns -1 on a dimension mismatch. */
int64_t orc_bs_count_pairs(const orc_bs *a, const orc_bs *b, int32_t n_legs,
                           const int32_t *legs);

/* Integer maps of contract(A, B, legs) in the reference's output order
 * (src/gpu/block_sparse.cuh:620-674, 700-732; src/tensor/block_sparse.h:
 * 455-551). Arrays have n_pairs (or n_pairs*out_rank) entries.
 * Returns the number of pairs, or -1 on a dimension mismatch. */
int64_t orc_bs_plan(const orc_bs *a, const orc_bs *b, int32_t n_legs,
                    const int32_t *legs, int64_t *a_block, int64_t *b_block,
                    int64_t *out_begin, int64_t *out_end, int64_t *a_offset,
                    int64_t *b_offset, int64_t *m, int64_t *n, int64_t *k,
                    int64_t *out_offsets);

/* BlockSparseTensor::contract, src/tensor/block_sparse.h:415-584 (D1 fixed:
 * each operand uses its own block's index table). Writes the output data
 * (out_n_elem values, the blocks of orc_bs_plan in order; a rank-0 result is
 * one value). n_threads > 1 runs pairs in parallel. Returns 0 or -1. */
int32_t orc_bs_contract_f64(const orc_bs *a, const double *a_data,
                            const orc_bs *b, const double *b_data,
                            int32_t n_legs, const int32_t *legs, double *c,
                            int32_t n_threads);

/* block_sparse::sort_into_matrix, src/tensor/block_sparse.h:16-80: pack a
 * strided (sub)block into an n_row x n_col column-major matrix, rows = legs
 * not in col_legs (original order), cols = col_legs in the given order. */
void orc_sort_into_matrix_f64(const double *block, int32_t rank,
                              const int64_t *dims, const int64_t *table,
                              int32_t n_col_legs, const int32_t *col_legs,
                              double *out);

/* BlockSparseTensor::hard_transpose, src/tensor/block_sparse.h:608-651
 * (general strided gather, as src/gpu/block_sparse.cu:9-64 reads blocks).
 * Output blocks are compact column-major in the same order. */
int32_t orc_bs_permute_f64(const orc_bs *in, const double *data,
                           const int32_t *perm, double *out);

/* DenseTensor::hard_transpose, src/tensor/dense.h:325-366: out dims[perm[k]]
 * on leg k; in_strides NULL means compact column-major. */
int32_t orc_dense_permute_f64(int32_t rank, const int64_t *dims,
                              const int64_t *in_strides, const int32_t *perm,
                              const double *in, double *out);

/* DenseTensor::contract, src/tensor/dense.h:182-292 (brute force). */
int32_t orc_dense_contract_f64(int32_t rank_a, const int64_t *dims_a,
                               const double *a, int32_t rank_b,
                               const int64_t *dims_b, const double *b,
                               int32_t n_legs, const int32_t *legs, double *c);

/* BlockSparseTensor::query summed over every covering block,
 * src/tensor/block_sparse.h:373-409, applied to every element: the dense
 * image of a block-sparse tensor (n = prod(dims) values). */
int32_t orc_bs_densify_f64(const orc_bs *t, const double *data, double *out);

/* SparseTensor::contract, src/tensor/sparse.h:244-356: element-sparse
 * tensors given by linear indices under their index tables (NULL = compact
 * column-major). Output indices are sorted and unique; a rank-0 result is
 * the one element {0: sum}. Returns the output nnz, or -1 on a dimension
 * mismatch. When out_idx is NULL only the count is returned. */
int64_t orc_sparse_contract_f64(int32_t rank_a, const int64_t *dims_a, const int64_t *table_a,
                                int64_t nnz_a, const int64_t *idx_a, const double *val_a,
                                int32_t rank_b, const int64_t *dims_b, const int64_t *table_b,
                                int64_t nnz_b, const int64_t *idx_b, const double *val_b,
                                int32_t n_legs, const int32_t *legs, int64_t *out_idx,
                                double *out_val);

/* Index part of SparseTensor::hard_transpose, src/tensor/sparse.h:371-405:
 * each index decoded under (dims, table) and re-encoded under the compact
 * table of dims[perm[k]]. */
int32_t orc_sparse_permute_indices(int32_t rank, const int64_t *dims, const int64_t *table,
                                   const int32_t *perm, int64_t nnz, const int64_t *in,
                                   int64_t *out);

#ifdef __cplusplus
}
#endif

#endif
