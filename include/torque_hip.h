/*
 * torque_hip.h -- C-ABI of the MI355X (gfx950) block-sparse tensor
 * contraction library.
 *
 * The reference library (Walter-Feng/Torque) is a header-only C++ template
 * library with no extern "C" surface; its GPU hot path is
 *   torque::gpu::BlockSparseTensor<T>::contract(cublasHandle_t, B, idx)
 *       src/gpu/block_sparse.cuh:558-1026
 *   torque::gpu::BlockSparseTensor<T>::hard_transpose(perm)
 *       src/gpu/block_sparse.cuh:1031-1122
 *   torque::gpu::DenseTensor<T>::hard_transpose(perm) / contract(...)
 *       src/gpu/dense.cuh:561-599 / :439-555
 *   torque::gpu::SparseTensor<T>::contract(...)
 *       src/gpu/sparse.cuh:270-375
 * Each entry point below replaces one of those calls (or one of the
 * library calls they make: cuTT, cuBLAS, thrust) and is what the C++
 * drop-in classes in torque_amd/include/torque/gpu/ bind to.
 *
 * Conventions
 *  - All device pointers are caller-owned unless an entry point says it
 *    allocates (tq_malloc, plans). No entry point calls exit().
 *  - Tensors are column-major: leg 0 is the fastest leg
 *    (src/util/space.h:10-24).
 *  - Block-sparse metadata uses the reference's data model
 *    (src/gpu/block_sparse.cuh:147-178): per block, inclusive begin/end
 *    points, an index table (element strides per leg) and an element offset
 *    into the concatenated data buffer. Metadata arrays are block-major:
 *    entry [b * rank + l] is leg l of block b (one reference umat column per
 *    block).
 *  - A handle serialises its work on one HIP stream; distinct handles may be
 *    used from distinct threads.
 */
#ifndef TORQUE_HIP_H
#define TORQUE_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TQ_MAX_RANK 12

typedef enum tq_status {
  TQ_SUCCESS = 0,
  TQ_ERROR_INVALID_VALUE = 1,      /* bad argument (null pointer, bad size)   */
  TQ_ERROR_DIMENSION_MISMATCH = 2, /* contracted legs have different dims     */
  TQ_ERROR_RANK_MISMATCH = 3,      /* permutation / index length != rank      */
  TQ_ERROR_OUT_OF_RANGE = 4,       /* index beyond the tensor dimension       */
  TQ_ERROR_ALLOCATION = 5,         /* device or host allocation failed        */
  TQ_ERROR_HIP = 6,                /* a HIP runtime call failed               */
  TQ_ERROR_UNSUPPORTED = 7,        /* rank above TQ_MAX_RANK, dtype/precision */
  TQ_ERROR_NOT_INITIALIZED = 8     /* null handle or plan                     */
} tq_status;

/* Element types. TQ_F16 is IEEE binary16 (the reference's `half`
 * instantiation, src/gpu/block_sparse.cu:94-148): its contractions run on
 * FP16 MFMA with FP32 accumulation and write FP16 (the reference's
 * cublasHgemmStridedBatched, src/gpu/block_sparse.cuh:918-929); its rank-0
 * results reduce in FP64 (:991-1003). */
typedef enum tq_dtype {
  TQ_F32 = 0,
  TQ_F64 = 1,
  TQ_F16 = 2
} tq_dtype;

/* Arithmetic used by the grouped GEMM of a contraction plan.
 * TQ_PREC_NATIVE: the data type's own MFMA (FP64 for TQ_F64, FP32 for TQ_F32).
 * TQ_PREC_F32:    FP64 data, products and sums in FP32 MFMA.
 * TQ_PREC_BF16X3: FP64 data split into three BF16 pieces per operand; the six
 *                 leading piece products run on BF16 MFMA with FP32
 *                 accumulation (the dropped terms are below 2^-24 relative).
 * TQ_PREC_BF16X3 and TQ_PREC_F32 write FP64 results. */
typedef enum tq_precision {
  TQ_PREC_NATIVE = 0,
  TQ_PREC_F32 = 1,
  TQ_PREC_BF16X3 = 2
} tq_precision;

typedef struct tq_handle_s *tq_handle_t;
typedef struct tq_plan_s *tq_plan_t;
typedef struct tq_batch_s *tq_batch_t;

/* Host-side view of a block-sparse tensor's metadata
 * (src/gpu/block_sparse.cuh:147-178; src/tensor/block_sparse.h:114-145). */
typedef struct tq_bs_desc {
  int32_t rank;
  int64_t n_blocks;
  const int64_t *dims;    /* [rank]            total dimension               */
  const int64_t *begin;   /* [n_blocks * rank] inclusive begin points        */
  const int64_t *end;     /* [n_blocks * rank] inclusive end points          */
  const int64_t *strides; /* [n_blocks * rank] index tables (element strides) */
  const int64_t *offsets; /* [n_blocks]        element offset of each block  */
} tq_bs_desc;

/* Summary of a contraction plan (sizes of the host arrays to request). */
typedef struct tq_plan_info {
  int32_t out_rank;
  int32_t n_legs;
  int64_t n_pairs;      /* output blocks: matched (A block, B block) pairs, or
                           distinct output blocks of a merged-output plan   */
  int64_t out_n_elem;   /* elements of the output data buffer (>= 1)         */
  double flops;         /* sum over pairs of 2*M*N*K                         */
  double bytes;         /* compulsory bytes: sizeof(T)*(sum nA + nB + nC)    */
  int32_t a_layout;     /* 0 = read in place m-fast, 1 = in place k-fast,
                           2 = packed by the permute kernel                 */
  int32_t b_layout;     /* same for B (n-fast / k-fast / packed)             */
  int64_t n_tiles;      /* MFMA tiles of the current configuration           */
  int64_t n_segments;   /* 64 x 64 segments of the streaming small-K kernel  */
  int64_t n_matched;    /* matched pairs (== n_pairs unless merged output)   */
} tq_plan_info;

/* ---- runtime --------------------------------------------------------- */

/* Create a handle on a HIP device. Replaces cublasCreate in the reference's
 * callers (test/gpu/block_sparse_test.cu:7-8). */
tq_status tq_create(tq_handle_t *handle, int device);
/* Replaces cublasDestroy (test/gpu/block_sparse_test.cu:234). */
tq_status tq_destroy(tq_handle_t handle);
/* Work is issued on `stream` (a hipStream_t; NULL = the default stream). */
tq_status tq_set_stream(tq_handle_t handle, void *stream);
tq_status tq_synchronize(tq_handle_t handle);
const char *tq_status_string(tq_status status);
/* Text of the last error raised on this handle ("" if none). */
const char *tq_last_error(tq_handle_t handle);

/* Device memory helpers used by the C++ classes (reference: cudaMalloc /
 * cudaMemcpy / cudaMemset calls at src/gpu/block_sparse.cuh:180-446).
 * kind: 0 = host->device, 1 = device->host, 2 = device->device. */
tq_status tq_malloc(tq_handle_t handle, void **ptr, size_t bytes);
tq_status tq_free(tq_handle_t handle, void *ptr);
tq_status tq_memcpy(tq_handle_t handle, void *dst, const void *src,
                    size_t bytes, int kind);
tq_status tq_memset_zero(tq_handle_t handle, void *ptr, size_t bytes);

/* ---- permutation (replaces cuTT) ----------------------------------- */

/* Dense permutation: out has dims[perm[k]] on leg k and is compact
 * column-major; element out[n] = in[sum_k n_k * in_strides[perm[k]]].
 * in_strides == NULL means the compact column-major table of dims.
 * Replaces cuttPlan/cuttExecute in DenseTensor::hard_transpose
 * (src/gpu/dense.cuh:561-599; CPU semantics src/tensor/dense.h:325-366). */
tq_status tq_permute(tq_handle_t handle, tq_dtype dtype, int32_t rank,
                     const int64_t *dims, const int64_t *in_strides,
                     const int32_t *perm, const void *in, void *out);

/* Block-sparse permutation: every block is permuted by perm into a compact
 * column-major block; block order is kept and output block b starts at
 * out_offsets[b] (the exclusive prefix sum of block sizes, written by the
 * call into a host array of n_blocks entries, may be NULL).
 * Replaces reshape_kernel + cuTT + reshape_kernel in
 * BlockSparseTensor::hard_transpose (src/gpu/block_sparse.cuh:1031-1122;
 * CPU semantics src/tensor/block_sparse.h:608-651). */
tq_status tq_bs_permute(tq_handle_t handle, tq_dtype dtype,
                        const tq_bs_desc *in_desc, const int32_t *perm,
                        const void *in, void *out, int64_t *out_offsets);

/* Splits every block into prod(divisor) equal sub-blocks (block extent on
 * leg l divided by divisor[l]; a remainder is an error, "current slice does
 * not support residues"). Sub-block j of block b, with j decoded leg 0
 * fastest over the divisor, becomes output block b * prod(divisor) + j; its
 * begin/end points go to out_begin/out_end (host, block-major, may be NULL)
 * and its data is the compact column-major copy of that region of block b.
 * Replaces BlockSparseTensor::slice (src/gpu/block_sparse.cuh:1124-1176),
 * which copies the data buffer unchanged under the new block list. */
tq_status tq_bs_slice(tq_handle_t handle, tq_dtype dtype, const tq_bs_desc *in_desc,
                      const int64_t *divisor, const void *in, void *out,
                      int64_t *out_begin, int64_t *out_end);

/* ---- box copies (leg fusion) ------------------------------------------ */

/* One box of a block copy: the box of extents dims[0..rank) is read at
 * in[src_off + sum_l i_l * src_strides[l]] and written to
 * out[dst_off + sum_l i_l * dst_strides[l]] (element units, strides >= 0,
 * rank 0 = one element). */
typedef struct tq_box {
  int64_t src_off;
  int64_t dst_off;
  int32_t rank;
  int32_t reserved;
  int64_t dims[TQ_MAX_RANK];
  int64_t src_strides[TQ_MAX_RANK];
  int64_t dst_strides[TQ_MAX_RANK];
} tq_box;

typedef struct tq_copy_plan_s *tq_copy_plan_t;

/* Plans the copy of n boxes from a buffer of in_elems elements into one of
 * out_elems elements: each box is checked to lie inside both buffers (an
 * error otherwise), and legs contiguous in both buffers are merged. Boxes
 * must not overlap in out and in must not overlap out (neither is checked).
 * The charge-aware front end fuses groups of U(1) legs into one leg, and
 * splits them again, as such copies: each block of the tensor is one box of
 * a block of the fused tensor. No counterpart in the reference, whose
 * tensors have no leg fusion; the nearest operation is its per-block
 * gather / scatter between strided blocks and a stacked layout
 * (reshape_kernel, src/gpu/block_sparse.cu:9-64). */
tq_status tq_copy_plan_create(tq_handle_t handle, tq_dtype dtype, int64_t n,
                              const tq_box *boxes, int64_t in_elems, int64_t out_elems,
                              tq_copy_plan_t *plan);
/* Runs the planned copies on the handle's stream: no allocation and no
 * synchronisation, so a capture may record it. Elements of out that no box
 * covers are left as they are. */
tq_status tq_copy_execute(tq_copy_plan_t plan, const void *in, void *out);
tq_status tq_copy_plan_destroy(tq_copy_plan_t plan);

/* ---- block-sparse contraction -------------------------------------- */

/* Plan C = contract(A, B, legs): legs is n_legs x 2, row k = (leg of A,
 * leg of B), the reference's `contracting_indices` umat
 * (src/gpu/block_sparse.cuh:558-561). Block matching and all output index
 * maps are built on the device (src/tensor/block_sparse.cpp:6-175 and
 * src/gpu/block_sparse.cuh:620-674,700-732 restated). The output keeps the
 * reference's layout: one block per matched pair, ordered by A block then
 * ascending B block, each block column-major over [A free legs, B free
 * legs]. The plan owns its device memory and may be executed many times on
 * any data with the same block structure. */
tq_status tq_bs_contract_plan(tq_handle_t handle, const tq_bs_desc *a,
                              const tq_bs_desc *b, int32_t n_legs,
                              const int32_t *legs, tq_dtype dtype,
                              tq_precision precision, tq_plan_t *plan);
/* As tq_bs_contract_plan, with the output merged: one output block per
 * distinct output block (pairs with equal begin and end points) holding the
 * sum of its pairs' products, instead of the reference's one block per
 * matched pair (src/tensor/block_sparse.h:548-551,
 * src/gpu/block_sparse.cuh:972-975). Blocks are listed in the order of their
 * first pair, column-major over [A free legs, B free legs]; the dense image
 * equals the unmerged output's. Each output block is one GEMM over the
 * concatenated contracted ranges of its pairs (operands packed per block);
 * when every pair has K <= 8 and no output block has more than two pairs
 * (not at TQ_PREC_BF16X3) the blocks are instead streamed from the unmerged
 * plan's operand panels, summing the pairs' products in the same order (the
 * same bits); such a plan cannot join a batch (tq_batch_create refuses it).
 * When every pair writes an output block of its own (a block-diagonal
 * product, for one) the merged output is the unmerged one and the plan is
 * the unmerged plan, operand layouts included.
 * tq_plan_get_pairs still lists the matched pairs (tq_plan_info.n_matched);
 * tq_plan_get_output and the worklist list the merged blocks. */
tq_status tq_bs_contract_plan_merged(tq_handle_t handle, const tq_bs_desc *a,
                                     const tq_bs_desc *b, int32_t n_legs,
                                     const int32_t *legs, tq_dtype dtype,
                                     tq_precision precision, tq_plan_t *plan);
/* tq_bs_contract_plan (merged = 0) or tq_bs_contract_plan_merged (merged =
 * 1) through a cache kept by the handle: a call whose inputs (dtype,
 * precision, legs and both tensors' metadata, compared byte for byte) match
 * a cached plan returns that plan without planning. The handle owns the plan:
 * the caller must not destroy it; tq_destroy releases the cache, which keeps
 * the 32 most recently used plans. No reference counterpart: the reference
 * re-derives every map on each contract call
 * (src/gpu/block_sparse.cuh:570-732). */
tq_status tq_bs_contract_plan_cached(tq_handle_t handle, const tq_bs_desc *a,
                                     const tq_bs_desc *b, int32_t n_legs,
                                     const int32_t *legs, tq_dtype dtype,
                                     tq_precision precision, int32_t merged,
                                     tq_plan_t *plan);
tq_status tq_plan_get_info(tq_plan_t plan, tq_plan_info *info);
/* Output metadata, host arrays: dims[out_rank], begin/end/strides
 * [n_pairs*out_rank] (block-major), offsets[n_pairs]. Any may be NULL. */
tq_status tq_plan_get_output(tq_plan_t plan, int64_t *dims, int64_t *begin,
                             int64_t *end, int64_t *strides,
                             int64_t *offsets);
/* Integer maps per matched pair, host arrays of n_matched entries (any may be
 * NULL): A block, B block, A sub-block element offset, B sub-block element
 * offset, GEMM shape M, N, K. These are the quantities the reference computes
 * in src/gpu/block_sparse.cuh:700-732 and src/tensor/block_sparse.h:516-525. */
tq_status tq_plan_get_pairs(tq_plan_t plan, int64_t *a_block, int64_t *b_block,
                            int64_t *a_offset, int64_t *b_offset, int64_t *m,
                            int64_t *n, int64_t *k);
/* Order in which the grouped GEMM visits the pairs (the size-sorted work
 * list): a permutation of 0..n_pairs-1, largest estimated cost first. */
tq_status tq_plan_get_worklist(tq_plan_t plan, int64_t *order);
/* Run the plan: C receives out_n_elem elements. A rank-0 output (full
 * contraction) is the network scalar, reduced on the GPU in FP64. The plan's
 * device state (operand panels, split-K workspace, the tile counters of its
 * GEMM launches) serves one execute at a time: executes of one plan on
 * different streams must not overlap. */
tq_status tq_bs_contract_execute(tq_plan_t plan, const void *a, const void *b,
                                 void *c);
/* The two stages of tq_bs_contract_execute, for callers that time or
 * schedule them separately: pack permutes the operand blocks that are not
 * already in GEMM layout into the plan's panels (the reference's
 * reshape_kernel + cuTT steps, src/gpu/block_sparse.cuh:755-865); multiply
 * runs the grouped GEMM (or the rank-0 reduction) and must follow a pack of
 * the same A and B. */
tq_status tq_bs_contract_pack(tq_plan_t plan, const void *a, const void *b);
tq_status tq_bs_contract_multiply(tq_plan_t plan, const void *a, const void *b,
                                  void *c);
/* MFMA pairs (K above the small-K threshold) fall in two tile classes, each
 * launched with its own tile configuration: class 1 holds the pairs with
 * M >= 128 and N >= 128, class 0 the rest. set_config selects the
 * configuration of both classes by index (see tq_gemm_config_count),
 * set_class_config of one; -1 restores the plan's automatic choice.
 * get_config writes the two classes' current configurations. */
tq_status tq_plan_set_config(tq_plan_t plan, int32_t config);
tq_status tq_plan_set_class_config(tq_plan_t plan, int32_t cls, int32_t config);
tq_status tq_plan_get_config(tq_plan_t plan, int32_t *config /* [2] */);
/* Split-K of one MFMA tile class: each pair's contracted range is cut into
 * up to `split` slices (1..64; ceil(K / split) rounded up to 32 values, so
 * pairs with K <= 32 stay whole) that run as GEMMs of their own into a
 * plan-owned workspace; a reduction adds each pair's slices into C in slice
 * order, so results repeat bit for bit (and differ from split = 1 in the
 * last bits). A new plan starts with an analytic split: 1, unless a class's
 * tiles would fill less than half of the GPU's tile slots, then the fewest
 * power-of-two slices (at most 16, each at least 256 k values) that fill
 * them. tq_plan_autotune times splits 1, 2, 4, 8 and 16 on each
 * class's chosen configuration and keeps the fastest; tq_plan_choose_config
 * predicts them. get_split writes both classes' splits. No reference
 * counterpart (src/gpu/block_sparse.cuh:894-930 runs K whole). */
tq_status tq_plan_set_class_split(tq_plan_t plan, int32_t cls, int32_t split);
tq_status tq_plan_get_split(tq_plan_t plan, int32_t *split /* [2] */);
tq_status tq_plan_destroy(tq_plan_t plan);

/* Pairs whose contracted extent K is at most max_k (0..8) are written by the
 * streaming small-K kernel (2K flops per output element: bound by writing C)
 * instead of MFMA tiles; 0 sends every pair to the MFMA GEMM. The default is
 * 8. The split-BF16 tier streams such pairs from its split panels (each
 * value rebuilt as hi + mid + lo in FP64, FP64 products). */
tq_status tq_plan_set_small_k(tq_plan_t plan, int32_t max_k);

/* Autotuner (no reference counterpart: the reference measures a cuTT plan
 * per call, src/gpu/block_sparse.cuh:836-863). Packs A and B once, then, for
 * each tile class, times `reps` grouped-GEMM launches of every tile
 * configuration with HIP events on the handle's stream and keeps the fastest
 * for that class. A split-BF16 plan in the automatic operand form (see
 * tq_plan_set_split_operands) is swept in both forms and keeps the form whose
 * whole execute (pack and multiply) is faster. times_ms (host, 2 * tq_gemm_config_count(precision)
 * entries, class 0 first, may be NULL) receives the mean time of each
 * (class, configuration), 0 for a class without tiles and -1 for a
 * configuration the plan cannot run (wave-tile configurations need leading
 * dimensions below 2^22); best (host, 2 entries, may be NULL) the chosen
 * configuration of each class. C is written by the timed launches. */
tq_status tq_plan_autotune(tq_plan_t plan, const void *a, const void *b,
                           void *c, int32_t reps, int32_t *best,
                           float *times_ms);

/* Operand form of a split-BF16 plan whose operands are both plain matrices
 * (m or k fastest with one leading dimension): it can read them in place,
 * each FP64 value split in registers as the GEMM stages it (no pack pass,
 * but a value is split once per tile that reads it), or pack them once into
 * split panels. form: 0 = automatic (in place until tq_plan_autotune times a
 * configuration sweep in both forms and keeps the faster whole execute),
 * 1 = split panels, 2 = in place (TQ_ERROR_UNSUPPORTED when an operand is
 * not plain). Both forms give the same bits. tq_plan_info's a_layout and
 * b_layout read 2 for panels. Releases a captured graph. No reference
 * counterpart. */
tq_status tq_plan_set_split_operands(tq_plan_t plan, int32_t form);

/* Input-domain check of the split-BF16 tier: *out_of_range = 1 when an FP64
 * operand value split by this plan's executes since the last check (or since
 * planning) lay outside the domain in which hi + mid + lo represents it to
 * 2^-26: |x| >= 2^127 (hi overflows), 0 < |x| < 2^-100 (lo loses bits),
 * NaN or infinity (they turn into NaN; the FP64 tier keeps infinities).
 * The pack and GEMM kernels set a device flag as they split; this call
 * reads and clears it (synchronises the handle's stream). Always 0 for the
 * other tiers. No reference counterpart. */
tq_status tq_plan_split_range_check(tq_plan_t plan, int32_t *out_of_range);

/* ---- batches and graphs (no reference counterpart) ------------------- */

/* Joins n plans into one batch plan: every pair of every plan goes into one
 * cost-sorted work list, so tq_bs_contract_execute(batch, a, b, c) packs all
 * operands with one launch per side and multiplies every pair with one
 * streaming launch and one grouped-GEMM launch per tile class. Plan i reads A at element
 * offset a_offset[i] of `a`, B at b_offset[i] of `b`, and writes its output
 * blocks at c_offset[i] of `c`. The plans must share dtype, precision and
 * small-K threshold and have tensor (rank > 0) results, and none may be a
 * merged plan that streams its output blocks (tq_bs_contract_plan_merged);
 * they may be destroyed after the call. The batch supports execute,
 * autotune, set_config, capture and destroy. */
tq_status tq_batch_create(tq_handle_t handle, int32_t n, const tq_plan_t *plans,
                          const int64_t *a_offset, const int64_t *b_offset,
                          const int64_t *c_offset, tq_plan_t *batch);

/* Captures tq_bs_contract_execute(plan, a, b, c) into a HIP graph owned by
 * the plan (the pointers are baked in); tq_plan_graph_launch replays it on
 * the handle's stream with no host work per replay. tq_plan_set_config,
 * tq_plan_set_class_config, tq_plan_set_small_k and tq_plan_autotune
 * release the captured graph (they may rebuild the work lists it reads):
 * tq_plan_graph_launch then returns TQ_ERROR_NOT_INITIALIZED until the plan
 * is captured again. */
tq_status tq_plan_capture(tq_plan_t plan, const void *a, const void *b, void *c);
tq_status tq_plan_graph_launch(tq_plan_t plan);

/* Records a sequence of calls into one HIP graph: between tq_capture_begin
 * and tq_capture_end the handle's work is captured on a stream of the
 * handle's own instead of running, and tq_capture_end returns the graph
 * (the pointers each call received are baked in). Only calls that neither
 * allocate nor synchronise may run in between: tq_bs_contract_execute /
 * _pack / _multiply of plans executed at least once before (their tile
 * lists exist), tq_scale_f64, tq_normalize_f64_noalloc; tq_set_stream is
 * refused. A call that cannot be captured makes tq_capture_end fail.
 * tq_graph_launch replays the graph on the handle's stream. Used for a whole
 * CTMRG step (absorb, projector contractions, FP64 normalisation). No
 * reference counterpart. */
typedef struct tq_graph_s *tq_graph_t;
tq_status tq_capture_begin(tq_handle_t handle);
tq_status tq_capture_end(tq_handle_t handle, tq_graph_t *graph);
tq_status tq_graph_launch(tq_graph_t graph, tq_handle_t handle);
tq_status tq_graph_destroy(tq_graph_t graph);

/* Number of grouped-GEMM tile configurations and a description of one:
 * the tile's M and N extents, its k step, wave_tile = 1 when one
 * wavefront (not a workgroup) computes a tile, and min_waves, the
 * wavefronts per SIMD the kernel keeps resident (2 for the variants whose
 * registers fit two, or whose eight-wave workgroup places two on each SIMD,
 * else 1). Any output pointer may be NULL. */
int32_t tq_gemm_config_count(tq_precision precision);
tq_status tq_gemm_config_describe(tq_precision precision, int32_t config,
                                  int32_t *bm, int32_t *bn, int32_t *bk,
                                  int32_t *wave_tile, int32_t *min_waves);

/* ---- dense contraction --------------------------------------------- */

/* Dense C = contract(A, B, legs) with compact column-major operands;
 * C has dims [A free legs, B free legs]. Replaces the permute + cublasXgemm
 * path of DenseTensor::contract (src/gpu/dense.cuh:439-555). The one-block
 * plan comes from the handle's plan cache (tq_bs_contract_plan_cached), so
 * repeated shapes plan once; returns after C is written. */
tq_status tq_dense_contract(tq_handle_t handle, tq_dtype dtype, int32_t rank_a,
                            const int64_t *dims_a, int32_t rank_b,
                            const int64_t *dims_b, int32_t n_legs,
                            const int32_t *legs, const void *a, const void *b,
                            void *c);

/* ---- element-sparse tensors ---------------------------------------- */

/* An element-sparse tensor (src/gpu/sparse.cuh:84-441): nnz stored elements,
 * element i at linear index indices[i] under the index table `table`
 * (element strides per leg; a soft-transposed tensor carries a permuted
 * table, src/gpu/sparse.cuh:380-392). */
typedef struct tq_sp_desc {
  int32_t rank;
  const int64_t *dims;    /* [rank] host: dimension                           */
  const int64_t *table;   /* [rank] host: index table; NULL = compact table   */
  int64_t nnz;
  const int64_t *indices; /* [nnz] device: linear index of each element       */
  const void *values;     /* [nnz] device: values (dtype)                     */
} tq_sp_desc;

/* C = contract(A, B, legs) of element-sparse tensors (legs n_legs x 2, row
 * k = (leg of A, leg of B)). Replaces SparseTensor<T>::contract
 * (src/gpu/sparse.cuh:270-375 and its handle_indices kernel,
 * src/gpu/sparse.cu:88-194; CPU semantics src/tensor/sparse.h:244-356).
 * C has dims [A free legs, B free legs] (written to out_dims, host, may be
 * NULL), the compact column-major index table, and holds every output
 * index that receives a product, in ascending order, each with the sum of
 * its products (FP64 accumulation); a rank-0 result is the one element
 * {index 0: sum}. The library allocates *out_indices (int64) and
 * *out_values (dtype) on the device (at least one element each); the
 * caller releases both with tq_free. Synchronous. */
tq_status tq_sparse_contract(tq_handle_t handle, tq_dtype dtype, const tq_sp_desc *a,
                             const tq_sp_desc *b, int32_t n_legs, const int32_t *legs,
                             int64_t *out_dims, int64_t *out_nnz, int64_t **out_indices,
                             void **out_values);

/* Re-encodes nnz linear indices (device) of a tensor with dims/table (host,
 * table NULL = compact) under the compact table of the permuted dimension
 * dims[perm[k]]: the index part of SparseTensor::hard_transpose
 * (src/gpu/sparse.cuh:398-432; the values keep their order). in_indices
 * and out_indices may alias. */
tq_status tq_sparse_permute_indices(tq_handle_t handle, int32_t rank, const int64_t *dims,
                                    const int64_t *table, const int32_t *perm, int64_t nnz,
                                    const int64_t *in_indices, int64_t *out_indices);

/* ---- FP64 reductions ------------------------------------------------ */

/* result = sum_i x[i] * y[i] in FP64 on the GPU (reference: cublasDdot +
 * add<<<1,1>>>, src/gpu/block_sparse.cuh:979-1010). Writes *result on the
 * host after synchronising the handle's stream. */
tq_status tq_dot_f64(tq_handle_t handle, int64_t n, const double *x,
                     const double *y, double *result);
/* result = sqrt(sum x[i]^2), FP64 on the GPU. */
tq_status tq_nrm2_f64(tq_handle_t handle, int64_t n, const double *x,
                      double *result);
/* x <- x * alpha on the GPU (normalisation of environment tensors). */
tq_status tq_scale_f64(tq_handle_t handle, int64_t n, double *x, double alpha);
/* Normalisation of an environment tensor between CTMRG steps, entirely on
 * the device: *norm (a device pointer) receives sqrt(sum x[i]^2), reduced in
 * FP64 in the fixed order of tq_nrm2_f64, and x <- x / *norm with the norm
 * read on the device (x is unchanged when the norm is 0). No reference
 * counterpart: the reference's PEPS tests normalise on the host
 * (test/gpu/peps_test.cu:798-842). Allocates its partial sums and
 * synchronises the handle's stream before returning. */
tq_status tq_normalize_f64(tq_handle_t handle, int64_t n, double *x, double *norm);
/* tq_normalize_f64 without any allocation, for callers that cannot allocate
 * device memory (code under HIP graph capture, serving loops). No reference
 * counterpart. The partial sums live in a static device array of 1024 FP64
 * values per GPU (8 KiB, held until the process ends; every n fits, each
 * partial covering a 1/1024 share of x). Stream-ordered: it returns without
 * synchronising. All calls on one GPU share that array, so calls that
 * overlap on one GPU (other streams or threads) race. Results are bitwise
 * those of tq_normalize_f64. */
tq_status tq_normalize_f64_noalloc(tq_handle_t handle, int64_t n, double *x, double *norm);

/* ---- learned tile-configuration cost model (no reference counterpart) -- */

/* A multilayer perceptron (TQ_COST_FEATURES inputs, two hidden layers of
 * TQ_COST_HIDDEN ReLU units, one output) predicting the log of a tile
 * class's grouped-GEMM time under one configuration. It is trained on the
 * GPU (full-batch Adam on the mean squared error) by the library's own
 * forward, backward and update kernels, on timings the autotuner records;
 * the autotuner stays the fallback and measures the model's regret. */
#define TQ_COST_FEATURES 34
#define TQ_COST_HIDDEN 64
typedef struct tq_costmodel_s *tq_costmodel_t;

/* Features of a set of block pairs (M, N, K per pair) under one tile
 * configuration, host-only: [0,8) [8,16) [16,24) flop-weighted histograms
 * of M, N, K over the buckets <16, 16-31, ..., 512-1023, >=1024; 24 log10
 * flops; 25 log10 pairs; 26 log10 tiles; 27 useful / padded flops; 28-30
 * log2 of the tile's M, N, K extents; 31 2 * wave_tile + min_waves (as
 * tq_gemm_config_describe reports them); 32 log10(tiles per concurrently
 * resident tile); 33 log10 output bytes. All zero for an empty set.
 * n_cu <= 0 means 256. */
tq_status tq_cost_features(const int64_t *m, const int64_t *n, const int64_t *k,
                           int64_t n_pairs, tq_precision precision, int32_t config,
                           int32_t n_cu, float *features);
/* The same for the pairs of MFMA tile class cls of a plan. */
tq_status tq_plan_features(tq_plan_t plan, int32_t cls, int32_t config, float *features);

int64_t tq_costmodel_n_params(void);
/* He-normal weights from `seed`, zero biases. */
tq_status tq_costmodel_create(tq_handle_t handle, uint64_t seed, tq_costmodel_t *model);
tq_status tq_costmodel_destroy(tq_costmodel_t model);
/* Parameters as one float array (tq_costmodel_n_params entries): W1
 * [features][hidden], b1, W2 [hidden][hidden], b2, w3, b3. */
tq_status tq_costmodel_get_params(tq_costmodel_t model, float *params);
tq_status tq_costmodel_set_params(tq_costmodel_t model, const float *params);
/* The normalisation a fit sets (identity until then): x_mean and x_std
 * (TQ_COST_FEATURES floats each) and y = {mean, std} of the log-times; with
 * the parameters they save and restore a trained model. */
tq_status tq_costmodel_get_scaling(tq_costmodel_t model, float *x_mean, float *x_std,
                                   float *y);
tq_status tq_costmodel_set_scaling(tq_costmodel_t model, const float *x_mean,
                                   const float *x_std, const float *y);
/* Gradient (tq_costmodel_n_params floats, may be NULL) and value of the
 * mean squared error of the raw network over n samples (x row-major
 * [n][TQ_COST_FEATURES], y [n]), computed on the GPU. */
tq_status tq_costmodel_gradient(tq_costmodel_t model, const float *x, const float *y,
                                int64_t n, float *grad, float *loss);
/* Trains on (x, y) for `steps` Adam steps (learning rate lr) after setting
 * the input and target normalisation from the data; *loss (may be NULL)
 * receives the final normalised mean squared error. */
tq_status tq_costmodel_fit(tq_costmodel_t model, const float *x, const float *y, int64_t n,
                           int32_t steps, float lr, float *loss);
/* Predictions for n samples; raw = 1 skips the normalisation (network
 * output for the given inputs), raw = 0 applies the fit's. */
tq_status tq_costmodel_predict(tq_costmodel_t model, const float *x, int64_t n, int32_t raw,
                               float *y);
/* Sets each tile class of the plan to the configuration and split-K with
 * the least predicted time (a split class is predicted as the grouped GEMM
 * of its slices plus the reduction's traffic); best (may be NULL) receives
 * the two configurations. */
tq_status tq_plan_choose_config(tq_plan_t plan, tq_costmodel_t model, int32_t *best);

#ifdef __cplusplus
}
#endif

#endif /* TORQUE_HIP_H */
