// tq_common.h -- internal declarations shared by the HIP sources of
// libtorque_hip.so (handle, error plumbing, device helpers).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "torque_hip.h"

using i64 = int64_t;
using u64 = uint64_t;

// A plan kept by the handle for tq_bs_contract_plan_cached: the key holds
// every input of the plan (flags, legs and both tensors' metadata).
struct PlanCacheEntry {
  std::string key;
  tq_plan_t plan = nullptr;
  uint64_t last_use = 0;
};

struct tq_handle_s {
  int device = 0;
  hipStream_t stream = nullptr;
  int n_cu = 256;
  std::string last_error;
  // pinned staging buffer for small host<->device transfers of metadata
  void *pinned = nullptr;
  size_t pinned_bytes = 0;
  // plans of tq_bs_contract_plan_cached, released by tq_destroy
  std::vector<PlanCacheEntry> plan_cache;
  uint64_t plan_clock = 0;
  // tq_capture_begin: work is recorded on capture_stream; the caller's
  // stream comes back at tq_capture_end
  hipStream_t capture_stream = nullptr;
  hipStream_t saved_stream = nullptr;
  bool capturing = false;
};

// Bytes of one element of a data type.
inline size_t tq_elem_size(tq_dtype dt) {
  return dt == TQ_F64 ? 8 : dt == TQ_F16 ? 2 : 4;
}
inline bool tq_valid_dtype(tq_dtype dt) {
  return dt == TQ_F32 || dt == TQ_F64 || dt == TQ_F16;
}

// Records an error message on the handle and returns the status.
tq_status tq_fail(tq_handle_t h, tq_status s, const std::string &msg);

// Checks a HIP call; on failure records the HIP error text and returns
// TQ_ERROR_HIP from the enclosing function.
#define TQ_HIP(h, expr)                                                     \
  do {                                                                      \
    hipError_t tq_e_ = (expr);                                              \
    if (tq_e_ != hipSuccess)                                                \
      return tq_fail((h), TQ_ERROR_HIP,                                     \
                     std::string(#expr) + ": " + hipGetErrorString(tq_e_)); \
  } while (0)

#define TQ_CHECK(expr)                  \
  do {                                  \
    tq_status tq_s_ = (expr);           \
    if (tq_s_ != TQ_SUCCESS) return tq_s_; \
  } while (0)

// Device buffer owned by a plan or a call; freed on destruction.
struct DevBuf {
  void *p = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
  hipError_t alloc(size_t b) {
    if (p && bytes >= b) return hipSuccess;
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
    if (b == 0) b = 8;
    hipError_t e = hipMalloc(&p, b);
    if (e == hipSuccess) bytes = b;
    return e;
  }
  template <class T>
  T *as() const {
    return reinterpret_cast<T *>(p);
  }
};

// Pinned host staging for metadata uploads (grown on demand).
tq_status tq_stage(tq_handle_t h, size_t bytes, void **host_ptr);

__host__ __device__ inline int64_t ceil_div(int64_t a, int64_t b) {
  return (a + b - 1) / b;
}

// ---- permutation / gather jobs (tq_permute.hip) -------------------------

// One strided-gather job: dst[dst_off + e] = src[src_off + sum_l d_l * s_l]
// where e is the column-major linear index over dims (leg 0 fastest).
struct GemmDesc;
struct RunDesc;

// For the split-BF16 pack (kdim > 0) the first nk legs are the contracted
// ones (they multiply to kdim) and the rest the rows (n_elem / kdim of
// them); each of the three BF16 planes is k-chunked: value (row, k) lands at
// dst_off + (k / 8) * rows * 8 + row * 8 + k % 8, with k padded with zeros
// to kpad (a multiple of 8), so the 8 values of one row in one chunk form a
// 16-byte vector.
struct CopyJob {
  i64 src_off;
  i64 dst_off;
  i64 n_elem;
  int32_t rank;
  int32_t kdim;
  int32_t kpad;
  int32_t nk;
  int32_t dims[TQ_MAX_RANK];
  i64 strides[TQ_MAX_RANK];
  // ceil(2^32 / dims[l]) (0 for a unit leg): the kernels split an index by
  // a leg with a multiply-high and one correction (job_div) instead of an
  // integer division
  uint32_t magic[TQ_MAX_RANK];
  // plain gathers only: dst_strided = 1 writes element e to dst_off +
  // sum_l i_l * dst_strides[l] instead of dst_off + e (a pair's slice of a
  // k-fast group panel, whose rows are longer than the pair's K)
  int32_t dst_strided;
  int32_t pad;
  i64 dst_strides[TQ_MAX_RANK];
};

// Last step of building any CopyJob. A plain gather (kdim == 0) drops its
// unit legs and merges each leg into the previous one when it continues it
// in the source and in the destination (a compact destination always
// continues): a block moved into a panel with its legs regrouped decodes
// few legs. Then the reciprocal of every leg.
__host__ __device__ inline void finalize_job(CopyJob &j) {
  if (j.kdim <= 0) {
    int r = 0;
    for (int l = 0; l < j.rank; ++l) {
      const int32_t d = j.dims[l];
      if (d == 1) continue;
      if (r > 0 && j.strides[r - 1] * (i64)j.dims[r - 1] == j.strides[l] &&
          (!j.dst_strided || j.dst_strides[r - 1] * (i64)j.dims[r - 1] == j.dst_strides[l]) &&
          (i64)j.dims[r - 1] * d < (i64(1) << 31)) {
        j.dims[r - 1] *= d;
        continue;
      }
      j.dims[r] = d;
      j.strides[r] = j.strides[l];
      j.dst_strides[r] = j.dst_strides[l];
      ++r;
    }
    for (int l = r; l < j.rank; ++l) {
      j.dims[l] = 1;
      j.strides[l] = 0;
      j.dst_strides[l] = 0;
    }
    j.rank = r;
  }
  for (int l = 0; l < TQ_MAX_RANK; ++l) {
    const uint64_t d = l < j.rank && j.dims[l] > 1 ? (uint64_t)j.dims[l] : 0;
    j.magic[l] = d ? (uint32_t)((((uint64_t)1 << 32) + d - 1) / d) : 0u;
  }
}

// e / d for any 32-bit e from m = ceil(2^32 / d) (m = 0: d = 1): the
// multiply-high exceeds the quotient by at most one.
__host__ __device__ inline uint32_t job_div(uint32_t e, uint32_t d, uint32_t m) {
  if (m == 0) return e;
  const uint32_t q = (uint32_t)(((uint64_t)e * m) >> 32);
  return (uint64_t)q * d > e ? q - 1 : q;
}

// Work items of the pack kernels per job, one wavefront each: chunks of
// kGatherPerWG elements for a plain gather. A split-BF16 job is cut into its
// (row, 8-value chunk) cells, numbered rows fastest, kSplitCells per item;
// when its first contracted leg is contiguous in memory (k fastest) into
// units of kSplitRows rows x one chunk instead, numbered chunks fastest, so
// that lanes read whole cache lines of a row and the chunks of one row band
// stay in neighbouring items. A workgroup thus covers up to four small
// blocks at once.
constexpr int kSplitCells = 64;
constexpr int kSplitRows = 64;
__host__ __device__ inline bool split_kfast(const CopyJob &j) {
  return j.nk > 0 && j.strides[0] == 1 && j.dims[0] > 1;
}
__host__ __device__ inline int64_t pack_job_wgs(const CopyJob &j);

// Split-BF16 pack: like tq_launch_gather_jobs, reading FP64 and writing the
// three BF16 pieces of each value to planes dst, dst + plane, dst + 2*plane.
// range_flag (may be null) is or-ed with 1 when a value lies outside the
// split's domain (tqs::split_out_of_range).
tq_status tq_launch_split_jobs(tq_handle_t h, const CopyJob *jobs,
                               const int32_t *job_wg_start, int64_t n_jobs,
                               int64_t total_wg, const double *src,
                               uint16_t *dst, int64_t plane, int32_t *range_flag);
// Split-BF16 grouped GEMM. direct = false: a and b are the split panels
// (planes a_plane / b_plane elements apart); direct = true: a and b are the
// FP64 operands read in place (the descriptors' offsets, leading dimensions
// and k-fast flags address them) and split on the way to LDS. tile_ctr (may
// be null): kTileCounters device words the single-buffered configurations use
// to hand out tiles per XCD, as in tq_launch_grouped_gemm.
tq_status tq_launch_bf16x3_gemm(tq_handle_t h, int cfg, bool direct, const GemmDesc *descs,
                                const RunDesc *runs, const int32_t *tile_pair,
                                const int32_t *tile_first, int64_t n_tiles, const void *a,
                                const void *b, int64_t a_plane, int64_t b_plane, double *c,
                                int32_t *range_flag, uint32_t *tile_ctr);

// Runs a list of gather jobs that already sit in device memory. job_wg_start
// is the exclusive prefix sum of workgroups per job (n_jobs + 1 entries).
tq_status tq_launch_gather_jobs(tq_handle_t h, tq_dtype dt, const CopyJob *jobs,
                                const int32_t *job_wg_start, int64_t n_jobs,
                                int64_t total_wg, const void *src, void *dst);
// Elements of one work item (a wavefront's chunk) of the job-list gather.
constexpr int kGatherPerWG = 512;

__host__ __device__ inline int64_t pack_job_wgs(const CopyJob &j) {
  if (j.n_elem <= 0) return 0;
  if (j.kdim <= 0) return ceil_div(j.n_elem, kGatherPerWG);
  const int64_t rows = j.n_elem / j.kdim;
  if (split_kfast(j)) return ceil_div(rows, kSplitRows) * (j.kpad / 8);
  return ceil_div(rows * (j.kpad / 8), kSplitCells);
}

// ---- grouped GEMM (tq_gemm.hip) ----------------------------------------

// One matched pair as the grouped GEMM sees it: C(MxN) = A(MxK) B(NxK)^T.
// flags: bit0 A is k-fast (element (m,k) at a_off + k + lda*m, else
// a_off + m + lda*k); bit1 A comes from the packed buffer; bit2/bit3 the
// same for B (n instead of m).
struct GemmDesc {
  i64 a_off;
  i64 b_off;
  i64 c_off;
  int32_t m, n, k;
  int32_t lda, ldb;
  int32_t flags;
};

enum : int32_t {
  kAKFast = 1,
  kAPacked = 2,
  kBKFast = 4,
  kBPacked = 8,
};

// A run: pairs p0 .. p0 + np - 1, consecutive in pair order, that share their
// A operand (same a_off, M, K, lda and A flags) and whose output blocks
// follow each other in C. Together they are one GEMM
// C(M x n) = A(M x K) [B_p0; ...; B_p(np-1)]^T with n the sum of their N and
// the output contiguous (ldc = M); the wave-tile kernel finds the B operand of
// each output column in the run's pairs.
struct RunDesc {
  i64 a_off, c_off;
  int32_t m, n, k, lda, flags;
  int32_t p0, np;
};

// Pairs with K <= kStreamMaxK are written by the streaming small-K kernel,
// one work item per segment of kStreamRows rows x kStreamCols columns of the
// output block, instead of MFMA tiles (the column cut keeps the items of a
// wide block small enough to balance across wavefronts).
constexpr int kStreamMaxK = 8;

// Work-list sort key of a block pair (larger first). A pair streamed by the
// small-K kernel costs about its output writes plus its K-long dot products,
// M * N * (K + 4). A pair cut into fixed-size MFMA tiles costs about K + 4
// per tile, and the tile list is dealt round-robin, so ordering pairs by K
// balances the tiles; it also keeps the pairs of one contracted sector (same
// K, hence the same operand blocks) next to each other in their (A block,
// B block) order, so tiles that share operands run together and share L2.
__host__ __device__ inline int64_t pair_order_key(int64_t m, int64_t n, int64_t k) {
  return k <= kStreamMaxK ? m * n * (k + 4) : k + 4;
}
constexpr int kStreamRows = 64;
constexpr int kStreamCols = 64;

// For the split-BF16 tier the operands are the split panels (a_packed /
// b_packed, planes a_plane / b_plane elements apart), or with split_direct
// the FP64 operands in place, each value rounded to its split (hi + mid + lo).
tq_status tq_launch_small_k(tq_handle_t h, tq_dtype dt, tq_precision prec,
                            bool split_direct, const GemmDesc *descs,
                            const int32_t *chunk_pair, const int32_t *chunk_first,
                            int64_t n_chunks, const void *a, const void *a_packed,
                            const void *b, const void *b_packed, void *c, int64_t a_plane,
                            int64_t b_plane, int32_t *range_flag);

// Streamed groups of a merged-output plan: an output block that is the sum
// of up to kGroupMaxMembers matched pairs, each with K <= kStreamMaxK, is
// streamed like a single pair; each lane keeps its row of every member's A
// in registers and every output column is one sum over the members'
// contracted values, written once. A slice is kStreamCols (or fewer)
// columns of a group from column n0: the members are members[mem0 ..
// mem0 + nmem) (their own operand descriptors; the group's M and N), the
// output at c_off with ldc = m.
constexpr int kGroupMaxMembers = 2;
struct GroupSlice {
  i64 c_off;
  int32_t m, n, n0, mem0, nmem, pad;
};
tq_status tq_launch_small_k_groups(tq_handle_t h, tq_dtype dt, tq_precision prec,
                                   const GroupSlice *slices, const GemmDesc *members,
                                   const int32_t *seg_slice, const int32_t *slice_first,
                                   int64_t n_segs, const void *a, const void *a_packed,
                                   const void *b, const void *b_packed, void *c);

// The wave-tile kernel addresses operands with 32-bit byte offsets from a
// per-tile base; it runs only when every leading dimension is below this.
constexpr int64_t kWaveMaxLd = int64_t(1) << 22;

// Tile shape of a configuration; wave_tile = 1 when one wavefront (not a
// four-wave workgroup) computes a tile; min_waves = wavefronts per SIMD the
// kernel's register budget keeps resident (2 for the variants compiled to
// fit two).
struct GemmConfig {
  int bm, bn, bk;
  int wave_tile;
  int min_waves;
};

int tq_gemm_n_configs(tq_precision prec);
GemmConfig tq_gemm_config(tq_precision prec, int cfg);

// Launch the grouped GEMM over `tile_pair` (pair index per tile, in work-list
// order) with tile_start[p] = first tile of pair p; or, with runs != null (a
// wave-tile configuration only), over runs: tile_pair then holds run indices.
// tile_ctr (may be null): kTileCounters device words the 64 x 64 wave-tile
// configurations use to hand out tiles per XCD (zeroed by the launch; a
// concurrent launch must not share them).
constexpr int kTileCounters = 8;
tq_status tq_launch_grouped_gemm(tq_handle_t h, tq_dtype dt, tq_precision prec,
                                 int cfg, const GemmDesc *descs, const RunDesc *runs,
                                 const int32_t *tile_pair,
                                 const int32_t *tile_first, int64_t n_tiles,
                                 const void *a, const void *a_packed,
                                 const void *b, const void *b_packed, void *c,
                                 uint32_t *tile_ctr);

// ---- scans (tq_scan.hip) ----------------------------------------------

// Exclusive prefix sum of n int64 values (in may equal out); writes the total
// to *total_dev (device pointer, may be null). Uses `tmp` of
// tq_scan_tmp_bytes(n) bytes.
size_t tq_scan_tmp_bytes(int64_t n);
tq_status tq_scan_i64(tq_handle_t h, const i64 *in, i64 *out, int64_t n,
                      i64 *total_dev, void *tmp);

// ---- reductions (tq_reduce.hip) ----------------------------------------

// Per-device set-up of the reductions' static scratch (called by tq_create
// with the handle's device current).
tq_status tq_reduce_setup_device(tq_handle_t h);

// Deterministic FP64 sum of n values into *out (device pointer).
tq_status tq_sum_f64_dev(tq_handle_t h, const double *x, int64_t n,
                         double *out, void *tmp);
size_t tq_sum_tmp_bytes(int64_t n);
