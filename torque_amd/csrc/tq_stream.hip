// tq_stream.hip -- the small-K path of the grouped contraction.
//
// A pair with a short contracted extent (K <= kStreamMaxK) does 2K flops per
// output element: its cost is writing C, not multiplying, so it skips the
// MFMA tiles. One wavefront takes a 64-row segment of the pair's column-major
// output: each lane holds its row of A (K values) in registers and walks the
// N columns, and each column is one contiguous 64-element store. B[n, :] is
// the same for the whole wavefront: the lanes load the B rows of 64 columns
// at once (lane j the K values of column j) into the wavefront's LDS slot and
// read them back as broadcasts, and the column walk is compiled once per K,
// so a column costs exactly K LDS reads and K products with no per-k branch
// or scalar-load wait. The planner hands wide
// pairs over as slices of at most kStreamCols columns, each a pair of its own
// here, so that the segments stay small enough to balance.
//
// The split-BF16 tier streams the same pairs from its split panels: a pair
// with K <= 8 has one 8-value chunk per row, so a row is one 16-byte vector
// per plane; each operand value is rebuilt as hi + mid + lo in FP64 (the
// tier's operand rounding, below 2^-24 relative) and the products are FP64
// FMAs -- BF16 MFMA tiles of 32 k would leave three quarters of each step
// empty on a pair whose cost is its output.
#include "tq_common.h"
#include "tq_split.h"

namespace {

constexpr int kThreads = 256;
constexpr int kWaves = kThreads / 64;

// The columns [0, n_cols) of one segment for a pair of K contracted values:
// av holds this lane's row of A, pb points at B row 0 (element (n, k) at
// n * sbn + k * sbk), pc at this lane's row of column 0 of C (column stride
// ldc), bs is the wavefront's LDS slot. Products are summed k = 0, 1, ... as
// fma(a_k, b_k, acc) from a_0 * b_0.
template <int K, class TIn, class TC, class TOut, class In>
__device__ inline void walk_columns(const TC (&av)[kStreamMaxK], const TIn *pb, int64_t sbn,
                                    int64_t sbk, int n_cols, int64_t ldc, TOut *pc, bool live,
                                    In in, TC (*bs)[kStreamMaxK + 1], int lane) {
  for (int n0 = 0; n0 < n_cols; n0 += 64) {
    // lane j loads column n0 + j (columns past the segment repeat its last)
    const TIn *rb = pb + (int64_t)min(n0 + lane, n_cols - 1) * sbn;
    TC bv[K];
#pragma unroll
    for (int k = 0; k < K; ++k) bv[k] = in(rb[k * sbk]);
    // the slot's previous rows were read by this wavefront before this point
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
#pragma unroll
    for (int k = 0; k < K; ++k) bs[lane][k] = bv[k];
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    const int nn = min(64, n_cols - n0);
    for (int j = 0; j < nn; ++j) {
      TC acc = av[0] * bs[j][0];
#pragma unroll
      for (int k = 1; k < K; ++k) acc = fma(av[k], bs[j][k], acc);
      if (live) pc[ldc * (n0 + j)] = (TOut)acc;
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
  }
}

// kRound3: each operand value enters as its split-BF16 representation
// hi + mid + lo (the split-BF16 tier reading plain operands in place).
template <class TIn, class TC, class TOut, bool kRound3 = false>
__global__ __launch_bounds__(kThreads) void small_k_kernel(
    const GemmDesc *__restrict__ descs, const int32_t *__restrict__ seg_pair,
    const int32_t *__restrict__ seg_first, int64_t n_segs,
    const TIn *__restrict__ a, const TIn *__restrict__ a_packed,
    const TIn *__restrict__ b, const TIn *__restrict__ b_packed,
    TOut *__restrict__ c, int32_t *__restrict__ range_flag) {
  static_assert(kStreamMaxK == 8, "one column walk per K = 1 .. 8");
  // B rows of 64 columns per wavefront (K + 1 values per row: rows start on
  // different banks)
  __shared__ TC brows[kWaves][64][kStreamMaxK + 1];
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  bool bad = false;  // kRound3: a value outside the split's domain
  // the segment index is the same on every lane of the wavefront: say so, so
  // that the descriptor reads become scalar loads
  const int wave_id = __builtin_amdgcn_readfirstlane((int)(blockIdx.x * kWaves + wave));
  const int n_waves = (int)(gridDim.x * kWaves);
  for (int sg = wave_id; sg < n_segs; sg += n_waves) {
    const int p = __builtin_amdgcn_readfirstlane(seg_pair[sg]);
    const GemmDesc d = descs[p];
    const int m = (int)(sg - seg_first[p]) * 64 + lane;
    const bool live = m < d.m;
    const TIn *pa = ((d.flags & kAPacked) ? a_packed : a) + d.a_off;
    const TIn *pb = ((d.flags & kBPacked) ? b_packed : b) + d.b_off;
    const int64_t sam = (d.flags & kAKFast) ? d.lda : 1;
    const int64_t sak = (d.flags & kAKFast) ? 1 : d.lda;
    const int64_t sbn = (d.flags & kBKFast) ? d.ldb : 1;
    const int64_t sbk = (d.flags & kBKFast) ? 1 : d.ldb;
    auto in = [&](TIn x) -> TC {
      if constexpr (kRound3) {
        bad = bad || tqs::split_out_of_range((double)x);
        return (TC)tqs::round3((double)x);
      } else {
        return (TC)x;
      }
    };
    TC av[kStreamMaxK];
#pragma unroll
    for (int k = 0; k < kStreamMaxK; ++k)
      av[k] = (live && k < d.k) ? in(pa[m * sam + k * sak]) : (TC)0;
    TOut *pc = c + d.c_off + m;
    // 1 <= K <= kStreamMaxK for every streamed pair
    switch (d.k) {
#define TQ_WALK(K)                                                                       \
  case K:                                                                                \
    walk_columns<K, TIn, TC, TOut>(av, pb, sbn, sbk, d.n, d.m, pc, live, in, brows[wave], \
                                   lane);                                                \
    break;
      TQ_WALK(1)
      TQ_WALK(2)
      TQ_WALK(3)
      TQ_WALK(4)
      TQ_WALK(5)
      TQ_WALK(6)
      TQ_WALK(7)
      default:
      TQ_WALK(8)
#undef TQ_WALK
    }
  }
  if (kRound3 && bad && range_flag) atomicOr(range_flag, 1);
}

// Streamed groups of a merged-output plan: a slice's output column sums the
// products of all its members, member by member in group order and each
// member's contracted values in order (the sum over the concatenated range),
// and is written once. Each lane holds its row of every member's A; the
// slice's B values (at most kStreamCols columns) are loaded once per segment,
// lane j taking column j of every member, into the wavefront's LDS slot and
// read back by every lane as broadcasts.
constexpr int kGroupRow = kGroupMaxMembers * kStreamMaxK + 1;  // odd: no bank conflicts

// Column walk of a slice for members of K0 and K1 contracted values (K1 = 0:
// one member): member 0's products, then member 1's, each summed k = 0, 1,
// ... from a_0 * b_0 of member 0 -- the order of the concatenated range.
template <int K0, int K1, class TC, class TOut>
__device__ inline void group_walk(const TC (*av)[kStreamMaxK], const TC (*brow)[kGroupRow],
                                  int n, int64_t ldc, TOut *pc, bool live) {
  for (int j = 0; j < n; ++j) {
    TC acc = av[0][0] * brow[j][0];
#pragma unroll
    for (int k = 1; k < K0; ++k) acc = fma(av[0][k], brow[j][k], acc);
#pragma unroll
    for (int k = 0; k < K1; ++k) acc = fma(av[1][k], brow[j][kStreamMaxK + k], acc);
    if (live) pc[ldc * j] = (TOut)acc;
  }
}

// group_walk for member 0 of K0 values and member 1 (if any) of k1 values.
template <int K0, int NM, class TC, class TOut>
__device__ inline void group_walk_k1(int k1, const TC (*av)[kStreamMaxK],
                                     const TC (*brow)[kGroupRow], int n, int64_t ldc,
                                     TOut *pc, bool live) {
  if constexpr (NM == 1) {
    group_walk<K0, 0, TC, TOut>(av, brow, n, ldc, pc, live);
  } else {
    switch (k1) {
#define TQ_GWALK(K1)                                                 \
  case K1:                                                           \
    group_walk<K0, K1, TC, TOut>(av, brow, n, ldc, pc, live);        \
    break;
      TQ_GWALK(1)
      TQ_GWALK(2)
      TQ_GWALK(3)
      TQ_GWALK(4)
      TQ_GWALK(5)
      TQ_GWALK(6)
      TQ_GWALK(7)
      default:
      TQ_GWALK(8)
#undef TQ_GWALK
    }
  }
}

template <class TIn, class TC, class TOut, int NM>
__device__ inline void group_slice(const GroupSlice &g, const GemmDesc *__restrict__ members,
                                   int row0, const TIn *__restrict__ a,
                                   const TIn *__restrict__ a_packed, const TIn *__restrict__ b,
                                   const TIn *__restrict__ b_packed, TOut *__restrict__ c,
                                   TC (*brow)[kGroupRow]) {
  const int lane = threadIdx.x & 63;
  const int m = row0 + lane;
  const bool live = m < g.m;
  TC av[NM][kStreamMaxK];
  int kq[NM];
  TC bl[NM][kStreamMaxK];
#pragma unroll
  for (int q = 0; q < NM; ++q) {
    const GemmDesc d = members[g.mem0 + q];
    const TIn *pa = ((d.flags & kAPacked) ? a_packed : a) + d.a_off;
    const int64_t sam = (d.flags & kAKFast) ? d.lda : 1;
    const int64_t sak = (d.flags & kAKFast) ? 1 : d.lda;
    const int64_t sbn = (d.flags & kBKFast) ? d.ldb : 1;
    const int64_t sbk = (d.flags & kBKFast) ? 1 : d.ldb;
    const TIn *pb = ((d.flags & kBPacked) ? b_packed : b) + d.b_off +
                    (int64_t)(g.n0 + min(lane, g.n - 1)) * sbn;
    kq[q] = d.k;
#pragma unroll
    for (int k = 0; k < kStreamMaxK; ++k) {
      av[q][k] = (live && k < d.k) ? (TC)pa[m * sam + k * sak] : (TC)0;
      bl[q][k] = k < d.k ? (TC)pb[k * sbk] : (TC)0;
    }
  }
  // the slot's previous rows were read by this wavefront before this point
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
#pragma unroll
  for (int q = 0; q < NM; ++q)
#pragma unroll
    for (int k = 0; k < kStreamMaxK; ++k) brow[lane][q * kStreamMaxK + k] = bl[q][k];
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  TOut *pc = c + g.c_off + m;
  // every member holds 1 <= K <= kStreamMaxK contracted values; one walk per
  // (K0, K1)
  const int k1 = NM > 1 ? kq[NM - 1] : 0;
  switch (kq[0]) {
#define TQ_GWALK0(K0)                                                                   \
  case K0:                                                                              \
    group_walk_k1<K0, NM, TC, TOut>(k1, av, brow, g.n, (int64_t)g.m, pc, live);         \
    break;
    TQ_GWALK0(1)
    TQ_GWALK0(2)
    TQ_GWALK0(3)
    TQ_GWALK0(4)
    TQ_GWALK0(5)
    TQ_GWALK0(6)
    TQ_GWALK0(7)
    default:
    TQ_GWALK0(8)
#undef TQ_GWALK0
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

template <class TIn, class TC, class TOut>
__global__ __launch_bounds__(kThreads) void small_k_group_kernel(
    const GroupSlice *__restrict__ slices, const GemmDesc *__restrict__ members,
    const int32_t *__restrict__ seg_slice, const int32_t *__restrict__ slice_first,
    int64_t n_segs, const TIn *__restrict__ a, const TIn *__restrict__ a_packed,
    const TIn *__restrict__ b, const TIn *__restrict__ b_packed, TOut *__restrict__ c) {
  static_assert(kGroupMaxMembers == 2, "one instance per member count");
  static_assert(kStreamCols <= 64, "a slice's columns fit the wavefront");
  __shared__ TC brow[kWaves][64][kGroupRow];
  const int wave = threadIdx.x >> 6;
  const int wave_id = __builtin_amdgcn_readfirstlane((int)(blockIdx.x * kWaves + wave));
  const int n_waves = (int)(gridDim.x * kWaves);
  for (int sg = wave_id; sg < n_segs; sg += n_waves) {
    const int si = __builtin_amdgcn_readfirstlane(seg_slice[sg]);
    const GroupSlice g = slices[si];
    const int row0 = (int)(sg - slice_first[si]) * 64;
    if (g.nmem == 1)
      group_slice<TIn, TC, TOut, 1>(g, members, row0, a, a_packed, b, b_packed, c, brow[wave]);
    else
      group_slice<TIn, TC, TOut, 2>(g, members, row0, a, a_packed, b, b_packed, c, brow[wave]);
  }
}

template <class TIn, class TC, class TOut>
tq_status launch_groups(tq_handle_t h, const GroupSlice *slices, const GemmDesc *members,
                        const int32_t *seg_slice, const int32_t *slice_first, int64_t n_segs,
                        const void *a, const void *ap, const void *b, const void *bp,
                        void *c) {
  static int per_cu = 0;
  if (per_cu == 0) {
    int occ = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &occ, small_k_group_kernel<TIn, TC, TOut>, kThreads, 0) != hipSuccess ||
        occ < 1)
      occ = 1;
    per_cu = occ;
  }
  const int64_t grid = std::max<int64_t>(
      1, std::min<int64_t>(ceil_div(n_segs, kWaves), (int64_t)h->n_cu * per_cu));
  hipLaunchKernelGGL((small_k_group_kernel<TIn, TC, TOut>), dim3((unsigned)grid),
                     dim3(kThreads), 0, h->stream, slices, members, seg_slice, slice_first,
                     n_segs, reinterpret_cast<const TIn *>(a),
                     reinterpret_cast<const TIn *>(ap), reinterpret_cast<const TIn *>(b),
                     reinterpret_cast<const TIn *>(bp), reinterpret_cast<TOut *>(c));
  TQ_HIP(h, hipGetLastError());
  return TQ_SUCCESS;
}

__device__ inline double bf16_bits_value(uint32_t h) {
  return (double)__uint_as_float(h << 16);
}

// The 8 values of one split-panel row chunk: hi + mid + lo per value.
__device__ inline void split_row(const uint16_t *p, int64_t plane, double (&v)[8]) {
  const uint4 h = *reinterpret_cast<const uint4 *>(p);
  const uint4 m = *reinterpret_cast<const uint4 *>(p + plane);
  const uint4 l = *reinterpret_cast<const uint4 *>(p + 2 * plane);
  const uint32_t hw[4] = {h.x, h.y, h.z, h.w}, mw[4] = {m.x, m.y, m.z, m.w},
                 lw[4] = {l.x, l.y, l.z, l.w};
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    v[2 * q] = bf16_bits_value(hw[q] & 0xffffu) + bf16_bits_value(mw[q] & 0xffffu) +
               bf16_bits_value(lw[q] & 0xffffu);
    v[2 * q + 1] = bf16_bits_value(hw[q] >> 16) + bf16_bits_value(mw[q] >> 16) +
                   bf16_bits_value(lw[q] >> 16);
  }
}

// nn columns of a segment from the rebuilt B rows in LDS (row j: column j),
// K products per column summed k = 0, 1, ... as in walk_columns.
template <int K>
__device__ inline void split_walk(const double (&av)[8], const double (*bs)[9], int nn,
                                  int64_t ldc, double *pc, bool live) {
  for (int j = 0; j < nn; ++j) {
    double acc = av[0] * bs[j][0];
#pragma unroll
    for (int k = 1; k < K; ++k) acc = fma(av[k], bs[j][k], acc);
    if (live) pc[ldc * j] = acc;
  }
}

// B rows are rebuilt 64 at a time, one per lane, into the wavefront's LDS
// slot, then read back by every lane as broadcasts: the rebuild costs one
// pass per 64 output columns instead of one per column and lane. The column
// walk is compiled once per K.
__global__ __launch_bounds__(kThreads) void small_k_split_kernel(
    const GemmDesc *__restrict__ descs, const int32_t *__restrict__ seg_pair,
    const int32_t *__restrict__ seg_first, int64_t n_segs, const uint16_t *__restrict__ a,
    int64_t a_plane, const uint16_t *__restrict__ b, int64_t b_plane,
    double *__restrict__ c) {
  __shared__ double brow[kWaves][64][9];  // 9: rows on distinct bank pairs
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int wave_id = __builtin_amdgcn_readfirstlane(
      (int)(blockIdx.x * kWaves + wave));
  const int n_waves = (int)(gridDim.x * kWaves);
  for (int sg = wave_id; sg < n_segs; sg += n_waves) {
    const int p = __builtin_amdgcn_readfirstlane(seg_pair[sg]);
    const GemmDesc d = descs[p];
    const int m = (int)(sg - seg_first[p]) * 64 + lane;
    const bool live = m < d.m;
    double av[8];
    // K <= 8: the panel holds one chunk, row r at r * 8
    split_row(a + d.a_off + (int64_t)min(m, d.m - 1) * 8, a_plane, av);
    double *pc = c + d.c_off + m;
    for (int n0 = 0; n0 < d.n; n0 += 64) {
      double bv[8];
      split_row(b + d.b_off + (int64_t)min(n0 + lane, d.n - 1) * 8, b_plane, bv);
      // the slot's previous rows were read by this wavefront before this
      // point; publish the new ones to the wavefront
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
#pragma unroll
      for (int k = 0; k < 8; ++k) brow[wave][lane][k] = bv[k];
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      const int nn = min(64, d.n - n0);
      switch (d.k) {
#define TQ_SPLIT_WALK(K)                                                         \
  case K:                                                                        \
    split_walk<K>(av, brow[wave], nn, (int64_t)d.m, pc + (int64_t)d.m * n0, live); \
    break;
        TQ_SPLIT_WALK(1)
        TQ_SPLIT_WALK(2)
        TQ_SPLIT_WALK(3)
        TQ_SPLIT_WALK(4)
        TQ_SPLIT_WALK(5)
        TQ_SPLIT_WALK(6)
        TQ_SPLIT_WALK(7)
        default:
        TQ_SPLIT_WALK(8)
#undef TQ_SPLIT_WALK
      }
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
    }
  }
}

template <class TIn, class TC, class TOut, bool kRound3 = false>
tq_status launch(tq_handle_t h, const GemmDesc *descs, const int32_t *chunk_pair,
                 const int32_t *chunk_first, int64_t n_chunks, const void *a,
                 const void *ap, const void *b, const void *bp, void *c,
                 int32_t *range_flag = nullptr) {
  static int per_cu = 0;
  if (per_cu == 0) {
    int occ = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &occ, small_k_kernel<TIn, TC, TOut, kRound3>, kThreads, 0) != hipSuccess ||
        occ < 1)
      occ = 1;
    per_cu = occ;
  }
  const int64_t grid = std::max<int64_t>(
      1, std::min<int64_t>(ceil_div(n_chunks, kWaves), (int64_t)h->n_cu * per_cu));
  hipLaunchKernelGGL((small_k_kernel<TIn, TC, TOut, kRound3>), dim3((unsigned)grid),
                     dim3(kThreads),
                     0, h->stream, descs, chunk_pair, chunk_first, n_chunks,
                     reinterpret_cast<const TIn *>(a), reinterpret_cast<const TIn *>(ap),
                     reinterpret_cast<const TIn *>(b), reinterpret_cast<const TIn *>(bp),
                     reinterpret_cast<TOut *>(c), range_flag);
  TQ_HIP(h, hipGetLastError());
  return TQ_SUCCESS;
}

}  // namespace

tq_status tq_launch_small_k(tq_handle_t h, tq_dtype dt, tq_precision prec,
                            bool split_direct, const GemmDesc *descs,
                            const int32_t *chunk_pair, const int32_t *chunk_first,
                            int64_t n_chunks, const void *a, const void *a_packed,
                            const void *b, const void *b_packed, void *c, int64_t a_plane,
                            int64_t b_plane, int32_t *range_flag) {
  if (n_chunks <= 0) return TQ_SUCCESS;
  if (prec == TQ_PREC_BF16X3 && split_direct)
    return launch<double, double, double, true>(h, descs, chunk_pair, chunk_first, n_chunks,
                                                a, a_packed, b, b_packed, c, range_flag);
  if (prec == TQ_PREC_BF16X3) {
    static int per_cu = 0;
    if (per_cu == 0) {
      int occ = 0;
      if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, small_k_split_kernel, kThreads,
                                                       0) != hipSuccess ||
          occ < 1)
        occ = 1;
      per_cu = occ;
    }
    const int64_t grid = std::max<int64_t>(
        1, std::min<int64_t>(ceil_div(n_chunks, kWaves), (int64_t)h->n_cu * per_cu));
    hipLaunchKernelGGL(small_k_split_kernel, dim3((unsigned)grid), dim3(kThreads), 0,
                       h->stream, descs, chunk_pair, chunk_first, n_chunks,
                       reinterpret_cast<const uint16_t *>(a_packed), a_plane,
                       reinterpret_cast<const uint16_t *>(b_packed), b_plane,
                       reinterpret_cast<double *>(c));
    TQ_HIP(h, hipGetLastError());
    return TQ_SUCCESS;
  }
  if (dt == TQ_F32)
    return launch<float, float, float>(h, descs, chunk_pair, chunk_first, n_chunks, a,
                                       a_packed, b, b_packed, c);
  if (dt == TQ_F16)
    return launch<_Float16, float, _Float16>(h, descs, chunk_pair, chunk_first, n_chunks, a,
                                             a_packed, b, b_packed, c);
  if (prec == TQ_PREC_F32)
    return launch<double, float, double>(h, descs, chunk_pair, chunk_first, n_chunks, a,
                                         a_packed, b, b_packed, c);
  return launch<double, double, double>(h, descs, chunk_pair, chunk_first, n_chunks, a,
                                        a_packed, b, b_packed, c);
}

tq_status tq_launch_small_k_groups(tq_handle_t h, tq_dtype dt, tq_precision prec,
                                   const GroupSlice *slices, const GemmDesc *members,
                                   const int32_t *seg_slice, const int32_t *slice_first,
                                   int64_t n_segs, const void *a, const void *a_packed,
                                   const void *b, const void *b_packed, void *c) {
  if (n_segs <= 0) return TQ_SUCCESS;
  if (prec == TQ_PREC_BF16X3)
    return tq_fail(h, TQ_ERROR_UNSUPPORTED, "contract: split-BF16 groups stream from panels");
  if (dt == TQ_F32)
    return launch_groups<float, float, float>(h, slices, members, seg_slice, slice_first,
                                              n_segs, a, a_packed, b, b_packed, c);
  if (dt == TQ_F16)
    return launch_groups<_Float16, float, _Float16>(h, slices, members, seg_slice,
                                                    slice_first, n_segs, a, a_packed, b,
                                                    b_packed, c);
  if (prec == TQ_PREC_F32)
    return launch_groups<double, float, double>(h, slices, members, seg_slice, slice_first,
                                                n_segs, a, a_packed, b, b_packed, c);
  return launch_groups<double, double, double>(h, slices, members, seg_slice, slice_first,
                                               n_segs, a, a_packed, b, b_packed, c);
}
