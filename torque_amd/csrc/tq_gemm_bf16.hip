// tq_gemm_bf16.hip -- the split-BF16 tier of the grouped GEMM.
//
// Each FP64 operand value x is taken as three BF16 pieces hi + mid + lo
// (tq_split.h). The product keeps the six piece products of order >= 2^-16
// relative,
//   hi*lo + mid*mid + hi*mid + lo*hi + mid*hi + hi*hi   (A piece * B piece,
// in this order per k step: grouped by B plane, smaller products first),
// on v_mfma_f32_16x16x32_bf16 with FP32 accumulation; the dropped terms
// (mid*lo, lo*mid, lo*lo) are below 2^-23 of |a||b| and the pieces represent
// each operand to 2^-26 relative. Results are written as FP64.
//
// Operands come in one of two forms, fixed per plan:
//  * packed: three BF16 planes written by the split pack (tq_permute.hip). A
//    panel of R rows is k-chunked: value (row, k) of a plane sits at
//    (k / 8) * R * 8 + row * 8 + k % 8, so the 8 values of one row in one
//    chunk are one 16-byte vector and consecutive rows are consecutive
//    vectors.
//  * in place: the FP64 operand as the contraction's data holds it (a plain
//    matrix, m or k fastest). The loaders read FP64 and the split happens in
//    registers as the step is written to LDS, so no pack pass runs at all.
// Both forms stage the same BF16 vectors in LDS and multiply them in the same
// order: a contraction gives the same bits either way.
//
// Schedule: persistent workgroups walk the same tile list as the FP64 kernel
// (tq_gemm.hip). A k step covers 32 values (four chunks): the workgroup
// stages the step's A and B vectors of all three planes in LDS
// ([plane][chunk][row] vectors, so the 16 lanes reading one MFMA operand read
// 256 contiguous bytes), and the values of the next step (or of the next
// tile's first step) are loaded into registers while the current step's
// MFMAs run. The six piece products are six passes over the wavefront's
// register tile, so consecutive MFMAs never depend on each other; the three
// A planes stay in registers and the B planes are read one at a time.
#include <type_traits>

#include "tq_gemm_common.h"
#include "tq_split.h"

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f4 __attribute__((ext_vector_type(4)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
typedef double f64x2 __attribute__((ext_vector_type(2)));

using tqg::kThreads;
constexpr int kBK = 32;     // k per step
constexpr int kChunks = 4;  // 8-value chunks per step

// One operand as a thread's loader reads it: p is the operand's first
// element (plane 0 of a split panel, or the FP64 value in place); the tile
// starts at row row0 of its `rows` rows (rows of A, or output columns for
// B); ld is the panel's padded K (packed) or the leading dimension (in
// place); kfast says whether k is the fast index in place.
struct OpRef {
  const void *p;
  int64_t ld;
  int rows, row0;
  int kfast;
};

template <int ROWS, int THREADS, bool DIRECT>
struct StepLoader;

// Packed operand: one k step is 3 planes x 4 chunks x ROWS rows of 16-byte
// vectors; vector v of the step is (row v % ROWS, chunk (v / ROWS) % 4,
// plane v / (4 ROWS)), so consecutive threads read consecutive rows.
template <int ROWS, int THREADS>
struct StepLoader<ROWS, THREADS, false> {
  static constexpr int kVec = 3 * kChunks * ROWS;
  static constexpr int kN = (kVec + THREADS - 1) / THREADS;
  static_assert(kN <= 32, "one mask bit per vector");
  u32x4 v[kN];
  uint32_t ok;  // bit q: v[q] is in range (otherwise it is stored as zeros)
  int32_t *flag = nullptr;  // unused: the split pack checks the values

  // Rows past R are read from row R - 1 (their products only reach
  // unstored output); chunks past the padded K are zeros. Branch-free and
  // without waiting on the loads: out-of-range vectors load a clamped
  // in-range vector and become zeros when stored. (Native vector types: the
  // loaded values stay in registers.)
  __device__ inline void load(const OpRef &o, int64_t plane, int /*K*/, int step) {
    ok = 0;
    const int n_chunks = (int)(o.ld >> 3);
    if (n_chunks <= 0) return;  // no k values: the panel holds nothing to read
    const uint16_t *base = static_cast<const uint16_t *>(o.p);
    const int chunk0 = step * kChunks;
    const int last = n_chunks - 1;
#pragma unroll
    for (int q = 0; q < kN; ++q) {
      const int e = threadIdx.x + q * THREADS;
      const int r = e % ROWS;
      const int c = (e / ROWS) % kChunks;
      const int p = min(e / (ROWS * kChunks), 2);
      const int row = min(o.row0 + r, o.rows - 1);
      const int chunk = chunk0 + c;
      ok |= (uint32_t)(chunk < n_chunks && (kVec % THREADS == 0 || e < kVec)) << q;
      v[q] = *reinterpret_cast<const u32x4 *>(
          base + p * plane + ((int64_t)min(chunk, last) * o.rows + row) * 8);
    }
  }

  __device__ inline void store(uint4 *lds) const {
#pragma unroll
    for (int q = 0; q < kN; ++q) {
      const int e = threadIdx.x + q * THREADS;
      const u32x4 x = ((ok >> q) & 1) ? v[q] : u32x4{0, 0, 0, 0};
      if (kVec % THREADS == 0 || e < kVec) reinterpret_cast<u32x4 *>(lds)[e] = x;
    }
  }
};

// FP64 operand in place: one k step is ROWS x 4 cells of 8 values (a row's
// chunk); cell e is (row e % ROWS, chunk e / ROWS), so consecutive threads
// take consecutive rows (coalesced when m is the fast index; k-fast rows are
// read as 16-byte pairs where aligned). Rows past R read row R - 1 and k past
// K read k = K - 1 (in bounds); the split on the way to LDS zeroes the
// values at k >= K.
template <int ROWS, int THREADS>
struct StepLoader<ROWS, THREADS, true> {
  static constexpr int kCells = kChunks * ROWS;
  static constexpr int kN = (kCells + THREADS - 1) / THREADS;
  double v[kN][8];
  int kleft;  // K minus the step's first k
  int32_t *flag = nullptr;  // or-ed with 1 when a value is outside the split's domain

  __device__ inline void load(const OpRef &o, int64_t /*plane*/, int K, int step) {
    const double *base = static_cast<const double *>(o.p);
    const int k0 = step * kBK;
    kleft = K - k0;
    if (kleft <= 0) return;  // stored as zeros
#pragma unroll
    for (int q = 0; q < kN; ++q) {
      const int e = threadIdx.x + q * THREADS;
      const int r = e % ROWS;
      const int c = min(e / ROWS, kChunks - 1);  // lanes past the cells load, never store
      const int64_t row = min(max(o.row0 + r, 0), o.rows - 1);
      const int kb = k0 + 8 * c;
      const bool full = kb + 8 <= K;  // the whole chunk lies below K
      if (o.kfast) {
        const double *p = base + row * o.ld + kb;
        if (full && (reinterpret_cast<uintptr_t>(p) & 15) == 0) {
#pragma unroll
          for (int j = 0; j < 4; ++j) {
            const f64x2 t = *reinterpret_cast<const f64x2 *>(p + 2 * j);
            v[q][2 * j] = t.x;
            v[q][2 * j + 1] = t.y;
          }
        } else if (full) {
#pragma unroll
          for (int j = 0; j < 8; ++j) v[q][j] = p[j];
        } else {
#pragma unroll
          for (int j = 0; j < 8; ++j) v[q][j] = base[row * o.ld + min(kb + j, K - 1)];
        }
      } else {
        const double *p = base + row + o.ld * kb;
        if (full) {
#pragma unroll
          for (int j = 0; j < 8; ++j) v[q][j] = p[j * o.ld];
        } else {
#pragma unroll
          for (int j = 0; j < 8; ++j) v[q][j] = base[row + o.ld * min(kb + j, K - 1)];
        }
      }
    }
  }

  __device__ inline void store(uint4 *lds) const {
    bool bad = false;
#pragma unroll
    for (int q = 0; q < kN; ++q) {
      const int e = threadIdx.x + q * THREADS;
      if (kCells % THREADS != 0 && e >= kCells) continue;
      const int r = e % ROWS, c = e / ROWS;
      const int cnt = kleft - 8 * c;  // values of the cell with k < K
      u32x4 w[3];
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const double x0 = 2 * j < cnt ? v[q][2 * j] : 0.0;
        const double x1 = 2 * j + 1 < cnt ? v[q][2 * j + 1] : 0.0;
        bad = bad || tqs::split_out_of_range(x0) || tqs::split_out_of_range(x1);
        uint32_t h, m, l;
        tqs::split3x2(x0, x1, h, m, l);
        w[0][j] = h;
        w[1][j] = m;
        w[2][j] = l;
      }
#pragma unroll
      for (int p = 0; p < 3; ++p)
        reinterpret_cast<u32x4 *>(lds)[(p * kChunks + c) * ROWS + r] = w[p];
    }
    if (bad && flag) atomicOr(flag, 1);
  }
};

// A tile: the GEMM it belongs to (a pair, or a run of pairs), its corner,
// and this thread's view of the A and B operands for the loaders. For a run
// the B rows of the thread (its output column) come from the run's pair
// holding that column.
struct TileRef {
  GemmDesc d;
  int m0, n0;
  OpRef a, b;
};

template <int BM, int BN, int THREADS, bool DIRECT>
__device__ inline TileRef fetch(int64_t t, const GemmDesc *__restrict__ descs,
                                const RunDesc *__restrict__ runs,
                                const int32_t *__restrict__ tile_pair,
                                const int32_t *__restrict__ tile_first,
                                const void *__restrict__ a, const void *__restrict__ b) {
  using Elem = typename std::conditional<DIRECT, double, uint16_t>::type;
  const Elem *pa = static_cast<const Elem *>(a);
  const Elem *pb = static_cast<const Elem *>(b);
  TileRef r;
  const int u = tile_pair[t];
  RunDesc run{};
  if (runs) {
    run = runs[u];
    r.d = GemmDesc{run.a_off, 0, run.c_off, run.m, run.n, run.k, run.lda, run.lda, run.flags};
  } else {
    r.d = descs[u];
  }
  const tqg::TilePos tp =
      tqg::tile_position((int)(t - tile_first[u]), (r.d.m + BM - 1) / BM, (r.d.n + BN - 1) / BN,
                         tqg::tile_group<BM, BN>(gridDim.x));
  r.m0 = tp.mt * BM;
  r.n0 = tp.nt * BN;
  r.a = OpRef{pa + r.d.a_off, r.d.lda, r.d.m, r.m0, (r.d.flags & kAKFast) != 0};
  if (!runs) {
    r.b = OpRef{pb + r.d.b_off, r.d.ldb, r.d.n, r.n0, (r.d.flags & kBKFast) != 0};
    return r;
  }
  // the pair of the run holding this thread's column (columns past the run
  // read its last one); the pairs of a run share K
  const int rt = BN >= THREADS ? (int)threadIdx.x : (int)threadIdx.x % BN;
  const int col = min(r.n0 + rt, r.d.n - 1);
  r.b = OpRef{pb, 8, 1, 0, 1};
  int start = 0;
  for (int q = 0; q < run.np; ++q) {
    const GemmDesc &pd = descs[run.p0 + q];
    if (col >= start && col < start + pd.n)
      r.b = OpRef{pb + pd.b_off, pd.ldb, pd.n, col - start - rt, (pd.flags & kBKFast) != 0};
    start += pd.n;
  }
  return r;
}

// k steps of a tile: a tile with no k values still runs one step over zeros.
__device__ inline int tile_steps(const GemmDesc &d) { return max(1, (d.k + kBK - 1) / kBK); }

// The six piece products of one k step for a wavefront's TM x TN blocks:
// all three A planes stay in registers; B planes are read one at a time
// (fewer live registers), each followed by its products. Every configuration
// adds the products to an accumulator in this one order, so all of them give
// the same bits. a_row / b_row: the LDS row of the wavefront's first block
// for this lane.
template <int BM, int BN, int TM, int TN>
__device__ inline void step_products(const uint4 *As, const uint4 *Bs, int a_row, int b_row,
                                     int lchunk, f4 (&acc)[TM][TN]) {
  bf16x8 af[3][TM];
#pragma unroll
  for (int p = 0; p < 3; ++p)
#pragma unroll
    for (int i = 0; i < TM; ++i)
      af[p][i] =
          *reinterpret_cast<const bf16x8 *>(&As[(p * kChunks + lchunk) * BM + a_row + 16 * i]);
#pragma unroll
  for (int pbq = 2; pbq >= 0; --pbq) {
    bf16x8 bfr[TN];
#pragma unroll
    for (int j = 0; j < TN; ++j)
      bfr[j] =
          *reinterpret_cast<const bf16x8 *>(&Bs[(pbq * kChunks + lchunk) * BN + b_row + 16 * j]);
    // A planes paired with this B plane, smaller products first:
    // b_lo: a_hi; b_mid: a_mid, a_hi; b_hi: a_lo, a_mid, a_hi
#pragma unroll
    for (int pa = 2 - pbq; pa >= 0; --pa)
#pragma unroll
      for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j)
          acc[i][j] =
              __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[pa][i], bfr[j], acc[i][j], 0, 0, 0);
  }
}

template <int BM, int BN>
struct Smem {
  static constexpr size_t kAB = sizeof(uint4) * 3 * kChunks * (size_t)(BM + BN);
  static constexpr size_t kStage = sizeof(double) * 16 * (size_t)(BM + 1);
  static constexpr size_t kBytes = kAB > kStage ? kAB : kStage;
};

// MINW: wavefronts per SIMD the register allocation must allow (2 makes
// the 128 x 128 tile fit two workgroups per CU at the price of a few spills).
template <int BM, int BN, int WAVES_M, int MINW, bool DIRECT>
__global__ __launch_bounds__(kThreads, MINW) void bf16x3_gemm_kernel(
    const GemmDesc *__restrict__ descs, const RunDesc *__restrict__ runs,
    const int32_t *__restrict__ tile_pair, const int32_t *__restrict__ tile_first,
    int64_t n_tiles, const void *__restrict__ a, const void *__restrict__ b,
    int64_t a_plane, int64_t b_plane, double *__restrict__ c,
    int32_t *__restrict__ range_flag, uint32_t *__restrict__ tile_ctr) {
  constexpr int WAVES_N = 4 / WAVES_M;
  constexpr int kT = kThreads;
  constexpr int WM = BM / WAVES_M, WN = BN / WAVES_N;
  constexpr int TM = WM / 16, TN = WN / 16;
  static_assert(TM >= 1 && TN >= 1, "tile shape");
  __shared__ __attribute__((aligned(16))) char smem[Smem<BM, BN>::kBytes];
  uint4 *As = reinterpret_cast<uint4 *>(smem);
  uint4 *Bs = As + 3 * kChunks * BM;
  double *stage = reinterpret_cast<double *>(smem);
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int wm0 = (wave % WAVES_M) * WM, wn0 = (wave / WAVES_M) * WN;
  const int lrow = lane & 15, lchunk = lane >> 4;

  // tiles from the XCD's counter (tqg::stripe_tile) when tile_ctr is given:
  // thread 0 draws the workgroup's next tile at the start of each tile into
  // one of two LDS slots, read by every thread after step 0's first barrier
  // (the slot is next written two tiles later, long after every read)
  __shared__ int64_t next_tile[2];
  const int part = (int)(blockIdx.x % kTileCounters);
  int64_t t;
  if (tile_ctr) {
    if (threadIdx.x == 0) next_tile[0] = tqg::stripe_tile(atomicAdd(tile_ctr + part, 1u), part);
    __syncthreads();
    t = next_tile[0];
  } else {
    t = tqg::xcd_virtual_id(blockIdx.x, gridDim.x);
  }
  if (t >= n_tiles) return;
  TileRef cur = fetch<BM, BN, kT, DIRECT>(t, descs, runs, tile_pair, tile_first, a, b);
  StepLoader<BM, kT, DIRECT> la;
  StepLoader<BN, kT, DIRECT> lb;
  la.flag = lb.flag = range_flag;
  la.load(cur.a, a_plane, cur.d.k, 0);
  lb.load(cur.b, b_plane, cur.d.k, 0);
  for (int it = 1;; ++it) {
    if (tile_ctr && threadIdx.x == 0)
      next_tile[it & 1] = tqg::stripe_tile(atomicAdd(tile_ctr + part, 1u), part);
    const GemmDesc &d = cur.d;
    const int steps = tile_steps(d);
    // every MFMA block runs (blocks past M or N compute on clamped rows and
    // are not stored), so the MFMA stream has no branches
    f4 acc[TM][TN];
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
      for (int j = 0; j < TN; ++j) acc[i][j] = f4{0.f, 0.f, 0.f, 0.f};

    int64_t t_next = t + gridDim.x;
    TileRef nxt = cur;
    for (int s = 0; s < steps; ++s) {
      __syncthreads();  // the previous step (or epilogue staging) is read
      if (tile_ctr && s == 0) t_next = next_tile[it & 1];
      la.store(As);
      lb.store(Bs);
      __syncthreads();
      if (s + 1 < steps) {
        la.load(cur.a, a_plane, d.k, s + 1);
        lb.load(cur.b, b_plane, d.k, s + 1);
      } else if (t_next < n_tiles) {
        // last step: the next tile's first step loads under these MFMAs and
        // the epilogue
        nxt = fetch<BM, BN, kT, DIRECT>(t_next, descs, runs, tile_pair, tile_first, a, b);
        la.load(nxt.a, a_plane, nxt.d.k, 0);
        lb.load(nxt.b, b_plane, nxt.d.k, 0);
      }
      step_products<BM, BN, TM, TN>(As, Bs, wm0 + lrow, wn0 + lrow, lchunk, acc);
    }
    tqg::store_tile<double, false, BM, BN, WAVES_M, TM, TN>(acc, c, d, cur.m0, cur.n0, stage);
    if (t_next >= n_tiles) break;
    t = t_next;
    cur = nxt;
  }
}

// Double-buffered variant for large tiles at one workgroup per CU: LDS holds
// two k steps, so each step needs one barrier, and that barrier orders LDS
// only (the registers' loads of the step after next stay in flight across
// it). A load cursor runs one step ahead of the MFMAs through the flattened
// sequence of (tile, k step) of this workgroup, so the last step of a tile
// already has the next tile's first step in LDS. The epilogue stages through
// the buffer the last step read. NW wavefronts (four, or eight for two per
// SIMD).
template <int BM, int BN, int WAVES_M, int NW, bool DIRECT>
__global__ __launch_bounds__(NW * 64, 1) void bf16x3_db_kernel(
    const GemmDesc *__restrict__ descs, const RunDesc *__restrict__ runs,
    const int32_t *__restrict__ tile_pair, const int32_t *__restrict__ tile_first,
    int64_t n_tiles, const void *__restrict__ a, const void *__restrict__ b,
    int64_t a_plane, int64_t b_plane, double *__restrict__ c,
    int32_t *__restrict__ range_flag, uint32_t *__restrict__ tile_ctr) {
  constexpr int WAVES_N = NW / WAVES_M;
  constexpr int kT = NW * 64;
  constexpr int WM = BM / WAVES_M, WN = BN / WAVES_N;
  constexpr int TM = WM / 16, TN = WN / 16;
  constexpr int kBuf = 3 * kChunks * (BM + BN);  // 16-byte vectors per buffer
  static_assert(TM >= 1 && TN >= 1, "tile shape");
  static_assert(sizeof(double) * 16 * (BM + 1) <= sizeof(uint4) * kBuf,
                "epilogue staging fits one buffer");
  __shared__ uint4 smem[2 * kBuf];
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int wm0 = (wave % WAVES_M) * WM, wn0 = (wave / WAVES_M) * WN;
  const int lrow = lane & 15, lchunk = lane >> 4;

  int64_t t = tqg::xcd_virtual_id(blockIdx.x, gridDim.x);
  if (t >= n_tiles) return;
  TileRef cur = fetch<BM, BN, kT, DIRECT>(t, descs, runs, tile_pair, tile_first, a, b);

  // load cursor: tile lt (lref), k step ls
  TileRef lref = cur;
  int64_t lt = t;
  int ls = 0;
  StepLoader<BM, kT, DIRECT> la;
  StepLoader<BN, kT, DIRECT> lb;
  la.flag = lb.flag = range_flag;
  auto load = [&]() {
    la.load(lref.a, a_plane, lref.d.k, ls);
    lb.load(lref.b, b_plane, lref.d.k, ls);
  };
  auto advance = [&]() -> bool {
    if (++ls < tile_steps(lref.d)) return true;
    lt += gridDim.x;
    if (lt >= n_tiles) return false;
    lref = fetch<BM, BN, kT, DIRECT>(lt, descs, runs, tile_pair, tile_first, a, b);
    ls = 0;
    return true;
  };
  load();
  la.store(smem);
  lb.store(smem + 3 * kChunks * BM);
  bool have = advance();  // the registers hold a step not yet in LDS
  if (have) load();
  tqg::lds_barrier();
  int buf = 0;
  while (true) {
    const int steps = tile_steps(cur.d);
    f4 acc[TM][TN];
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
      for (int j = 0; j < TN; ++j) acc[i][j] = f4{0.f, 0.f, 0.f, 0.f};
    for (int s = 0; s < steps; ++s) {
      const uint4 *As = smem + buf * kBuf;
      if (have) {
        // the other buffer was last read before the previous barrier
        uint4 *An = smem + (buf ^ 1) * kBuf;
        la.store(An);
        lb.store(An + 3 * kChunks * BM);
        have = advance();
        if (have) load();
      }
      step_products<BM, BN, TM, TN>(As, As + 3 * kChunks * BM, wm0 + lrow, wn0 + lrow, lchunk,
                                    acc);
      tqg::lds_barrier();
      buf ^= 1;
    }
    // buf now holds the next tile's first step; buf ^ 1 is free for staging
    tqg::store_tile<double, false, BM, BN, WAVES_M, TM, TN, NW>(
        acc, c, cur.d, cur.m0, cur.n0, reinterpret_cast<double *>(smem + (buf ^ 1) * kBuf));
    t += gridDim.x;
    if (t >= n_tiles) break;
    cur = fetch<BM, BN, kT, DIRECT>(t, descs, runs, tile_pair, tile_first, a, b);
  }
}

// As bf16x3_db_kernel, with the loads two k steps ahead of the MFMAs: two
// register sets take turns, so a load issued at the start of step g is stored
// to LDS at the start of step g + 2 (two steps of MFMAs hide its latency where
// the single set gives one). The work of the workgroup is one flat sequence
// of k steps over its tiles, walked two steps per loop turn (one per
// register set); a tile ends with its epilogue inside the step that finishes
// it. NW wavefronts: with eight, a 256 x 128 tile keeps 64 x 64 wavefront
// tiles and two wavefronts per SIMD at one workgroup per CU. Packed operands
// only: two sets of in-place FP64 values do not fit two wavefronts per SIMD
// (those configurations run bf16x3_db_kernel with eight wavefronts instead).
template <int BM, int BN, int WAVES_M, int NW>
__global__ __launch_bounds__(NW * 64, 1) void bf16x3_db2_kernel(
    const GemmDesc *__restrict__ descs, const RunDesc *__restrict__ runs,
    const int32_t *__restrict__ tile_pair, const int32_t *__restrict__ tile_first,
    int64_t n_tiles, const void *__restrict__ a, const void *__restrict__ b,
    int64_t a_plane, int64_t b_plane, double *__restrict__ c,
    int32_t *__restrict__ range_flag, uint32_t *__restrict__ tile_ctr) {
  constexpr int WAVES_N = NW / WAVES_M;
  constexpr int kT = NW * 64;
  constexpr int WM = BM / WAVES_M, WN = BN / WAVES_N;
  constexpr int TM = WM / 16, TN = WN / 16;
  constexpr int kBuf = 3 * kChunks * (BM + BN);  // 16-byte vectors per buffer
  static_assert(TM >= 1 && TN >= 1, "tile shape");
  static_assert(sizeof(double) * 16 * (BM + 1) <= sizeof(uint4) * kBuf,
                "epilogue staging fits one buffer");
  using LA = StepLoader<BM, kT, false>;
  using LB = StepLoader<BN, kT, false>;
  __shared__ uint4 smem[2 * kBuf];
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int wm0 = (wave % WAVES_M) * WM, wn0 = (wave / WAVES_M) * WN;
  const int lrow = lane & 15, lchunk = lane >> 4;

  int64_t t = tqg::xcd_virtual_id(blockIdx.x, gridDim.x);
  if (t >= n_tiles) return;
  TileRef cur = fetch<BM, BN, kT, false>(t, descs, runs, tile_pair, tile_first, a, b);

  // load cursor: the next step to load is step ls of tile lt (lref)
  TileRef lref = cur;
  int64_t lt = t;
  int ls = 0;
  bool lvalid = true;
  // loads the step at the cursor into (la, lb) and advances the cursor;
  // false once the workgroup's steps are all loaded
  auto load_next = [&](LA &la, LB &lb) -> bool {
    if (!lvalid) return false;
    la.load(lref.a, a_plane, lref.d.k, ls);
    lb.load(lref.b, b_plane, lref.d.k, ls);
    if (++ls >= tile_steps(lref.d)) {
      lt += gridDim.x;
      lvalid = lt < n_tiles;
      if (lvalid) lref = fetch<BM, BN, kT, false>(lt, descs, runs, tile_pair, tile_first, a, b);
      ls = 0;
    }
    return true;
  };
  LA xa, ya;
  LB xb, yb;
  bool hx = load_next(xa, xb);  // step 0
  bool hy = load_next(ya, yb);  // step 1
  xa.store(smem);
  xb.store(smem + 3 * kChunks * BM);
  hx = load_next(xa, xb);  // step 2
  tqg::lds_barrier();

  int buf = 0, s = 0, steps = tile_steps(cur.d);
  f4 acc[TM][TN];
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j) acc[i][j] = f4{0.f, 0.f, 0.f, 0.f};
  // one k step: the register set (la, lb) holds the step after this one; it
  // goes to the other buffer (last read before the previous barrier) and is
  // refilled with the step three ahead. False once the workgroup is done.
  auto step = [&](LA &la, LB &lb, bool &h) -> bool {
    if (h) {
      uint4 *An = smem + (buf ^ 1) * kBuf;
      la.store(An);
      lb.store(An + 3 * kChunks * BM);
      h = load_next(la, lb);
    }
    const uint4 *As = smem + buf * kBuf;
    step_products<BM, BN, TM, TN>(As, As + 3 * kChunks * BM, wm0 + lrow, wn0 + lrow, lchunk,
                                  acc);
    tqg::lds_barrier();
    buf ^= 1;
    if (++s < steps) return true;
    // buf holds the next step; buf ^ 1 is free for staging
    tqg::store_tile<double, false, BM, BN, WAVES_M, TM, TN, NW>(
        acc, c, cur.d, cur.m0, cur.n0, reinterpret_cast<double *>(smem + (buf ^ 1) * kBuf));
    t += gridDim.x;
    if (t >= n_tiles) return false;
    cur = fetch<BM, BN, kT, false>(t, descs, runs, tile_pair, tile_first, a, b);
    s = 0;
    steps = tile_steps(cur.d);
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
      for (int j = 0; j < TN; ++j) acc[i][j] = f4{0.f, 0.f, 0.f, 0.f};
    return true;
  };
  while (step(ya, yb, hy) && step(xa, xb, hx)) {
  }
}

struct Cfg {
  int bm, bn, min_waves;
};
// keep in step with the switch in tq_launch_bf16x3_gemm(); configurations 9
// and 10 are the double-buffered kernel, 11 and 12 its two-step-ahead form
// (eight wavefronts)
constexpr Cfg kCfgs[] = {{64, 64, 1},   {128, 64, 1},  {64, 128, 1},  {128, 128, 1},
                         {256, 32, 1},  {32, 32, 1},   {128, 16, 1},  {256, 16, 1},
                         {128, 128, 2}, {256, 128, 1}, {128, 256, 1}, {256, 128, 2},
                         {128, 256, 2}};
constexpr int kNCfg = sizeof(kCfgs) / sizeof(kCfgs[0]);

typedef void (*KernelFn)(const GemmDesc *, const RunDesc *, const int32_t *, const int32_t *,
                         int64_t, const void *, const void *, int64_t, int64_t, double *,
                         int32_t *, uint32_t *);

template <int BM, int BN, int WAVES_M, int MINW, int DB, bool DIRECT>
tq_status launch(tq_handle_t h, const GemmDesc *descs, const RunDesc *runs,
                 const int32_t *tile_pair, const int32_t *tile_first, int64_t n_tiles,
                 const void *a, const void *b, int64_t ap, int64_t bp, double *c,
                 int32_t *range_flag, uint32_t *tile_ctr) {
  KernelFn kernel;
  const int threads = DB == 2 ? 512 : kThreads;
  if constexpr (DB == 2 && DIRECT)
    kernel = bf16x3_db_kernel<BM, BN, WAVES_M, 8, true>;
  else if constexpr (DB == 2)
    kernel = bf16x3_db2_kernel<BM, BN, WAVES_M, 8>;
  else if constexpr (DB == 1)
    kernel = bf16x3_db_kernel<BM, BN, WAVES_M, 4, DIRECT>;
  else
    kernel = bf16x3_gemm_kernel<BM, BN, WAVES_M, MINW, DIRECT>;
  // persistent grid: exactly the workgroups that can be resident at once
  // (occupancy from registers and LDS), a multiple of the 8 XCDs
  static int per_cu = 0;
  if (per_cu == 0) {
    int occ = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, kernel, threads, 0) != hipSuccess ||
        occ < 1)
      occ = 1;
    per_cu = occ;
  }
  int64_t grid = std::min<int64_t>(n_tiles, (int64_t)h->n_cu * per_cu);
  if (grid >= 8) grid -= grid % 8;
  if (grid < 1) grid = 1;
  // per-XCD counters for the single-buffered kernels (the double-buffered
  // ones walk a fixed sequence of tiles), with every XCD holding workgroups
  uint32_t *ctr = (DB == 0 && grid % kTileCounters == 0) ? tile_ctr : nullptr;
  if (ctr)
    hipLaunchKernelGGL(tqg::zero_tile_counters_kernel, dim3(1), dim3(64), 0, h->stream, ctr);
  hipLaunchKernelGGL(kernel, dim3((unsigned)grid), dim3(threads), 0, h->stream, descs, runs,
                     tile_pair, tile_first, n_tiles, a, b, ap, bp, c, range_flag, ctr);
  TQ_HIP(h, hipGetLastError());
  return TQ_SUCCESS;
}

template <bool DIRECT>
tq_status dispatch(tq_handle_t h, int cfg, const GemmDesc *descs, const RunDesc *runs,
                   const int32_t *tile_pair, const int32_t *tile_first, int64_t n_tiles,
                   const void *a, const void *b, int64_t a_plane, int64_t b_plane, double *c,
                   int32_t *range_flag, uint32_t *tile_ctr) {
  switch (cfg) {
#define TQ_CFG(I, BM, BN, WM, MINW, DB)                                                       \
  case I:                                                                                     \
    return launch<BM, BN, WM, MINW, DB, DIRECT>(h, descs, runs, tile_pair, tile_first, n_tiles, \
                                                a, b, a_plane, b_plane, c, range_flag,          \
                                                tile_ctr);
    TQ_CFG(0, 64, 64, 2, 1, 0)
    TQ_CFG(1, 128, 64, 2, 1, 0)
    TQ_CFG(2, 64, 128, 2, 1, 0)
    TQ_CFG(3, 128, 128, 2, 1, 0)
    TQ_CFG(4, 256, 32, 4, 1, 0)
    TQ_CFG(5, 32, 32, 2, 1, 0)
    TQ_CFG(6, 128, 16, 4, 1, 0)
    TQ_CFG(7, 256, 16, 4, 1, 0)
    TQ_CFG(8, 128, 128, 2, 2, 0)
    TQ_CFG(9, 256, 128, 2, 1, 1)
    TQ_CFG(10, 128, 256, 2, 1, 1)
    TQ_CFG(11, 256, 128, 4, 1, 2)
    TQ_CFG(12, 128, 256, 2, 1, 2)
#undef TQ_CFG
  }
  return tq_fail(h, TQ_ERROR_INVALID_VALUE, "gemm: unknown tile configuration");
}

}  // namespace

int tq_bf16x3_n_configs() { return kNCfg; }

GemmConfig tq_bf16x3_config(int cfg) {
  if (cfg < 0 || cfg >= kNCfg) cfg = 0;
  return GemmConfig{kCfgs[cfg].bm, kCfgs[cfg].bn, kBK, 0, kCfgs[cfg].min_waves};
}

tq_status tq_launch_bf16x3_gemm(tq_handle_t h, int cfg, bool direct, const GemmDesc *descs,
                                const RunDesc *runs, const int32_t *tile_pair,
                                const int32_t *tile_first, int64_t n_tiles, const void *a,
                                const void *b, int64_t a_plane, int64_t b_plane, double *c,
                                int32_t *range_flag, uint32_t *tile_ctr) {
  if (n_tiles <= 0) return TQ_SUCCESS;
  if (direct)
    return dispatch<true>(h, cfg, descs, runs, tile_pair, tile_first, n_tiles, a, b, a_plane,
                          b_plane, c, range_flag, tile_ctr);
  return dispatch<false>(h, cfg, descs, runs, tile_pair, tile_first, n_tiles, a, b, a_plane,
                         b_plane, c, range_flag, tile_ctr);
}
