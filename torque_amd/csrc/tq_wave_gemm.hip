// tq_wave_gemm.hip -- grouped GEMM with one wavefront per output tile and
// operands loaded straight from global memory into MFMA registers.
//
// Same contract as the workgroup-tile kernel in tq_gemm.hip (every pair p:
// C_p(MxN) = A_p(MxK) B_p(NxK)^T, C_p column-major with ldc = M, operands in
// place with either index fastest or from packed panels), different
// schedule: each wavefront owns a (16 TM) x (16 TN) tile of one pair and
// walks K on its own, so there is no LDS and no workgroup barrier. Four
// independent wavefronts share a workgroup only to fill the SIMDs.
//
// The FP64 MFMA (v_mfma_f64_16x16x4_f64) takes 64 cycles on a CDNA4 SIMD,
// so a 2x2 register tile consumes operands at about 32 B/clk per CU: what the
// L1/L2 path delivers without staging in LDS. Small block pairs (the bulk of
// a U(1) contraction's pairs) then cost no barrier or staging round trips.
//
// k mapping: MFMA step s of a k tile sums k = k0 + 4s + g over the four
// 16-lane groups g, the grouping the workgroup-tile kernel uses, so every
// configuration of both kernels adds the same products in the same order and
// gives bit-identical results. For a fixed step the 16 lanes of a group read
// 16 consecutive rows (one cache line when the row index is fastest) and the
// four groups 4 consecutive k of each row (the next steps read the rest of
// the same lines when k is fastest).
#include "tq_gemm_common.h"

namespace {

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

__device__ inline d4 mfma(double a, double b, d4 c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}
__device__ inline f4 mfma(float a, float b, f4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

template <class TC>
struct Acc;
template <>
struct Acc<double> {
  typedef d4 type;
};
template <>
struct Acc<float> {
  typedef f4 type;
};

__device__ inline int64_t uniform64(int64_t v) {
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
  const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)((uint64_t)v >> 32));
  return (int64_t)(((uint64_t)hi << 32) | lo);
}

// Where one wave tile reads and writes (all wave-uniform). pa and pb point
// at the tile's first row of A and of B; element (row, k) of the tile lies
// at row * sr + k * sk from there.
template <class TIn>
struct WaveTile {
  const TIn *pa, *pb;
  int32_t sra, ska, srb, skb;
  int64_t c_off;
  int m, n, k, m0, n0;
};

template <class TIn, int BM, int BN>
__device__ inline WaveTile<TIn> wave_tile(int64_t t, const GemmDesc *__restrict__ descs,
                                          const int32_t *__restrict__ tile_pair,
                                          const int32_t *__restrict__ tile_first,
                                          const TIn *a, const TIn *a_packed, const TIn *b,
                                          const TIn *b_packed) {
  WaveTile<TIn> w;
  const int p = __builtin_amdgcn_readfirstlane(tile_pair[t]);
  const GemmDesc d = descs[p];
  const int lt = (int)(t - tile_first[p]);
  const int tiles_m = (d.m + BM - 1) / BM;
  w.m = d.m;
  w.n = d.n;
  w.k = d.k;
  w.m0 = (lt % tiles_m) * BM;
  w.n0 = (lt / tiles_m) * BN;
  w.c_off = d.c_off;
  const bool akf = d.flags & kAKFast, bkf = d.flags & kBKFast;
  w.sra = akf ? d.lda : 1;
  w.ska = akf ? 1 : d.lda;
  w.srb = bkf ? d.ldb : 1;
  w.skb = bkf ? 1 : d.ldb;
  w.pa = ((d.flags & kAPacked) ? a_packed : a) + d.a_off + (int64_t)w.m0 * w.sra;
  w.pb = ((d.flags & kBPacked) ? b_packed : b) + d.b_off + (int64_t)w.n0 * w.srb;
  return w;
}

// The A side and output of tile t of a run (B comes from run_columns).
template <class TIn, int BM, int BN>
__device__ inline WaveTile<TIn> run_tile(int64_t t, const RunDesc *__restrict__ runs,
                                         const int32_t *__restrict__ tile_pair,
                                         const int32_t *__restrict__ tile_first, const TIn *a,
                                         const TIn *a_packed, RunDesc &r) {
  WaveTile<TIn> w;
  const int u = __builtin_amdgcn_readfirstlane(tile_pair[t]);
  r = runs[u];
  const int lt = (int)(t - tile_first[u]);
  const int tiles_m = (r.m + BM - 1) / BM;
  w.m = r.m;
  w.n = r.n;
  w.k = r.k;
  w.m0 = (lt % tiles_m) * BM;
  w.n0 = (lt / tiles_m) * BN;
  w.c_off = r.c_off;
  const bool akf = r.flags & kAKFast;
  w.sra = akf ? r.lda : 1;
  w.ska = akf ? 1 : r.lda;
  w.srb = w.skb = 0;
  w.pa = ((r.flags & kAPacked) ? a_packed : a) + r.a_off + (int64_t)w.m0 * w.sra;
  w.pb = nullptr;
  return w;
}

// Byte offsets, from element (row0, k0) of the operand, of this lane's KPL
// k values (k0 + 4s + g) in each of its T row blocks, for a k tile that lies
// inside K. Rows past the block are clamped to its last row (their products
// only reach output rows that are never stored). Offsets stay below 2^32
// because the leading dimensions are below kWaveMaxLd.
template <class TIn, int T, int KPL>
__device__ inline void tile_offsets(uint32_t (&off)[T][KPL], int32_t sr, int32_t sk, int rows,
                                    int row0, int lane) {
  const uint32_t g = lane >> 4;
#pragma unroll
  for (int i = 0; i < T; ++i) {
    const uint32_t rl = (uint32_t)(min(row0 + 16 * i + (lane & 15), rows - 1) - row0);
#pragma unroll
    for (int s = 0; s < KPL; ++s)
      off[i][s] = (rl * (uint32_t)sr + (g + 4 * s) * (uint32_t)sk) * (uint32_t)sizeof(TIn);
  }
}

// Loads one k tile of one operand: base is element (row0, k0) (wave-uniform,
// so the loads take a scalar base and a 32-bit lane offset). In a k tile
// that crosses K (rem = K - k0 < 4 KPL) k is clamped to K - 1 and the MFMAs
// zero those values (tail_select); either way the loads are one straight run
// whose results nothing consumes before the next k tile, so they stay in
// flight under the current tile's MFMAs.
template <class TIn, class TC, int T, int KPL>
__device__ inline void load_k(TC (&r)[T][KPL], const TIn *base, const uint32_t (&off)[T][KPL],
                              int32_t sk, int rem, int lane) {
  const char *pb = reinterpret_cast<const char *>(base);
  uint32_t o[T][KPL];
#pragma unroll
  for (int i = 0; i < T; ++i)
#pragma unroll
    for (int s = 0; s < KPL; ++s) o[i][s] = off[i][s];
  if (rem < 4 * KPL) {
    // move each clamped k back onto k = rem - 1 of the same row
    const int g = lane >> 4;
#pragma unroll
    for (int s = 0; s < KPL; ++s) {
      const int over = max(0, g + 4 * s - (rem - 1));
      const uint32_t back = (uint32_t)over * (uint32_t)sk * (uint32_t)sizeof(TIn);
#pragma unroll
      for (int i = 0; i < T; ++i) o[i][s] -= back;
    }
  }
#pragma unroll
  for (int i = 0; i < T; ++i)
#pragma unroll
    for (int s = 0; s < KPL; ++s) r[i][s] = (TC)*reinterpret_cast<const TIn *>(pb + o[i][s]);
}

// The B operand of a run tile, per lane: for each of the lane's TN column
// blocks, element (column, k = 0) of its column and the k stride. Columns
// past the run are clamped to its last column (their products only reach
// unstored output). The run's pairs are walked in order, summing their N.
template <class TIn, int TN>
struct LaneCols {
  const TIn *p[TN];
  int32_t sk[TN];
};

template <class TIn, int TN>
__device__ inline void run_columns(LaneCols<TIn, TN> &lc, const RunDesc &r, int n0,
                                   const GemmDesc *__restrict__ descs, const TIn *b,
                                   const TIn *b_packed, int lane) {
  int col[TN];
#pragma unroll
  for (int j = 0; j < TN; ++j) {
    col[j] = min(n0 + 16 * j + (lane & 15), r.n - 1);
    lc.p[j] = b;
    lc.sk[j] = 1;
  }
  int start = 0;
  for (int q = 0; q < r.np; ++q) {
    const GemmDesc d = descs[r.p0 + q];
    const bool bkf = d.flags & kBKFast;
    const TIn *base = ((d.flags & kBPacked) ? b_packed : b) + d.b_off;
#pragma unroll
    for (int j = 0; j < TN; ++j)
      if (col[j] >= start && col[j] < start + d.n) {
        lc.p[j] = base + (int64_t)(col[j] - start) * (bkf ? d.ldb : 1);
        lc.sk[j] = bkf ? 1 : d.ldb;
      }
    start += d.n;
  }
}

// Loads one k tile from per-lane columns: the lane's KPL k values (k0 + 4s +
// g) of each of its T columns, k clamped to K - 1 in the k tile that crosses
// K (tail_select zeroes those values).
template <class TIn, class TC, int T, int KPL>
__device__ inline void load_k_cols(TC (&r)[T][KPL], const LaneCols<TIn, T> &lc, int k0, int K,
                                   int lane) {
  const int g = lane >> 4;
#pragma unroll
  for (int s = 0; s < KPL; ++s) {
    const int64_t kk = min(k0 + 4 * s + g, K - 1);
#pragma unroll
    for (int j = 0; j < T; ++j) r[j][s] = (TC)lc.p[j][kk * lc.sk[j]];
  }
}

// Zeroes this lane's operand values past K (kvalid = K - k0 - g) in the k
// tile that crosses K.
template <class TC, int T, int KPL>
__device__ inline void tail_select(TC (&r)[T][KPL], int kvalid) {
#pragma unroll
  for (int s = 0; s < KPL; ++s)
#pragma unroll
    for (int i = 0; i < T; ++i) r[i][s] = 4 * s < kvalid ? r[i][s] : (TC)0;
}

// Two output elements side by side in memory, stored as one vector (the
// address is aligned to the element only).
template <class T>
struct Pair {
  typedef T type __attribute__((ext_vector_type(2), aligned(sizeof(T))));
};

// Accumulator registers 2h and 2h + 1 (x, y) of a 16 x 16 MFMA block as two
// vertically adjacent elements (v0 above v1) of this lane's column, and the
// row of v0 within the block. The FP32-accumulating MFMAs hold rows
// 4g + r (g = lane / 16), already adjacent. The FP64 MFMA holds rows
// g + 4r, so lanes swap halves across 16-lane rows (permlane16 swap: rows
// 1 and 3 of x with rows 0 and 2 of y): afterwards group g holds rows
// b, b + 1 with b = 8h + {0, 4, 2, 6}[g]. Every lane runs the swaps.
template <bool F64, class TOut, class TC>
__device__ inline void row_pair(TC x, TC y, int lane, int h, TOut &v0, TOut &v1, int &row) {
  const int g = lane >> 4;
  if constexpr (F64) {
    typedef uint32_t u2 __attribute__((ext_vector_type(2)));
    const u2 xb = __builtin_bit_cast(u2, (double)x), yb = __builtin_bit_cast(u2, (double)y);
    const auto lo = __builtin_amdgcn_permlane16_swap(xb.x, yb.x, false, false);
    const auto hi = __builtin_amdgcn_permlane16_swap(xb.y, yb.y, false, false);
    v0 = (TOut)__builtin_bit_cast(double, u2{lo[0], hi[0]});
    v1 = (TOut)__builtin_bit_cast(double, u2{lo[1], hi[1]});
    row = 8 * h + (((g & 1) << 2) | (g & 2));
  } else {
    v0 = (TOut)x;
    v1 = (TOut)y;
    row = 4 * g + 2 * h;
  }
}

// Tile order. With counters (the 64 x 64 configurations), every wavefront
// takes its next tile from its XCD's counter (tqg::stripe_tile): the
// wavefronts of an XCD then work through neighbouring tiles, which share
// their operand blocks, at the same time, and find them in that XCD's L2. A
// fixed assignment (tile t + k * stride to one wavefront) drifts apart as
// tiles of different K finish, and loses that sharing. Without counters each
// XCD still walks a contiguous stretch of the list per pass of the grid.
// Either way every tile is computed once, the same way: the order changes no
// bits.
__device__ inline int64_t grab_tile(uint32_t *ctr, int part) {
  uint32_t c = 0;
  if ((threadIdx.x & 63) == 0) c = atomicAdd(ctr + part, 1u);
  return tqg::stripe_tile(__builtin_amdgcn_readfirstlane(c), part);
}

// MINW: wavefronts per SIMD the register allocation must allow; 2 keeps a
// second wavefront resident to cover the first one's operand waits. RUNS:
// the tile list indexes runs (RunDesc) rather than pairs. tile_ctr: the
// XCDs' tile counters (zeroed), or null for the fixed assignment.
template <class TIn, class TC, class TOut, int TM, int TN, int KPL, int MINW, bool RUNS>
__global__ __launch_bounds__(256, MINW) void wave_gemm_kernel(
    const GemmDesc *__restrict__ descs, const RunDesc *__restrict__ runs,
    const int32_t *__restrict__ tile_pair, const int32_t *__restrict__ tile_first,
    int64_t n_tiles, const TIn *__restrict__ a, const TIn *__restrict__ a_packed,
    const TIn *__restrict__ b, const TIn *__restrict__ b_packed, TOut *__restrict__ c,
    uint32_t *__restrict__ tile_ctr) {
  typedef typename Acc<TC>::type AccT;
  constexpr int BM = 16 * TM, BN = 16 * TN, BK = 4 * KPL;
  constexpr bool kF64 = sizeof(TC) == 8;
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int64_t stride = (int64_t)gridDim.x * 4;
  const int part = (int)(blockIdx.x % kTileCounters);
  int64_t t = tile_ctr ? uniform64(grab_tile(tile_ctr, part))
                       : uniform64(tqg::xcd_virtual_id(blockIdx.x, gridDim.x) * 4 + wave);
  if (t >= n_tiles) return;

  RunDesc run_cur, run_nxt;
  WaveTile<TIn> cur;
  if constexpr (RUNS)
    cur = run_tile<TIn, BM, BN>(t, runs, tile_pair, tile_first, a, a_packed, run_cur);
  else
    cur = wave_tile<TIn, BM, BN>(t, descs, tile_pair, tile_first, a, a_packed, b, b_packed);
  uint32_t offa[TM][KPL], offb[TN][KPL];
  LaneCols<TIn, TN> cols;
  tile_offsets<TIn, TM, KPL>(offa, cur.sra, cur.ska, cur.m, cur.m0, lane);
  if constexpr (RUNS)
    run_columns<TIn, TN>(cols, run_cur, cur.n0, descs, b, b_packed, lane);
  else
    tile_offsets<TIn, TN, KPL>(offb, cur.srb, cur.skb, cur.n, cur.n0, lane);
  // (na, nb) always hold the prefetched next k tile; (ra, rb) the current one
  TC ra[TM][KPL], rb[TN][KPL], na[TM][KPL], nb[TN][KPL];
  load_k<TIn, TC, TM, KPL>(na, cur.pa, offa, cur.ska, cur.k, lane);
  if constexpr (RUNS)
    load_k_cols<TIn, TC, TN, KPL>(nb, cols, 0, cur.k, lane);
  else
    load_k<TIn, TC, TN, KPL>(nb, cur.pb, offb, cur.skb, cur.k, lane);
  while (true) {
    // the next wave tile is known from the start, so the prefetch of its first
    // k tile needs no wait (the last tile prefetches itself again)
    const int64_t t_next = tile_ctr ? uniform64(grab_tile(tile_ctr, part)) : t + stride;
    const bool more = t_next < n_tiles;
    WaveTile<TIn> nxt = cur;
    if constexpr (RUNS) run_nxt = run_cur;
    if (more) {
      if constexpr (RUNS)
        nxt = run_tile<TIn, BM, BN>(t_next, runs, tile_pair, tile_first, a, a_packed, run_nxt);
      else
        nxt = wave_tile<TIn, BM, BN>(t_next, descs, tile_pair, tile_first, a, a_packed, b,
                                     b_packed);
    }
    AccT acc[TM][TN];
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
      for (int j = 0; j < TN; ++j) acc[i][j] = AccT{0, 0, 0, 0};
    for (int k0 = 0; k0 < cur.k; k0 += BK) {
#pragma unroll
      for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int s = 0; s < KPL; ++s) ra[i][s] = na[i][s];
#pragma unroll
      for (int j = 0; j < TN; ++j)
#pragma unroll
        for (int s = 0; s < KPL; ++s) rb[j][s] = nb[j][s];
      const int kn = k0 + BK;
      if (kn < cur.k) {
        load_k<TIn, TC, TM, KPL>(na, cur.pa + (int64_t)kn * cur.ska, offa, cur.ska, cur.k - kn,
                                 lane);
        if constexpr (RUNS)
          load_k_cols<TIn, TC, TN, KPL>(nb, cols, kn, cur.k, lane);
        else
          load_k<TIn, TC, TN, KPL>(nb, cur.pb + (int64_t)kn * cur.skb, offb, cur.skb,
                                   cur.k - kn, lane);
      } else {
        // the current tile's loads are all issued: its offsets become the
        // next tile's
        tile_offsets<TIn, TM, KPL>(offa, nxt.sra, nxt.ska, nxt.m, nxt.m0, lane);
        load_k<TIn, TC, TM, KPL>(na, nxt.pa, offa, nxt.ska, nxt.k, lane);
        if constexpr (RUNS) {
          run_columns<TIn, TN>(cols, run_nxt, nxt.n0, descs, b, b_packed, lane);
          load_k_cols<TIn, TC, TN, KPL>(nb, cols, 0, nxt.k, lane);
        } else {
          tile_offsets<TIn, TN, KPL>(offb, nxt.srb, nxt.skb, nxt.n, nxt.n0, lane);
          load_k<TIn, TC, TN, KPL>(nb, nxt.pb, offb, nxt.skb, nxt.k, lane);
        }
      }
      // keep these loads ahead of the MFMAs: their latency hides under them
      __builtin_amdgcn_sched_barrier(0);
      if (kn > cur.k) {
        tail_select<TC, TM, KPL>(ra, cur.k - k0 - (lane >> 4));
        tail_select<TC, TN, KPL>(rb, cur.k - k0 - (lane >> 4));
      }
      // every MFMA block runs (blocks past M or N compute on clamped rows and
      // are not stored), so the MFMA stream has no branches
#pragma unroll
      for (int s = 0; s < KPL; ++s)
#pragma unroll
        for (int i = 0; i < TM; ++i)
#pragma unroll
          for (int j = 0; j < TN; ++j) acc[i][j] = mfma(ra[i][s], rb[j][s], acc[i][j]);
    }
    // column-major store of the tile's valid rows and columns, two adjacent
    // rows of a column per store
    TOut *cb = c + cur.c_off;
#pragma unroll
    for (int j = 0; j < TN; ++j) {
      const int col = cur.n0 + 16 * j + (lane & 15);
#pragma unroll
      for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          TOut v0, v1;
          int row;
          row_pair<kF64, TOut>(acc[i][j][2 * h], acc[i][j][2 * h + 1], lane, h, v0, v1, row);
          row += cur.m0 + 16 * i;
          if (col < cur.n) {
            TOut *p = cb + row + (int64_t)cur.m * col;
            if (row + 1 < cur.m)
              *reinterpret_cast<typename Pair<TOut>::type *>(p) = typename Pair<TOut>::type{v0, v1};
            else if (row < cur.m)
              *p = v0;
          }
        }
    }
    if (!more) break;
    t = t_next;
    cur = nxt;
    if constexpr (RUNS) run_cur = run_nxt;
  }
}

struct WaveCfg {
  int tm, tn, kpl, min_waves;
};
// keep in step with the switch in dispatch()
constexpr WaveCfg kWaveCfgs[] = {{1, 1, 4, 1}, {2, 1, 4, 1}, {1, 2, 4, 1},
                                 {2, 2, 4, 1}, {2, 2, 2, 1}, {4, 2, 2, 1},
                                 {2, 4, 2, 1}, {4, 4, 2, 1}, {4, 4, 2, 2}};
constexpr int kNWave = sizeof(kWaveCfgs) / sizeof(kWaveCfgs[0]);

template <class TIn, class TC, class TOut, int TM, int TN, int KPL, int MINW, bool RUNS>
tq_status launch_one(tq_handle_t h, const GemmDesc *descs, const RunDesc *runs,
                     const int32_t *tile_pair, const int32_t *tile_first, int64_t n_tiles,
                     const void *a, const void *ap, const void *b, const void *bp, void *c,
                     uint32_t *tile_ctr) {
  // persistent grid: the workgroups that fit at once, a multiple of the 8 XCDs
  static int per_cu = 0;
  if (per_cu == 0) {
    int occ = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &occ, (wave_gemm_kernel<TIn, TC, TOut, TM, TN, KPL, MINW, RUNS>), 256, 0) !=
            hipSuccess ||
        occ < 1)
      occ = 1;
    per_cu = occ;
  }
  int64_t grid = std::min<int64_t>(ceil_div(n_tiles, (int64_t)4), (int64_t)h->n_cu * per_cu);
  if (grid >= 8) grid -= grid % 8;
  if (grid < 1) grid = 1;
  // counters for the 64 x 64 tiles (one atomic per smaller tile costs more
  // than it saves), with every XCD holding workgroups
  uint32_t *ctr = (TM * TN >= 16 && grid % kTileCounters == 0) ? tile_ctr : nullptr;
  if (ctr)
    hipLaunchKernelGGL(tqg::zero_tile_counters_kernel, dim3(1), dim3(64), 0, h->stream, ctr);
  hipLaunchKernelGGL((wave_gemm_kernel<TIn, TC, TOut, TM, TN, KPL, MINW, RUNS>),
                     dim3((unsigned)grid), dim3(256), 0, h->stream, descs, runs, tile_pair,
                     tile_first, n_tiles, reinterpret_cast<const TIn *>(a),
                     reinterpret_cast<const TIn *>(ap), reinterpret_cast<const TIn *>(b),
                     reinterpret_cast<const TIn *>(bp), reinterpret_cast<TOut *>(c), ctr);
  TQ_HIP(h, hipGetLastError());
  return TQ_SUCCESS;
}

template <class TIn, class TC, class TOut, int TM, int TN, int KPL, int MINW>
tq_status launch_cfg(tq_handle_t h, const GemmDesc *descs, const RunDesc *runs,
                     const int32_t *tile_pair, const int32_t *tile_first, int64_t n_tiles,
                     const void *a, const void *ap, const void *b, const void *bp, void *c,
                     uint32_t *tile_ctr) {
  if (runs)
    return launch_one<TIn, TC, TOut, TM, TN, KPL, MINW, true>(h, descs, runs, tile_pair,
                                                              tile_first, n_tiles, a, ap, b,
                                                              bp, c, tile_ctr);
  return launch_one<TIn, TC, TOut, TM, TN, KPL, MINW, false>(h, descs, runs, tile_pair,
                                                             tile_first, n_tiles, a, ap, b, bp,
                                                             c, tile_ctr);
}

template <class TIn, class TC, class TOut>
tq_status dispatch(tq_handle_t h, int cfg, const GemmDesc *descs, const RunDesc *runs,
                   const int32_t *tile_pair, const int32_t *tile_first, int64_t n_tiles,
                   const void *a, const void *ap, const void *b, const void *bp, void *c,
                   uint32_t *tile_ctr) {
#define TQ_WCFG(I, TM, TN, KPL, MINW)                                                     \
  case I:                                                                                 \
    return launch_cfg<TIn, TC, TOut, TM, TN, KPL, MINW>(h, descs, runs, tile_pair,        \
                                                        tile_first, n_tiles, a, ap, b, bp, \
                                                        c, tile_ctr);
  switch (cfg) {
    TQ_WCFG(0, 1, 1, 4, 1)
    TQ_WCFG(1, 2, 1, 4, 1)
    TQ_WCFG(2, 1, 2, 4, 1)
    TQ_WCFG(3, 2, 2, 4, 1)
    TQ_WCFG(4, 2, 2, 2, 1)
    TQ_WCFG(5, 4, 2, 2, 1)
    TQ_WCFG(6, 2, 4, 2, 1)
    TQ_WCFG(7, 4, 4, 2, 1)
    TQ_WCFG(8, 4, 4, 2, 2)
  }
#undef TQ_WCFG
  return tq_fail(h, TQ_ERROR_INVALID_VALUE, "gemm: unknown wave tile configuration");
}

}  // namespace

int tq_wave_n_configs() { return kNWave; }

GemmConfig tq_wave_config(int cfg) {
  if (cfg < 0 || cfg >= kNWave) cfg = 0;
  const WaveCfg &w = kWaveCfgs[cfg];
  return GemmConfig{16 * w.tm, 16 * w.tn, 4 * w.kpl, 1, w.min_waves};
}

tq_status tq_launch_wave_gemm(tq_handle_t h, tq_dtype dt, tq_precision prec, int cfg,
                              const GemmDesc *descs, const RunDesc *runs,
                              const int32_t *tile_pair, const int32_t *tile_first,
                              int64_t n_tiles, const void *a, const void *a_packed,
                              const void *b, const void *b_packed, void *c,
                              uint32_t *tile_ctr) {
  if (n_tiles <= 0) return TQ_SUCCESS;
  if (dt == TQ_F32)
    return dispatch<float, float, float>(h, cfg, descs, runs, tile_pair, tile_first, n_tiles,
                                         a, a_packed, b, b_packed, c, tile_ctr);
  if (prec == TQ_PREC_F32)
    return dispatch<double, float, double>(h, cfg, descs, runs, tile_pair, tile_first, n_tiles,
                                           a, a_packed, b, b_packed, c, tile_ctr);
  return dispatch<double, double, double>(h, cfg, descs, runs, tile_pair, tile_first, n_tiles,
                                          a, a_packed, b, b_packed, c, tile_ctr);
}
