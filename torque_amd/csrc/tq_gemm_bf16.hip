// tq_gemm_bf16.hip -- the split-BF16 tier of the grouped GEMM.
//
// Each FP64 operand value x is stored as three BF16 pieces hi + mid + lo in
// three planes written by the split pack (tq_permute.hip). A panel of R rows
// is k-chunked: value (row, k) of a plane sits at (k / 8) * R * 8 + row * 8 +
// k % 8, so the 8 values of one row in one chunk are one 16-byte vector and
// consecutive rows are consecutive vectors. The product keeps the six piece
// products of order >= 2^-16 relative,
//   hi*lo + mid*mid + hi*mid + lo*hi + mid*hi + hi*hi   (A piece * B piece,
// in this order per k step: grouped by B plane, smaller products first),
// on v_mfma_f32_16x16x32_bf16 with FP32 accumulation; the dropped terms
// (mid*lo, lo*mid, lo*lo) are below 2^-23 of |a||b| and the pieces represent
// each operand to 2^-24 relative. Results are written as FP64.
//
// Schedule: persistent workgroups of four wavefronts walk the same tile list
// as the FP64 kernel (tq_gemm.hip). A k step covers 32 values (four chunks):
// the workgroup stages the step's A and B vectors of all three planes in LDS
// ([plane][chunk][row] vectors, so the 16 lanes reading one MFMA operand read
// 256 contiguous bytes), and the vectors of the next step (or of the next
// tile's first step) are loaded into registers while the current step's
// MFMAs run. The six piece products are six passes over the wavefront's
// register tile, so consecutive MFMAs never depend on each other; the three
// A planes stay in registers and the B planes are read one at a time.
#include "tq_gemm_common.h"

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f4 __attribute__((ext_vector_type(4)));

using tqg::kThreads;
constexpr int kBK = 32;     // k per step
constexpr int kChunks = 4;  // 8-value chunks per step

// One operand's vectors of one k step: 3 planes x 4 chunks x ROWS rows of
// 16 bytes; vector v of the step is (row v % ROWS, chunk (v / ROWS) % 4,
// plane v / (4 ROWS)), so consecutive threads read consecutive rows.
template <int ROWS>
struct StepLoader {
  static constexpr int kVec = 3 * kChunks * ROWS;
  static constexpr int kN = (kVec + kThreads - 1) / kThreads;
  uint4 v[kN];

  // base: plane 0 of the panel; rows R of the panel; rows past R are read
  // from row R - 1 (their products only reach unstored output); chunks past
  // the padded K are zeros.
  __device__ inline void load(const uint16_t *__restrict__ base, int64_t plane, int rows,
                              int row0, int n_chunks, int chunk0) {
#pragma unroll
    for (int q = 0; q < kN; ++q) {
      const int e = threadIdx.x + q * kThreads;
      const int r = e % ROWS;
      const int c = (e / ROWS) % kChunks;
      const int p = e / (ROWS * kChunks);
      const int row = min(row0 + r, rows - 1);
      const int chunk = chunk0 + c;
      uint4 x = make_uint4(0, 0, 0, 0);
      if (e < kVec && chunk < n_chunks)
        x = *reinterpret_cast<const uint4 *>(base + p * plane +
                                             ((int64_t)chunk * rows + row) * 8);
      v[q] = x;
    }
  }

  __device__ inline void store(uint4 *lds) const {
#pragma unroll
    for (int q = 0; q < kN; ++q) {
      const int e = threadIdx.x + q * kThreads;
      if (e < kVec) lds[e] = v[q];
    }
  }
};

struct TileRef {
  GemmDesc d;
  int m0, n0;
};

template <int BM, int BN>
__device__ inline TileRef fetch(int64_t t, const GemmDesc *__restrict__ descs,
                                const int32_t *__restrict__ tile_pair,
                                const int32_t *__restrict__ tile_first) {
  TileRef r;
  const int p = tile_pair[t];
  r.d = descs[p];
  const int lt = (int)(t - tile_first[p]);
  const int tiles_m = (r.d.m + BM - 1) / BM;
  r.m0 = (lt % tiles_m) * BM;
  r.n0 = (lt / tiles_m) * BN;
  return r;
}

template <int BM, int BN>
struct Smem {
  static constexpr size_t kAB = sizeof(uint4) * 3 * kChunks * (size_t)(BM + BN);
  static constexpr size_t kStage = sizeof(double) * 16 * (size_t)(BM + 1);
  static constexpr size_t kBytes = kAB > kStage ? kAB : kStage;
};

// MINW: wavefronts per SIMD the register allocation must allow (2 makes
// the 128 x 128 tile fit two workgroups per CU at the price of a few spills).
template <int BM, int BN, int WAVES_M, int MINW>
__global__ __launch_bounds__(kThreads, MINW) void bf16x3_gemm_kernel(
    const GemmDesc *__restrict__ descs, const int32_t *__restrict__ tile_pair,
    const int32_t *__restrict__ tile_first, int64_t n_tiles,
    const uint16_t *__restrict__ a, const uint16_t *__restrict__ b,
    int64_t a_plane, int64_t b_plane, double *__restrict__ c) {
  constexpr int WAVES_N = 4 / WAVES_M;
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

  int64_t t = tqg::xcd_virtual_id(blockIdx.x, gridDim.x);
  if (t >= n_tiles) return;
  TileRef cur = fetch<BM, BN>(t, descs, tile_pair, tile_first);
  StepLoader<BM> la;
  StepLoader<BN> lb;
  la.load(a + cur.d.a_off, a_plane, cur.d.m, cur.m0, cur.d.lda >> 3, 0);
  lb.load(b + cur.d.b_off, b_plane, cur.d.n, cur.n0, cur.d.ldb >> 3, 0);
  while (true) {
    const GemmDesc &d = cur.d;
    // panels store kpad = lda (A) / ldb (B) values per row: kpad / 8 chunks
    const int n_chunks = d.lda >> 3, n_chunks_b = d.ldb >> 3;
    // every MFMA block runs (blocks past M or N compute on clamped rows and
    // are not stored), so the MFMA stream has no branches
    f4 acc[TM][TN];
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
      for (int j = 0; j < TN; ++j) acc[i][j] = f4{0.f, 0.f, 0.f, 0.f};

    const int64_t t_next = t + gridDim.x;
    TileRef nxt = cur;
    for (int ch0 = 0; ch0 < n_chunks; ch0 += kChunks) {
      __syncthreads();  // the previous step (or epilogue staging) is read
      la.store(As);
      lb.store(Bs);
      __syncthreads();
      if (ch0 + kChunks < n_chunks) {
        la.load(a + d.a_off, a_plane, d.m, cur.m0, n_chunks, ch0 + kChunks);
        lb.load(b + d.b_off, b_plane, d.n, cur.n0, n_chunks_b, ch0 + kChunks);
      } else if (t_next < n_tiles) {
        // last step: the next tile's first step loads under these MFMAs and
        // the epilogue
        nxt = fetch<BM, BN>(t_next, descs, tile_pair, tile_first);
        la.load(a + nxt.d.a_off, a_plane, nxt.d.m, nxt.m0, nxt.d.lda >> 3, 0);
        lb.load(b + nxt.d.b_off, b_plane, nxt.d.n, nxt.n0, nxt.d.ldb >> 3, 0);
      }
      // all three A planes stay in registers; B planes are read one at a
      // time (fewer live registers), each followed by its products
      bf16x8 af[3][TM];
#pragma unroll
      for (int p = 0; p < 3; ++p)
#pragma unroll
        for (int i = 0; i < TM; ++i)
          af[p][i] = *reinterpret_cast<const bf16x8 *>(
              &As[(p * kChunks + lchunk) * BM + wm0 + 16 * i + lrow]);
#pragma unroll
      for (int pbq = 2; pbq >= 0; --pbq) {
        bf16x8 bfr[TN];
#pragma unroll
        for (int j = 0; j < TN; ++j)
          bfr[j] = *reinterpret_cast<const bf16x8 *>(
              &Bs[(pbq * kChunks + lchunk) * BN + wn0 + 16 * j + lrow]);
        // A planes paired with this B plane, smaller products first:
        // b_lo: hi; b_mid: mid, hi; b_hi: lo, mid, hi
#pragma unroll
        for (int pa = 2 - pbq; pa >= 0; --pa)
#pragma unroll
          for (int i = 0; i < TM; ++i)
#pragma unroll
            for (int j = 0; j < TN; ++j)
              acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[pa][i], bfr[j],
                                                                  acc[i][j], 0, 0, 0);
      }
    }
    tqg::store_tile<double, false, BM, BN, WAVES_M, TM, TN>(acc, c, d, cur.m0, cur.n0, stage);
    if (t_next >= n_tiles) break;
    t = t_next;
    cur = nxt;
  }
}

struct Cfg {
  int bm, bn, min_waves;
};
// keep in step with the switch in tq_launch_bf16x3_gemm()
constexpr Cfg kCfgs[] = {{64, 64, 1},  {128, 64, 1}, {64, 128, 1}, {128, 128, 1}, {256, 32, 1},
                         {32, 32, 1},  {128, 16, 1}, {256, 16, 1}, {128, 128, 2}};
constexpr int kNCfg = sizeof(kCfgs) / sizeof(kCfgs[0]);

template <int BM, int BN, int WAVES_M, int MINW>
tq_status launch(tq_handle_t h, const GemmDesc *descs, const int32_t *tile_pair,
                 const int32_t *tile_first, int64_t n_tiles, const uint16_t *a,
                 const uint16_t *b, int64_t ap, int64_t bp, double *c) {
  // persistent grid: exactly the workgroups that can be resident at once
  // (occupancy from registers and LDS), a multiple of the 8 XCDs
  static int per_cu = 0;
  if (per_cu == 0) {
    int occ = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ,
                                                     (bf16x3_gemm_kernel<BM, BN, WAVES_M, MINW>),
                                                     kThreads, 0) != hipSuccess ||
        occ < 1)
      occ = 1;
    per_cu = occ;
  }
  int64_t grid = std::min<int64_t>(n_tiles, (int64_t)h->n_cu * per_cu);
  if (grid >= 8) grid -= grid % 8;
  if (grid < 1) grid = 1;
  hipLaunchKernelGGL((bf16x3_gemm_kernel<BM, BN, WAVES_M, MINW>), dim3((unsigned)grid),
                     dim3(kThreads), 0, h->stream, descs, tile_pair, tile_first, n_tiles,
                     a, b, ap, bp, c);
  TQ_HIP(h, hipGetLastError());
  return TQ_SUCCESS;
}

}  // namespace

int tq_bf16x3_n_configs() { return kNCfg; }

GemmConfig tq_bf16x3_config(int cfg) {
  if (cfg < 0 || cfg >= kNCfg) cfg = 0;
  return GemmConfig{kCfgs[cfg].bm, kCfgs[cfg].bn, kBK, 0, kCfgs[cfg].min_waves};
}

tq_status tq_launch_bf16x3_gemm(tq_handle_t h, int cfg, const GemmDesc *descs,
                                const int32_t *tile_pair, const int32_t *tile_first,
                                int64_t n_tiles, const uint16_t *a_split,
                                const uint16_t *b_split, int64_t a_plane,
                                int64_t b_plane, double *c) {
  if (n_tiles <= 0) return TQ_SUCCESS;
  switch (cfg) {
#define TQ_CFG(I, BM, BN, WM, MINW)                                                  \
  case I:                                                                            \
    return launch<BM, BN, WM, MINW>(h, descs, tile_pair, tile_first, n_tiles, a_split, \
                                    b_split, a_plane, b_plane, c);
    TQ_CFG(0, 64, 64, 2, 1)
    TQ_CFG(1, 128, 64, 2, 1)
    TQ_CFG(2, 64, 128, 2, 1)
    TQ_CFG(3, 128, 128, 2, 1)
    TQ_CFG(4, 256, 32, 4, 1)
    TQ_CFG(5, 32, 32, 2, 1)
    TQ_CFG(6, 128, 16, 4, 1)
    TQ_CFG(7, 256, 16, 4, 1)
    TQ_CFG(8, 128, 128, 2, 2)
#undef TQ_CFG
  }
  return tq_fail(h, TQ_ERROR_INVALID_VALUE, "gemm: unknown tile configuration");
}
