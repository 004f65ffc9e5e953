// tq_gemm.hip -- persistent grouped GEMM over all matched block pairs
// (replaces the per-A-block cublas{S,D}gemmStridedBatched loop of
// src/gpu/block_sparse.cuh:870-930 and the dense cublasXgemm of
// src/gpu/dense.cuh:531-553).
//
// Every pair p computes C_p(MxN) = A_p(MxK) * B_p(NxK)^T, C_p column-major
// at c + c_off (ldc = M) -- the reference's (N,T) product whose output is
// already the column-major [A free..., B free...] block, so the reference's
// reverse reshape disappears. Operands are read in place with either index
// fastest, or from packed panels.
//
// One launch covers every tile of every pair: tile t belongs to pair
// tile_pair[t]; tiles were laid out largest pair first. Workgroups are
// persistent and walk the tile list with an XCD-aware stride so that
// neighbouring tiles (sharing operand panels) land on the same L2.
//
// The wave-tile variant without LDS is in tq_wave_gemm.hip.
//
// Arithmetic: FP64 on v_mfma_f64_16x16x4_f64; FP32 on v_mfma_f32_16x16x4_f32
// (for FP32 data, or FP64 data at the FP32 tier); FP16 data on
// v_mfma_f32_16x16x16_f16 with FP32 accumulation; split-BF16 (three BF16
// pieces per operand, six piece products on v_mfma_f32_16x16x32_bf16 with
// FP32 accumulation) in tq_gemm_bf16.hip.
#include <type_traits>

#include "tq_gemm_common.h"

namespace {

using tqg::kThreads;

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

__device__ inline d4 mfma(double a, double b, d4 c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}
__device__ inline f4 mfma(float a, float b, f4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

typedef _Float16 h4 __attribute__((ext_vector_type(4)));

// FP16: v_mfma_f32_16x16x16_f16, lane l supplying row (or column) l % 16 and
// k = 4 (l / 16) .. 4 (l / 16) + 3 of each operand, FP32 accumulation.
__device__ inline f4 mfma(h4 a, h4 b, f4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x16f16(a, b, c, 0, 0, 0);
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
template <>
struct Acc<_Float16> {
  typedef f4 type;
};

// Loads one ROWS x BK operand tile (rows = m or n) into registers and then
// into LDS laid out [k][row] with row length ROWS + 1. Element e = t + 256 q
// of the tile is (row e % ROWS, k e / ROWS) when the row index is fastest in
// memory (p[row + ld*k]) and (row e / BK, k e % BK) when k is fastest
// (p[k + ld*row]), so consecutive lanes read consecutive addresses.
template <class TIn, class TC, int ROWS, int BK>
struct TileLoader {
  static constexpr int kElems = ROWS * BK;
  static constexpr int kN = kElems >= kThreads ? kElems / kThreads : 1;
  TC v[kN];

  __device__ inline void load(const TIn *__restrict__ base, bool kfast, int ld,
                              int rows, int row0, int kcount, int k0) {
    const int t = threadIdx.x;
#pragma unroll
    for (int q = 0; q < kN; ++q) {
      const int e = t + q * kThreads;
      const int r = kfast ? e / BK : e % ROWS;
      const int k = kfast ? e % BK : e / ROWS;
      const bool ok = e < kElems && row0 + r < rows && k0 + k < kcount;
      const int64_t off = kfast ? (int64_t)(k0 + k) + (int64_t)ld * (row0 + r)
                                : (int64_t)(row0 + r) + (int64_t)ld * (k0 + k);
      v[q] = ok ? (TC)base[off] : (TC)0;
    }
  }

  __device__ inline void store(TC *lds, bool kfast) const {
    const int t = threadIdx.x;
#pragma unroll
    for (int q = 0; q < kN; ++q) {
      const int e = t + q * kThreads;
      if (e >= kElems) continue;
      const int r = kfast ? e / BK : e % ROWS;
      const int k = kfast ? e % BK : e / ROWS;
      lds[k * (ROWS + 1) + r] = v[q];
    }
  }
};

// Where one tile of the work list reads and writes.
template <class TIn>
struct TileMeta {
  GemmDesc d;
  int m0, n0;
  bool a_kfast, b_kfast;
  const TIn *pa, *pb;
};

template <class TIn, int BM, int BN>
__device__ inline TileMeta<TIn> fetch_tile(int64_t t, const GemmDesc *__restrict__ descs,
                                           const int32_t *__restrict__ tile_pair,
                                           const int32_t *__restrict__ tile_first,
                                           const TIn *a, const TIn *a_packed,
                                           const TIn *b, const TIn *b_packed) {
  TileMeta<TIn> m;
  const int p = tile_pair[t];
  m.d = descs[p];
  const int lt = (int)(t - tile_first[p]);
  const int tiles_m = (m.d.m + BM - 1) / BM;
  m.m0 = (lt % tiles_m) * BM;
  m.n0 = (lt / tiles_m) * BN;
  m.a_kfast = m.d.flags & kAKFast;
  m.b_kfast = m.d.flags & kBKFast;
  m.pa = ((m.d.flags & kAPacked) ? a_packed : a) + m.d.a_off;
  m.pb = ((m.d.flags & kBPacked) ? b_packed : b) + m.d.b_off;
  return m;
}

template <class TC, class TOut, int BM, int BN, int BK>
struct Smem {
  static constexpr size_t kAB = sizeof(TC) * (size_t)BK * (BM + 1 + BN + 1);
  static constexpr size_t kStage = sizeof(TOut) * 16 * (size_t)(BM + 1);
  static constexpr size_t kBytes = kAB > kStage ? kAB : kStage;
};

// One workgroup of 4 wavefronts (WAVES_M along M x 4/WAVES_M along N) per
// BM x BN tile; each wavefront owns a (BM/WAVES_M) x (BN*WAVES_M/4) sub-tile
// of 16x16 MFMA blocks. The next k tile is prefetched into registers while
// the current one is multiplied from LDS.
template <class TIn, class TC, class TOut, int BM, int BN, int BK, int WAVES_M>
__global__ __launch_bounds__(kThreads) void grouped_gemm_kernel(
    const GemmDesc *__restrict__ descs, const int32_t *__restrict__ tile_pair,
    const int32_t *__restrict__ tile_first, int64_t n_tiles,
    const TIn *__restrict__ a, const TIn *__restrict__ a_packed,
    const TIn *__restrict__ b, const TIn *__restrict__ b_packed,
    TOut *__restrict__ c) {
  typedef typename Acc<TC>::type AccT;
  constexpr int WAVES_N = 4 / WAVES_M;
  constexpr int WM = BM / WAVES_M, WN = BN / WAVES_N;
  constexpr int TM = WM / 16, TN = WN / 16;
  static_assert(TM >= 1 && TN >= 1 && BK % 4 == 0, "tile shape");
  __shared__ __attribute__((aligned(16))) char smem[Smem<TC, TOut, BM, BN, BK>::kBytes];
  TC *As = reinterpret_cast<TC *>(smem);
  TC *Bs = As + BK * (BM + 1);
  TOut *stage = reinterpret_cast<TOut *>(smem);
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int wm0 = (wave % WAVES_M) * WM, wn0 = (wave / WAVES_M) * WN;
  const int64_t vid = tqg::xcd_virtual_id(blockIdx.x, gridDim.x);

  int64_t t = vid;
  if (t >= n_tiles) return;
  TileMeta<TIn> cur = fetch_tile<TIn, BM, BN>(t, descs, tile_pair, tile_first, a, a_packed,
                                              b, b_packed);
  TileLoader<TIn, TC, BM, BK> la;
  TileLoader<TIn, TC, BN, BK> lb;
  la.load(cur.pa, cur.a_kfast, cur.d.lda, cur.d.m, cur.m0, cur.d.k, 0);
  lb.load(cur.pb, cur.b_kfast, cur.d.ldb, cur.d.n, cur.n0, cur.d.k, 0);
  while (true) {
    const GemmDesc &d = cur.d;
    // MFMA blocks of this wavefront that hold output (wave-uniform)
    const int tm_live = min(TM, max(0, (d.m - cur.m0 - wm0 + 15) / 16));
    const int tn_live = min(TN, max(0, (d.n - cur.n0 - wn0 + 15) / 16));
    AccT acc[TM][TN];
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
      for (int j = 0; j < TN; ++j) acc[i][j] = AccT{0, 0, 0, 0};

    const int64_t t_next = t + gridDim.x;
    TileMeta<TIn> nxt;
    for (int k0 = 0; k0 < d.k; k0 += BK) {
      __syncthreads();  // the previous k tile (or epilogue staging) is done
      la.store(As, cur.a_kfast);
      lb.store(Bs, cur.b_kfast);
      __syncthreads();
      if (k0 + BK < d.k) {
        la.load(cur.pa, cur.a_kfast, d.lda, d.m, cur.m0, d.k, k0 + BK);
        lb.load(cur.pb, cur.b_kfast, d.ldb, d.n, cur.n0, d.k, k0 + BK);
      } else if (t_next < n_tiles) {
        // last k tile: start the next tile's loads so their latency overlaps
        // this tile's MFMAs and epilogue stores
        nxt = fetch_tile<TIn, BM, BN>(t_next, descs, tile_pair, tile_first, a, a_packed,
                                      b, b_packed);
        la.load(nxt.pa, nxt.a_kfast, nxt.d.lda, nxt.d.m, nxt.m0, nxt.d.k, 0);
        lb.load(nxt.pb, nxt.b_kfast, nxt.d.ldb, nxt.d.n, nxt.n0, nxt.d.k, 0);
      }
      if constexpr (std::is_same<TC, _Float16>::value) {
        // 16 k per MFMA; k past K reads the loaders' zeros
        const int ksteps = min(BK / 16, (d.k - k0 + 15) / 16);
        for (int ks = 0; ks < ksteps; ++ks) {
          const int kr = ks * 16 + 4 * (lane >> 4);
          h4 af[TM], bf[TN];
#pragma unroll
          for (int i = 0; i < TM; ++i)
#pragma unroll
            for (int q = 0; q < 4; ++q)
              af[i][q] = As[(kr + q) * (BM + 1) + wm0 + i * 16 + (lane & 15)];
#pragma unroll
          for (int j = 0; j < TN; ++j)
#pragma unroll
            for (int q = 0; q < 4; ++q)
              bf[j][q] = Bs[(kr + q) * (BN + 1) + wn0 + j * 16 + (lane & 15)];
#pragma unroll
          for (int i = 0; i < TM; ++i)
#pragma unroll
            for (int j = 0; j < TN; ++j)
              if (i < tm_live && j < tn_live) acc[i][j] = mfma(af[i], bf[j], acc[i][j]);
        }
      } else {
        const int ksteps = min(BK / 4, (d.k - k0 + 3) / 4);
        for (int ks = 0; ks < ksteps; ++ks) {
          const int kr = ks * 4 + (lane >> 4);
          TC af[TM], bf[TN];
#pragma unroll
          for (int i = 0; i < TM; ++i)
            af[i] = As[kr * (BM + 1) + wm0 + i * 16 + (lane & 15)];
#pragma unroll
          for (int j = 0; j < TN; ++j)
            bf[j] = Bs[kr * (BN + 1) + wn0 + j * 16 + (lane & 15)];
#pragma unroll
          for (int i = 0; i < TM; ++i)
#pragma unroll
            for (int j = 0; j < TN; ++j)
              if (i < tm_live && j < tn_live) acc[i][j] = mfma(af[i], bf[j], acc[i][j]);
        }
      }
    }
    tqg::store_tile<TOut, sizeof(TC) == 8, BM, BN, WAVES_M, TM, TN>(acc, c, d, cur.m0,
                                                                    cur.n0, stage);
    if (t_next >= n_tiles) break;
    t = t_next;
    cur = nxt;
  }
}

struct Cfg {
  int bm, bn, bk;
};
// keep in step with the switch in dispatch()
constexpr Cfg kCfgs[] = {{64, 64, 16},  {128, 64, 16}, {64, 128, 16}, {128, 128, 16},
                         {256, 32, 16}, {32, 256, 16}, {32, 32, 16},  {128, 16, 16},
                         {256, 16, 8},  {512, 16, 4}};
constexpr int kNCfg = sizeof(kCfgs) / sizeof(kCfgs[0]);

template <class TIn, class TC, class TOut, int BM, int BN, int BK, int WAVES_M>
tq_status launch_one(tq_handle_t h, const GemmDesc *descs,
                     const int32_t *tile_pair, const int32_t *tile_first,
                     int64_t n_tiles, const void *a, const void *ap,
                     const void *b, const void *bp, void *c) {
  // persistent grid: a few workgroups per CU, a multiple of the 8 XCDs
  // persistent grid: exactly the workgroups that can be resident at once
  // (occupancy from registers and LDS), a multiple of the 8 XCDs
  static int per_cu = 0;
  if (per_cu == 0) {
    int occ = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, (grouped_gemm_kernel<TIn, TC, TOut, BM, BN, BK, WAVES_M>), kThreads, 0) !=
            hipSuccess || occ < 1)
      occ = 1;
    per_cu = occ;
  }
  int64_t grid = std::min<int64_t>(n_tiles, (int64_t)h->n_cu * per_cu);
  if (grid >= 8) grid -= grid % 8;
  if (grid < 1) grid = 1;
  hipLaunchKernelGGL((grouped_gemm_kernel<TIn, TC, TOut, BM, BN, BK, WAVES_M>),
                     dim3((unsigned)grid), dim3(kThreads), 0, h->stream, descs,
                     tile_pair, tile_first, n_tiles,
                     reinterpret_cast<const TIn *>(a),
                     reinterpret_cast<const TIn *>(ap),
                     reinterpret_cast<const TIn *>(b),
                     reinterpret_cast<const TIn *>(bp),
                     reinterpret_cast<TOut *>(c));
  TQ_HIP(h, hipGetLastError());
  return TQ_SUCCESS;
}

template <class TIn, class TC, class TOut>
tq_status dispatch(tq_handle_t h, int cfg, const GemmDesc *descs,
                   const int32_t *tile_pair, const int32_t *tile_first,
                   int64_t n_tiles, const void *a, const void *ap, const void *b,
                   const void *bp, void *c) {
#define TQ_CFG(I, BM, BN, BK, WM)                                                \
  case I:                                                                        \
    return launch_one<TIn, TC, TOut, BM, BN, BK, WM>(h, descs, tile_pair,        \
                                                     tile_first, n_tiles, a, ap, \
                                                     b, bp, c);
  if (std::is_same<TC, _Float16>::value && cfg >= 0 && cfg < kNCfg && kCfgs[cfg].bk % 16 != 0)
    return tq_fail(h, TQ_ERROR_UNSUPPORTED, "gemm: FP16 needs a 16-value k step");
  if constexpr (std::is_same<TC, _Float16>::value) {
    switch (cfg) {
      TQ_CFG(0, 64, 64, 16, 2)
      TQ_CFG(1, 128, 64, 16, 2)
      TQ_CFG(2, 64, 128, 16, 2)
      TQ_CFG(3, 128, 128, 16, 2)
      TQ_CFG(4, 256, 32, 16, 2)
      TQ_CFG(5, 32, 256, 16, 2)
      TQ_CFG(6, 32, 32, 16, 2)
      TQ_CFG(7, 128, 16, 16, 4)
    }
    return tq_fail(h, TQ_ERROR_INVALID_VALUE, "gemm: unknown tile configuration");
  }
  switch (cfg) {
    TQ_CFG(0, 64, 64, 16, 2)
    TQ_CFG(1, 128, 64, 16, 2)
    TQ_CFG(2, 64, 128, 16, 2)
    TQ_CFG(3, 128, 128, 16, 2)
    TQ_CFG(4, 256, 32, 16, 2)
    TQ_CFG(5, 32, 256, 16, 2)
    TQ_CFG(6, 32, 32, 16, 2)
    TQ_CFG(7, 128, 16, 16, 4)
    TQ_CFG(8, 256, 16, 8, 4)
    TQ_CFG(9, 512, 16, 4, 4)
  }
#undef TQ_CFG
  return tq_fail(h, TQ_ERROR_INVALID_VALUE, "gemm: unknown tile configuration");
}

}  // namespace

int tq_bf16x3_n_configs();
GemmConfig tq_bf16x3_config(int cfg);
int tq_wave_n_configs();
GemmConfig tq_wave_config(int cfg);
tq_status tq_launch_wave_gemm(tq_handle_t h, tq_dtype dt, tq_precision prec, int cfg,
                              const GemmDesc *descs, const RunDesc *runs,
                              const int32_t *tile_pair, const int32_t *tile_first,
                              int64_t n_tiles, const void *a, const void *a_packed,
                              const void *b, const void *b_packed, void *c,
                              uint32_t *tile_ctr);

// FP64 / FP32 configurations: the workgroup-tile kernel's (0 .. kNCfg-1),
// then the wave-tile kernel's (tq_wave_gemm.hip).
int tq_gemm_n_configs(tq_precision prec) {
  if (prec == TQ_PREC_BF16X3) return tq_bf16x3_n_configs();
  return kNCfg + tq_wave_n_configs();
}

GemmConfig tq_gemm_config(tq_precision prec, int cfg) {
  if (prec == TQ_PREC_BF16X3) return tq_bf16x3_config(cfg);
  if (cfg >= kNCfg) {
    GemmConfig g = tq_wave_config(cfg - kNCfg);
    g.wave_tile = 1;
    return g;
  }
  if (cfg < 0) cfg = 0;
  return GemmConfig{kCfgs[cfg].bm, kCfgs[cfg].bn, kCfgs[cfg].bk, 0, 1};
}

tq_status tq_launch_grouped_gemm(tq_handle_t h, tq_dtype dt, tq_precision prec,
                                 int cfg, const GemmDesc *descs, const RunDesc *runs,
                                 const int32_t *tile_pair,
                                 const int32_t *tile_first, int64_t n_tiles,
                                 const void *a, const void *a_packed,
                                 const void *b, const void *b_packed, void *c,
                                 uint32_t *tile_ctr) {
  if (n_tiles <= 0) return TQ_SUCCESS;
  if (prec == TQ_PREC_BF16X3)
    return tq_fail(h, TQ_ERROR_INVALID_VALUE, "split-BF16 tier runs on split panels");
  if (cfg >= kNCfg) {
    if (dt == TQ_F16)
      return tq_fail(h, TQ_ERROR_UNSUPPORTED, "gemm: FP16 has no wave-tile configuration");
    return tq_launch_wave_gemm(h, dt, prec, cfg - kNCfg, descs, runs, tile_pair, tile_first,
                               n_tiles, a, a_packed, b, b_packed, c, tile_ctr);
  }
  if (runs)
    return tq_fail(h, TQ_ERROR_INVALID_VALUE, "gemm: runs need a wave-tile configuration");
  if (dt == TQ_F16)
    return dispatch<_Float16, _Float16, _Float16>(h, cfg, descs, tile_pair, tile_first,
                                                  n_tiles, a, a_packed, b, b_packed, c);
  if (dt == TQ_F32)
    return dispatch<float, float, float>(h, cfg, descs, tile_pair, tile_first,
                                         n_tiles, a, a_packed, b, b_packed, c);
  if (prec == TQ_PREC_F32)
    return dispatch<double, float, double>(h, cfg, descs, tile_pair, tile_first,
                                           n_tiles, a, a_packed, b, b_packed, c);
  return dispatch<double, double, double>(h, cfg, descs, tile_pair, tile_first,
                                          n_tiles, a, a_packed, b, b_packed, c);
}

extern "C" {

int32_t tq_gemm_config_count(tq_precision precision) {
  return tq_gemm_n_configs(precision);
}

tq_status tq_gemm_config_describe(tq_precision precision, int32_t config,
                                  int32_t *bm, int32_t *bn, int32_t *bk,
                                  int32_t *wave_tile, int32_t *min_waves) {
  if (config < 0 || config >= tq_gemm_n_configs(precision))
    return TQ_ERROR_INVALID_VALUE;
  GemmConfig c = tq_gemm_config(precision, config);
  if (bm) *bm = c.bm;
  if (bn) *bn = c.bn;
  if (bk) *bk = c.bk;
  if (wave_tile) *wave_tile = c.wave_tile;
  if (min_waves) *min_waves = c.min_waves;
  return TQ_SUCCESS;
}

}  // extern "C"
