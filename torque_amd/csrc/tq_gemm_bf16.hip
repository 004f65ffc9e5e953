// tq_gemm_bf16.hip -- the split-BF16 tier of the grouped GEMM.
//
// Each FP64 operand value x is stored as three BF16 pieces hi + mid + lo
// (planes written by the split pack, k fastest, rows padded to multiples of
// 8). The product keeps the six piece products of order >= 2^-16 relative:
//   hi*hi + hi*mid + mid*hi + hi*lo + mid*mid + lo*hi
// on v_mfma_f32_16x16x32_bf16 with FP32 accumulation; the dropped terms
// (mid*lo, lo*mid, lo*lo) are below 2^-23 of |a||b| and the pieces represent
// each operand to 2^-24 relative, so the error per product term is about
// 2^-22 |a||b| plus FP32 accumulation. Results are written as FP64.
//
// Same work list, tile mapping and persistence as the FP64 kernel
// (tq_gemm.hip); operands are always the packed split panels.
#include "tq_gemm_common.h"

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f4 __attribute__((ext_vector_type(4)));

using tqg::kThreads;
constexpr int kBK = 32;
constexpr int kLDK = kBK + 8;  // LDS row stride in BF16 elements (80 bytes)

// Loads a ROWS x 32 tile of the three planes into LDS (16-byte chunks:
// thread t covers row t / 4, k chunk (t % 4) * 8, stepping 64 rows).
template <int ROWS>
__device__ inline void load_planes(const uint16_t *__restrict__ base, int64_t plane,
                                   int ld, int rows, int row0, int kpad, int k0,
                                   uint4 *lds) {
  const int t = threadIdx.x;
  const int kc = (t & 3) * 8;
#pragma unroll
  for (int r = t >> 2; r < ROWS; r += kThreads / 4) {
    const bool ok = (row0 + r < rows) && (k0 + kc < kpad);
    const uint16_t *p = base + (int64_t)(row0 + r) * ld + k0 + kc;
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      uint4 v = make_uint4(0, 0, 0, 0);
      if (ok) v = *reinterpret_cast<const uint4 *>(p + q * plane);
      lds[(q * ROWS * kLDK + r * kLDK + kc) / 8] = v;
    }
  }
}

template <int BM, int BN>
struct Smem {
  static constexpr size_t kAB = 2 * 3 * (size_t)(BM + BN) * kLDK;
  static constexpr size_t kStage = 8 * 16 * (size_t)(BM + 1);
  static constexpr size_t kBytes = kAB > kStage ? kAB : kStage;
};

template <int BM, int BN, int WAVES_M>
__global__ __launch_bounds__(kThreads) void bf16x3_gemm_kernel(
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
  uint4 *Bs = As + 3 * BM * kLDK / 8;
  double *stage = reinterpret_cast<double *>(smem);
  const __bf16 *Ab = reinterpret_cast<const __bf16 *>(As);
  const __bf16 *Bb = reinterpret_cast<const __bf16 *>(Bs);
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int wm0 = (wave % WAVES_M) * WM, wn0 = (wave / WAVES_M) * WN;
  const int64_t vid = tqg::xcd_virtual_id(blockIdx.x, gridDim.x);

  for (int64_t t = vid; t < n_tiles; t += gridDim.x) {
    const int p = tile_pair[t];
    const GemmDesc d = descs[p];
    const int lt = (int)(t - tile_first[p]);
    const int tiles_m = (d.m + BM - 1) / BM;
    const int m0 = (lt % tiles_m) * BM;
    const int n0 = (lt / tiles_m) * BN;
    const int kpad = (d.k + 7) & ~7;
    const uint16_t *pa = a + d.a_off;
    const uint16_t *pb = b + d.b_off;
    const int tm_live = min(TM, max(0, (d.m - m0 - wm0 + 15) / 16));
    const int tn_live = min(TN, max(0, (d.n - n0 - wn0 + 15) / 16));
    f4 acc[TM][TN];
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
      for (int j = 0; j < TN; ++j) acc[i][j] = f4{0.f, 0.f, 0.f, 0.f};

    for (int k0 = 0; k0 < kpad; k0 += kBK) {
      __syncthreads();
      load_planes<BM>(pa, a_plane, d.lda, d.m, m0, kpad, k0, As);
      load_planes<BN>(pb, b_plane, d.ldb, d.n, n0, kpad, k0, Bs);
      __syncthreads();
      const int kk = 8 * (lane >> 4);
      bf16x8 ah[TM], am[TM], al[TM];
#pragma unroll
      for (int i = 0; i < TM; ++i) {
        const int r = (wm0 + i * 16 + (lane & 15)) * kLDK + kk;
        ah[i] = *reinterpret_cast<const bf16x8 *>(Ab + r);
        am[i] = *reinterpret_cast<const bf16x8 *>(Ab + BM * kLDK + r);
        al[i] = *reinterpret_cast<const bf16x8 *>(Ab + 2 * BM * kLDK + r);
      }
#pragma unroll
      for (int j = 0; j < TN; ++j) {
        if (j >= tn_live) continue;
        const int r = (wn0 + j * 16 + (lane & 15)) * kLDK + kk;
        const bf16x8 bh = *reinterpret_cast<const bf16x8 *>(Bb + r);
        const bf16x8 bm = *reinterpret_cast<const bf16x8 *>(Bb + BN * kLDK + r);
        const bf16x8 bl = *reinterpret_cast<const bf16x8 *>(Bb + 2 * BN * kLDK + r);
#pragma unroll
        for (int i = 0; i < TM; ++i) {
          if (i >= tm_live) continue;
          f4 x = acc[i][j];
          // smallest terms first
          x = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[i], bh, x, 0, 0, 0);
          x = __builtin_amdgcn_mfma_f32_16x16x32_bf16(am[i], bm, x, 0, 0, 0);
          x = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], bl, x, 0, 0, 0);
          x = __builtin_amdgcn_mfma_f32_16x16x32_bf16(am[i], bh, x, 0, 0, 0);
          x = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], bm, x, 0, 0, 0);
          x = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], bh, x, 0, 0, 0);
          acc[i][j] = x;
        }
      }
    }
    tqg::store_tile<double, false, BM, BN, WAVES_M, TM, TN>(acc, c, d, m0, n0, stage);
  }
}

struct Cfg {
  int bm, bn;
};
// keep in step with the switch in tq_launch_bf16x3_gemm()
constexpr Cfg kCfgs[] = {{64, 64}, {128, 64}, {64, 128}, {128, 128}, {256, 32},
                         {32, 32}, {128, 16},  {256, 16}};
constexpr int kNCfg = sizeof(kCfgs) / sizeof(kCfgs[0]);

template <int BM, int BN, int WAVES_M>
tq_status launch(tq_handle_t h, const GemmDesc *descs, const int32_t *tile_pair,
                 const int32_t *tile_first, int64_t n_tiles, const uint16_t *a,
                 const uint16_t *b, int64_t ap, int64_t bp, double *c) {
  // persistent grid: exactly the workgroups that can be resident at once
  // (occupancy from registers and LDS), a multiple of the 8 XCDs
  static int per_cu = 0;
  if (per_cu == 0) {
    int occ = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, (bf16x3_gemm_kernel<BM, BN, WAVES_M>), kThreads, 0) !=
            hipSuccess || occ < 1)
      occ = 1;
    per_cu = occ;
  }
  int64_t grid = std::min<int64_t>(n_tiles, (int64_t)h->n_cu * per_cu);
  if (grid >= 8) grid -= grid % 8;
  if (grid < 1) grid = 1;
  hipLaunchKernelGGL((bf16x3_gemm_kernel<BM, BN, WAVES_M>), dim3((unsigned)grid),
                     dim3(kThreads), 0, h->stream, descs, tile_pair, tile_first, n_tiles,
                     a, b, ap, bp, c);
  TQ_HIP(h, hipGetLastError());
  return TQ_SUCCESS;
}

}  // namespace

int tq_bf16x3_n_configs() { return kNCfg; }

GemmConfig tq_bf16x3_config(int cfg) {
  if (cfg < 0 || cfg >= kNCfg) cfg = 0;
  return GemmConfig{kCfgs[cfg].bm, kCfgs[cfg].bn, kBK, 0};
}

tq_status tq_launch_bf16x3_gemm(tq_handle_t h, int cfg, const GemmDesc *descs,
                                const int32_t *tile_pair, const int32_t *tile_first,
                                int64_t n_tiles, const uint16_t *a_split,
                                const uint16_t *b_split, int64_t a_plane,
                                int64_t b_plane, double *c) {
  if (n_tiles <= 0) return TQ_SUCCESS;
  switch (cfg) {
#define TQ_CFG(I, BM, BN, WM)                                                        \
  case I:                                                                            \
    return launch<BM, BN, WM>(h, descs, tile_pair, tile_first, n_tiles, a_split,     \
                              b_split, a_plane, b_plane, c);
    TQ_CFG(0, 64, 64, 2)
    TQ_CFG(1, 128, 64, 2)
    TQ_CFG(2, 64, 128, 2)
    TQ_CFG(3, 128, 128, 2)
    TQ_CFG(4, 256, 32, 2)
    TQ_CFG(5, 32, 32, 2)
    TQ_CFG(6, 128, 16, 4)
    TQ_CFG(7, 256, 16, 4)
#undef TQ_CFG
  }
  return tq_fail(h, TQ_ERROR_INVALID_VALUE, "gemm: unknown tile configuration");
}
