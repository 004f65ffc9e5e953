// tq_gemm_common.h -- pieces shared by the grouped-GEMM kernels: the
// XCD-aware walk of the tile list and the LDS-staged epilogue.
#pragma once

#include "tq_common.h"

namespace tqg {

constexpr int kThreads = 256;

// Workgroups are dealt to the 8 XCDs round-robin; this virtual index makes
// each XCD walk a contiguous stretch of the tile list (tiles of one pair share
// operand panels, so they should share an L2).
__device__ inline int64_t xcd_virtual_id(int64_t bid, int64_t grid) {
  const int64_t per = grid / 8;
  if (per == 0 || grid % 8 != 0) return bid;
  return (bid % 8) * per + bid / 8;
}

// Row of accumulator register r in a 16x16 MFMA block for this lane: the
// FP64 MFMA interleaves rows (lane/16 + 4r), the FP32-accumulating ones keep
// four consecutive rows per lane (4*(lane/16) + r). The column is lane % 16.
template <bool kF64Layout>
__device__ inline int acc_row(int lane, int r) {
  return kF64Layout ? (lane >> 4) + 4 * r : 4 * (lane >> 4) + r;
}

// Writes a BM x BN accumulator tile to the column-major output block of the
// pair (ldc = M) through LDS, one 16-column slab at a time: waves owning the
// slab stage it, then all threads write the slab's valid rows x columns in
// linear column-major order, so a wavefront stores 64 consecutive elements
// (contiguous across columns when the tile spans the whole column).
// `stage` holds 16 * (BM + 1) TOut values.
template <class TOut, bool kF64Layout, int BM, int BN, int WAVES_M, int TM, int TN,
          class AccT>
__device__ inline void store_tile(const AccT (&acc)[TM][TN], TOut *__restrict__ c,
                                  const GemmDesc &d, int m0, int n0, TOut *stage) {
  constexpr int WM = BM / WAVES_M;
  constexpr int WN = BN / (4 / WAVES_M);
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int wm0 = (wave % WAVES_M) * WM, wn0 = (wave / WAVES_M) * WN;
  const int rows = min(BM, d.m - m0);
  TOut *cbase = c + d.c_off + m0;
  __syncthreads();  // operand tiles in LDS are no longer read
#pragma unroll
  for (int cb = 0; cb < BN / 16; ++cb) {
    const int col0 = n0 + cb * 16;
    if (col0 >= d.n) break;  // uniform over the workgroup
    const int cols = min(16, d.n - col0);
#pragma unroll
    for (int j = 0; j < TN; ++j) {
      if (wn0 + j * 16 != cb * 16) continue;
      const int col = lane & 15;
#pragma unroll
      for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int r = 0; r < 4; ++r)
          stage[col * (BM + 1) + wm0 + i * 16 + acc_row<kF64Layout>(lane, r)] =
              (TOut)acc[i][j][r];
    }
    __syncthreads();
    const int total = rows * cols;
    TOut *base = cbase + (int64_t)d.m * col0;
    if (rows == d.m && sizeof(TOut) == 8) {
      // the slab is one contiguous run of output: 16-byte stores of aligned
      // pairs, with a scalar head and tail
      const int head = (reinterpret_cast<uintptr_t>(base) & 8) ? 1 : 0;
      const int npair = (total - head) >> 1;
      if (tid == 0 && head && total > 0) base[0] = stage[0];
      for (int q = tid; q < npair; q += kThreads) {
        const int e = head + 2 * q;
        const int c0 = e / rows, r0 = e - c0 * rows;
        const int c1 = (r0 + 1 == rows) ? c0 + 1 : c0;
        const int r1 = (r0 + 1 == rows) ? 0 : r0 + 1;
        typedef TOut pair_t __attribute__((ext_vector_type(2)));
        pair_t v;
        v.x = stage[c0 * (BM + 1) + r0];
        v.y = stage[c1 * (BM + 1) + r1];
        *reinterpret_cast<pair_t *>(base + e) = v;
      }
      if (tid == 0 && ((total - head) & 1)) {
        const int e = total - 1;
        const int cc = e / rows, rr = e - cc * rows;
        base[e] = stage[cc * (BM + 1) + rr];
      }
    } else if (rows == d.m) {
      for (int e = tid; e < total; e += kThreads) {
        const int cc = e / rows, rr = e - cc * rows;
        base[e] = stage[cc * (BM + 1) + rr];
      }
    } else {
      for (int e = tid; e < total; e += kThreads) {
        const int cc = e / rows, rr = e - cc * rows;
        base[rr + (int64_t)d.m * cc] = stage[cc * (BM + 1) + rr];
      }
    }
    __syncthreads();
  }
}

}  // namespace tqg
