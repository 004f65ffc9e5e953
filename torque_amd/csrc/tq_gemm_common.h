// tq_gemm_common.h -- pieces shared by the grouped-GEMM kernels: the
// XCD-aware walk of the tile list and the LDS-staged epilogue.
#pragma once

#include "tq_common.h"

namespace tqg {

constexpr int kThreads = 256;

// Workgroups are dealt to the 8 XCDs round-robin; this virtual index makes
// each XCD walk a contiguous stretch of the tile list (tiles of one pair share
// operand panels, so they should share an L2).
// Tile handed out by per-XCD counters: the tile list is dealt to the 8 XCDs
// in stripes of kTileStripe consecutive tiles (stripe s to XCD s % 8, a
// workgroup's XCD taken as blockIdx.x % 8 as the dispatcher deals them), and
// value c of XCD part's counter is tile (c / kTileStripe) * 8 kTileStripe +
// part * kTileStripe + c % kTileStripe -- increasing in c, so a draw past the
// list means the XCD's tiles are all taken.
constexpr int kTileStripe = 64;
__device__ inline int64_t stripe_tile(uint32_t c, int part) {
  return (int64_t)(c / kTileStripe) * (kTileStripe * kTileCounters) +
         (int64_t)part * kTileStripe + (int64_t)(c % kTileStripe);
}

// Clears the kTileCounters counters before a launch that draws from them (a
// kernel, so that a captured graph clears them on every replay; one copy per
// translation unit).
static __global__ void zero_tile_counters_kernel(uint32_t *__restrict__ ctr) {
  if (threadIdx.x < kTileCounters) ctr[threadIdx.x] = 0;
}

__device__ inline int64_t xcd_virtual_id(int64_t bid, int64_t grid) {
  const int64_t per = grid / 8;
  if (per == 0 || grid % 8 != 0) return bid;
  return (bid % 8) * per + bid / 8;
}

// Position (tile row, tile column) of tile lt of a pair with tiles_m x
// tiles_n tiles: bands of `group` tile rows are walked column by column, so a
// run of consecutive tiles (what one XCD holds at once) covers a block of the
// output and reuses both operand panels from L2, where a plain m-fastest walk
// reuses only the B panel.
struct TilePos {
  int mt, nt;
};
__device__ inline TilePos tile_position(int lt, int tiles_m, int tiles_n, int group) {
  const int per_group = group * tiles_n;
  const int g = lt / per_group;
  const int first = g * group;
  const int gs = min(tiles_m - first, group);
  const int r = lt - g * per_group;
  return TilePos{first + r % gs, r / gs};
}

// Band height for tile_position: the workgroups of one XCD (a persistent
// grid's eighth) hold `per_xcd` BM x BN tiles at once; a band of
// sqrt(per_xcd * BN / BM) tile rows makes that block of the output square.
template <int BM, int BN>
__device__ inline int tile_group(int64_t grid) {
  const float per_xcd = (float)max<int64_t>(grid / 8, 1);
  const int g = (int)(sqrtf(per_xcd * BN / BM) + 0.5f);
  return min(max(g, 1), 32);
}

// Workgroup barrier ordering LDS accesses only: the fences are restricted to
// the local address space, so global loads in flight stay in flight across it
// (a full __syncthreads() waits for every outstanding global access).
__device__ inline void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
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
// NW: wavefronts in the workgroup (WAVES_M along M, NW / WAVES_M along N).
template <class TOut, bool kF64Layout, int BM, int BN, int WAVES_M, int TM, int TN,
          int NW = 4, class AccT>
__device__ inline void store_tile(const AccT (&acc)[TM][TN], TOut *__restrict__ c,
                                  const GemmDesc &d, int m0, int n0, TOut *stage) {
  constexpr int kThreads = NW * 64;
  constexpr int WM = BM / WAVES_M;
  constexpr int WN = BN / (NW / WAVES_M);
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
