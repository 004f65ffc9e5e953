// tq_scan.hip -- device exclusive prefix sums over int64 values: the
// exclusive scans the reference takes on the host with util::nest_sum
// (src/util/space.h:76-88) and arma::cumsum (src/gpu/block_sparse.cuh:
// 662-671). Reduce-then-scan over 2048-element tiles, recursing on the tile
// sums; wavefront scans use DPP-free shuffles of 64-bit values.
#include "tq_common.h"

namespace {

constexpr int kThreads = 256;
constexpr int kItems = 8;
constexpr int kTile = kThreads * kItems;  // 2048

__device__ inline i64 wave_inclusive_scan(i64 v) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    i64 o = __shfl_up(v, d, 64);
    if (lane >= d) v += o;
  }
  return v;
}

// Exclusive scan of one value per thread across the workgroup; returns the
// thread's exclusive prefix and writes the workgroup total to *total.
__device__ inline i64 block_exclusive_scan(i64 v, i64 *total) {
  __shared__ i64 wave_sums[kThreads / 64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  i64 inc = wave_inclusive_scan(v);
  if (lane == 63) wave_sums[wave] = inc;
  __syncthreads();
  i64 wave_off = 0, tot = 0;
#pragma unroll
  for (int w = 0; w < kThreads / 64; ++w) {
    if (w < wave) wave_off += wave_sums[w];
    tot += wave_sums[w];
  }
  __syncthreads();
  *total = tot;
  return wave_off + inc - v;
}

__global__ __launch_bounds__(kThreads) void tile_reduce_kernel(const i64 *in,
                                                               int64_t n,
                                                               i64 *sums) {
  const int64_t base = (int64_t)blockIdx.x * kTile;
  i64 s = 0;
#pragma unroll
  for (int q = 0; q < kItems; ++q) {
    int64_t idx = base + (int64_t)q * kThreads + threadIdx.x;
    if (idx < n) s += in[idx];
  }
  i64 tot;
  (void)block_exclusive_scan(s, &tot);
  if (threadIdx.x == 0) sums[blockIdx.x] = tot;
}

// Each thread owns kItems consecutive elements of the tile.
__global__ __launch_bounds__(kThreads) void tile_scan_kernel(
    const i64 *in, i64 *out, int64_t n, const i64 *tile_offsets,
    i64 *total_dev) {
  const int64_t base = (int64_t)blockIdx.x * kTile + threadIdx.x * kItems;
  i64 v[kItems];
  i64 s = 0;
#pragma unroll
  for (int q = 0; q < kItems; ++q) {
    v[q] = (base + q < n) ? in[base + q] : 0;
    s += v[q];
  }
  i64 tot;
  i64 run = block_exclusive_scan(s, &tot);
  if (tile_offsets) run += tile_offsets[blockIdx.x];
#pragma unroll
  for (int q = 0; q < kItems; ++q) {
    if (base + q < n) out[base + q] = run;
    run += v[q];
  }
  if (total_dev && !tile_offsets && threadIdx.x == 0) *total_dev = tot;
}

}  // namespace

size_t tq_scan_tmp_bytes(int64_t n) {
  size_t b = 0;
  while (n > kTile) {
    n = ceil_div(n, kTile);
    b += (size_t)n * sizeof(i64);
  }
  return b + 8;
}

tq_status tq_scan_i64(tq_handle_t h, const i64 *in, i64 *out, int64_t n,
                      i64 *total_dev, void *tmp) {
  if (n <= 0) {
    if (total_dev) TQ_HIP(h, hipMemsetAsync(total_dev, 0, sizeof(i64), h->stream));
    return TQ_SUCCESS;
  }
  if (n <= kTile) {
    hipLaunchKernelGGL(tile_scan_kernel, dim3(1), dim3(kThreads), 0, h->stream,
                       in, out, n, nullptr, total_dev);
    TQ_HIP(h, hipGetLastError());
    return TQ_SUCCESS;
  }
  const int64_t nt = ceil_div(n, kTile);
  i64 *sums = reinterpret_cast<i64 *>(tmp);
  hipLaunchKernelGGL(tile_reduce_kernel, dim3((unsigned)nt), dim3(kThreads), 0,
                     h->stream, in, n, sums);
  TQ_HIP(h, hipGetLastError());
  // scan the tile sums in place (recursively), which also yields the total
  TQ_CHECK(tq_scan_i64(h, sums, sums, nt, total_dev, sums + nt));
  hipLaunchKernelGGL(tile_scan_kernel, dim3((unsigned)nt), dim3(kThreads), 0,
                     h->stream, in, out, n, sums, nullptr);
  TQ_HIP(h, hipGetLastError());
  return TQ_SUCCESS;
}
