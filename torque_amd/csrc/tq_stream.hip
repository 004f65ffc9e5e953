// tq_stream.hip -- the small-K path of the grouped contraction.
//
// A pair with a short contracted extent (K <= kStreamMaxK) does 2K flops per
// output element: its cost is writing C, not multiplying. Such pairs skip the
// MFMA tiles and are written as a stream: workgroup chunk c covers
// kStreamChunk consecutive elements of one pair's column-major output block,
// lane-contiguous like a fill, each element the FP64 (or FP32) fused
// multiply-add chain sum_k A[m,k] * B[n,k] over the pair's operand views
// (read through L2; the operand panels are small next to the output).
#include "tq_common.h"

namespace {

constexpr int kThreads = 256;
constexpr int kPer = kStreamChunk / kThreads;

template <class TIn, class TC, class TOut>
__global__ __launch_bounds__(kThreads) void small_k_kernel(
    const GemmDesc *__restrict__ descs, const int32_t *__restrict__ chunk_pair,
    const int32_t *__restrict__ chunk_first, int64_t n_chunks,
    const TIn *__restrict__ a, const TIn *__restrict__ a_packed,
    const TIn *__restrict__ b, const TIn *__restrict__ b_packed,
    TOut *__restrict__ c) {
  for (int64_t ch = blockIdx.x; ch < n_chunks; ch += gridDim.x) {
    const int p = chunk_pair[ch];
    const GemmDesc d = descs[p];
    const int64_t e0 = (int64_t)(ch - chunk_first[p]) * kStreamChunk;
    const int64_t mn = (int64_t)d.m * d.n;
    const TIn *pa = ((d.flags & kAPacked) ? a_packed : a) + d.a_off;
    const TIn *pb = ((d.flags & kBPacked) ? b_packed : b) + d.b_off;
    // element strides of m and k in A, of n and k in B
    const int64_t sam = (d.flags & kAKFast) ? d.lda : 1;
    const int64_t sak = (d.flags & kAKFast) ? 1 : d.lda;
    const int64_t sbn = (d.flags & kBKFast) ? d.ldb : 1;
    const int64_t sbk = (d.flags & kBKFast) ? 1 : d.ldb;
    TOut *pc = c + d.c_off;
    // first element of this thread, then step kThreads elements at a time
    int64_t e = e0 + threadIdx.x;
    int64_t n = e / d.m;
    int64_t m = e - n * d.m;
#pragma unroll 4
    for (int q = 0; q < kPer; ++q) {
      if (e >= mn) break;
      const TIn *ra = pa + m * sam;
      const TIn *rb = pb + n * sbn;
      TC acc = (TC)ra[0] * (TC)rb[0];
      for (int k = 1; k < d.k; ++k) acc = fma((TC)ra[k * sak], (TC)rb[k * sbk], acc);
      pc[e] = (TOut)acc;
      e += kThreads;
      m += kThreads;
      if (m >= d.m) {
        const uint32_t wrap = (uint32_t)m / (uint32_t)d.m;
        n += wrap;
        m -= (int64_t)wrap * d.m;
      }
    }
  }
}

template <class TIn, class TC, class TOut>
tq_status launch(tq_handle_t h, const GemmDesc *descs, const int32_t *chunk_pair,
                 const int32_t *chunk_first, int64_t n_chunks, const void *a,
                 const void *ap, const void *b, const void *bp, void *c) {
  static int per_cu = 0;
  if (per_cu == 0) {
    int occ = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &occ, small_k_kernel<TIn, TC, TOut>, kThreads, 0) != hipSuccess || occ < 1)
      occ = 1;
    per_cu = occ;
  }
  const int64_t grid = std::max<int64_t>(1, std::min<int64_t>(n_chunks, (int64_t)h->n_cu * per_cu));
  hipLaunchKernelGGL((small_k_kernel<TIn, TC, TOut>), dim3((unsigned)grid), dim3(kThreads),
                     0, h->stream, descs, chunk_pair, chunk_first, n_chunks,
                     reinterpret_cast<const TIn *>(a), reinterpret_cast<const TIn *>(ap),
                     reinterpret_cast<const TIn *>(b), reinterpret_cast<const TIn *>(bp),
                     reinterpret_cast<TOut *>(c));
  TQ_HIP(h, hipGetLastError());
  return TQ_SUCCESS;
}

}  // namespace

tq_status tq_launch_small_k(tq_handle_t h, tq_dtype dt, tq_precision prec,
                            const GemmDesc *descs, const int32_t *chunk_pair,
                            const int32_t *chunk_first, int64_t n_chunks, const void *a,
                            const void *a_packed, const void *b, const void *b_packed,
                            void *c) {
  if (n_chunks <= 0) return TQ_SUCCESS;
  if (dt == TQ_F32)
    return launch<float, float, float>(h, descs, chunk_pair, chunk_first, n_chunks, a,
                                       a_packed, b, b_packed, c);
  if (prec == TQ_PREC_F32)
    return launch<double, float, double>(h, descs, chunk_pair, chunk_first, n_chunks, a,
                                         a_packed, b, b_packed, c);
  return launch<double, double, double>(h, descs, chunk_pair, chunk_first, n_chunks, a,
                                        a_packed, b, b_packed, c);
}
