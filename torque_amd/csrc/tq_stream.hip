// tq_stream.hip -- the small-K path of the grouped contraction.
//
// A pair with a short contracted extent (K <= kStreamMaxK) does 2K flops per
// output element: its cost is writing C, not multiplying, so it skips the
// MFMA tiles. One wavefront takes a 64-row segment of the pair's column-major
// output: each lane holds its row of A (K values) in registers and walks the
// N columns; B[n, :] is the same for the whole wavefront (scalar loads), and
// each column is one contiguous 64-element store. The next segment's
// descriptor and A rows are loaded while the current one is written.
#include "tq_common.h"

namespace {

constexpr int kThreads = 256;
constexpr int kWaves = kThreads / 64;

// Operands of one 64-row segment: the lane's row of A and the column walk.
template <class TIn, class TC>
struct Segment {
  GemmDesc d;
  int row;
  bool live;
  const TIn *pb;
  int64_t sbn, sbk;
  TC av[kStreamMaxK];
};

template <class TIn, class TC>
__device__ inline void load_segment(Segment<TIn, TC> &s, int sg,
                                    const GemmDesc *__restrict__ descs,
                                    const int32_t *__restrict__ seg_pair,
                                    const int32_t *__restrict__ seg_first,
                                    const TIn *a, const TIn *a_packed, const TIn *b,
                                    const TIn *b_packed, int lane) {
  const int p = __builtin_amdgcn_readfirstlane(seg_pair[sg]);
  s.d = descs[p];
  s.row = (sg - seg_first[p]) * kStreamRows + lane;
  s.live = s.row < s.d.m;
  const TIn *pa = ((s.d.flags & kAPacked) ? a_packed : a) + s.d.a_off;
  s.pb = ((s.d.flags & kBPacked) ? b_packed : b) + s.d.b_off;
  const int64_t sam = (s.d.flags & kAKFast) ? s.d.lda : 1;
  const int64_t sak = (s.d.flags & kAKFast) ? 1 : s.d.lda;
  s.sbn = (s.d.flags & kBKFast) ? s.d.ldb : 1;
  s.sbk = (s.d.flags & kBKFast) ? 1 : s.d.ldb;
#pragma unroll
  for (int k = 0; k < kStreamMaxK; ++k)
    s.av[k] = (s.live && k < s.d.k) ? (TC)pa[s.row * sam + k * sak] : (TC)0;
}

template <class TIn, class TC, class TOut>
__global__ __launch_bounds__(kThreads) void small_k_kernel(
    const GemmDesc *__restrict__ descs, const int32_t *__restrict__ seg_pair,
    const int32_t *__restrict__ seg_first, int64_t n_segs,
    const TIn *__restrict__ a, const TIn *__restrict__ a_packed,
    const TIn *__restrict__ b, const TIn *__restrict__ b_packed,
    TOut *__restrict__ c) {
  const int lane = threadIdx.x & 63;
  // the segment index is the same on every lane of the wavefront: say so, so
  // that the descriptor and B reads become scalar loads
  const int wave_id = __builtin_amdgcn_readfirstlane(
      (int)(blockIdx.x * kWaves + (threadIdx.x >> 6)));
  const int n_waves = (int)(gridDim.x * kWaves);
  int sg = wave_id;
  if (sg >= n_segs) return;
  Segment<TIn, TC> cur, nxt;
  load_segment(cur, sg, descs, seg_pair, seg_first, a, a_packed, b, b_packed, lane);
  while (true) {
    const int sg_next = sg + n_waves;
    // the next segment's loads overlap this segment's stores
    if (sg_next < n_segs)
      load_segment(nxt, sg_next, descs, seg_pair, seg_first, a, a_packed, b, b_packed,
                   lane);
    TOut *col = c + cur.d.c_off + cur.row;
    for (int n = 0; n < cur.d.n; ++n) {
      const TIn *rb = cur.pb + n * cur.sbn;
      TC acc = cur.av[0] * (TC)rb[0];
#pragma unroll
      for (int k = 1; k < kStreamMaxK; ++k)
        if (k < cur.d.k) acc = fma(cur.av[k], (TC)rb[k * cur.sbk], acc);
      if (cur.live) col[(int64_t)cur.d.m * n] = (TOut)acc;
    }
    if (sg_next >= n_segs) break;
    sg = sg_next;
    cur = nxt;
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
  const int64_t grid = std::max<int64_t>(
      1, std::min<int64_t>(ceil_div(n_chunks, kWaves), (int64_t)h->n_cu * per_cu));
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
