// tq_plan.hip -- block-sparse contraction planning on the device.
//
// Restates, as device kernels, the host metadata work of the reference's GPU
// contract (src/gpu/block_sparse.cuh:570-732) and of
// block_sparse::block_in_range (src/tensor/block_sparse.cpp:6-175):
//   1. match:   one wavefront per (A block, slice of B blocks); lane j tests
//               B block j for a non-empty intersection on every contracted
//               leg pair (max(begin) <= min(end), block_sparse.cpp:42-60) and
//               a ballot packs the matches. Counting pass, exclusive scan,
//               then a fill pass that writes pairs at scanned positions with
//               the ballot's lane prefix, so pairs come out ordered by A block
//               then ascending B block (block_sparse.cuh:626-652).
//   2. geometry: per pair, output begin/end (A free ++ B free,
//               block_sparse.cpp:66-119), GEMM shape M, N, K, the A and B
//               sub-block offsets (block_sparse.cuh:700-732), whether each
//               operand can be read in place, and the output block offsets
//               (exclusive scan of M*N, block_sparse.cuh:667-671).
//   3. pack jobs for operands that must be permuted to GEMM layout.
//   4. work list: a stable device counting sort of the pairs by estimated
//               cost (largest first) using wavefront ballots to rank equal
//               keys, then the tile list of the grouped GEMM.
#include <algorithm>
#include <cstring>
#include <map>
#include <memory>

#include "tq_common.h"
#include "tq_plan.h"

namespace {

constexpr int kWave = 64;

struct MatchArgs {
  int32_t nl;
  int32_t a_rank;
  int64_t n_a, n_b;
  int32_t chunks_per_split;  // 64-block chunks each wavefront scans
  int64_t n_split;
  const i64 *a_begin, *a_end;      // [n_a * a_rank]
  const int32_t *b_lo, *b_hi;      // [nl * n_b]  contracted legs, SoA
  int32_t a_legs[TQ_MAX_RANK];
};

__device__ inline uint64_t lanemask_lt() {
  const int lane = threadIdx.x & 63;
  return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}

// Predicate of block_in_range for (A block i, B block j).
__device__ inline bool block_match(const MatchArgs &g, const i64 *alo,
                                   const i64 *ahi, int64_t j) {
  bool ok = true;
  for (int k = 0; k < g.nl; ++k) {
    const i64 lo = max(alo[k], (i64)g.b_lo[(int64_t)k * g.n_b + j]);
    const i64 hi = min(ahi[k], (i64)g.b_hi[(int64_t)k * g.n_b + j]);
    ok = ok && (lo <= hi);
  }
  return ok;
}

// kFill = false: counts[item] = number of matches of the item's B slice.
// kFill = true:  writes pair_a/pair_b at starts[item] + rank.
template <bool kFill>
__global__ __launch_bounds__(256) void match_kernel(
    MatchArgs g, i64 *__restrict__ counts, const i64 *__restrict__ starts,
    int32_t *__restrict__ pair_a, int32_t *__restrict__ pair_b,
    uint8_t *__restrict__ a_used, uint8_t *__restrict__ b_used) {
  const int64_t item = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (item >= g.n_a * g.n_split) return;  // whole wavefront exits together
  const int64_t i = item / g.n_split;
  const int64_t s = item - i * g.n_split;
  const int lane = threadIdx.x & 63;
  i64 alo[TQ_MAX_RANK], ahi[TQ_MAX_RANK];
  for (int k = 0; k < g.nl; ++k) {
    alo[k] = g.a_begin[i * g.a_rank + g.a_legs[k]];
    ahi[k] = g.a_end[i * g.a_rank + g.a_legs[k]];
  }
  const int64_t j0 = s * (int64_t)g.chunks_per_split * kWave;
  int64_t pos = kFill ? starts[item] : 0;
  i64 cnt = 0;
  for (int c = 0; c < g.chunks_per_split; ++c) {
    const int64_t j = j0 + (int64_t)c * kWave + lane;
    const bool m = (j < g.n_b) && block_match(g, alo, ahi, j);
    const uint64_t mask = __ballot(m);
    if (kFill) {
      if (m) {
        const int64_t p = pos + __popcll(mask & lanemask_lt());
        pair_a[p] = (int32_t)i;
        pair_b[p] = (int32_t)j;
        b_used[j] = 1;
      }
      if (mask && lane == 0) a_used[i] = 1;
      pos += __popcll(mask);
    } else {
      cnt += __popcll(mask);
    }
  }
  if (!kFill && lane == 0) counts[item] = cnt;
}

__global__ void soa_legs_kernel(int64_t n_b, int32_t b_rank, int32_t nl,
                                const i64 *__restrict__ b_begin,
                                const i64 *__restrict__ b_end, LegList legs,
                                int32_t *__restrict__ lo,
                                int32_t *__restrict__ hi) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n_b) return;
  for (int k = 0; k < nl; ++k) {
    lo[(int64_t)k * n_b + j] = (int32_t)b_begin[j * b_rank + legs.b[k]];
    hi[(int64_t)k * n_b + j] = (int32_t)b_end[j * b_rank + legs.b[k]];
  }
}

// Layout test of one operand view: legs listed [free..., contracted...] with
// view dims vd and strides vs; returns the leading dimension if the view is
// a plain matrix with the free index fastest (kfast = false) or the
// contracted index fastest (kfast = true), or -1.
__device__ inline i64 plain_ld(int nf, int nl, const i64 *vd, const i64 *vs,
                               bool kfast, i64 rows, i64 cols) {
  // fast group: must be a unit-stride chain; slow group: chain from ld
  const int f0 = kfast ? nf : 0, f1 = kfast ? nf + nl : nf;
  const int s0 = kfast ? 0 : nf, s1 = kfast ? nf : nf + nl;
  i64 expect = 1;
  for (int q = f0; q < f1; ++q) {
    if (vd[q] == 1) continue;
    if (vs[q] != expect) return -1;
    expect *= vd[q];
  }
  i64 ld = -1;
  for (int q = s0; q < s1; ++q) {
    if (vd[q] == 1) continue;
    if (ld < 0) {
      ld = vs[q];
      expect = vs[q] * vd[q];
    } else {
      if (vs[q] != expect) return -1;
      expect *= vd[q];
    }
  }
  const i64 fast_extent = kfast ? cols : rows;
  if (ld < 0) ld = fast_extent > 0 ? fast_extent : 1;
  if (ld < fast_extent && (kfast ? rows : cols) > 1) return -1;
  return ld;
}

__global__ void geometry_kernel(GeomArgs g, const int32_t *__restrict__ pair_a,
                                const int32_t *__restrict__ pair_b,
                                GemmDesc *__restrict__ desc,
                                i64 *__restrict__ sub_off,
                                i64 *__restrict__ out_begin,
                                i64 *__restrict__ out_end,
                                i64 *__restrict__ mn, i64 *__restrict__ cost,
                                int32_t *__restrict__ flags_any) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= g.n_pairs) return;
  const int64_t i = pair_a[p], j = pair_b[p];
  const int ra = g.a_rank, rb = g.b_rank, nl = g.nl;
  const i64 *ab = g.a_begin + i * ra, *ae = g.a_end + i * ra,
            *as = g.a_str + i * ra;
  const i64 *bb = g.b_begin + j * rb, *be = g.b_end + j * rb,
            *bs = g.b_str + j * rb;
  i64 avd[TQ_MAX_RANK], avs[TQ_MAX_RANK], bvd[TQ_MAX_RANK], bvs[TQ_MAX_RANK];
  i64 a_off = g.a_off[i], b_off = g.b_off[j];
  i64 M = 1, N = 1, K = 1;
  const int naf = ra - nl, nbf = rb - nl;
  const int ro = naf + nbf;
  for (int q = 0; q < naf; ++q) {
    const int l = g.a_free[q];
    avd[q] = ae[l] - ab[l] + 1;
    avs[q] = as[l];
    M *= avd[q];
    out_begin[p * ro + q] = ab[l];
    out_end[p * ro + q] = ae[l];
  }
  for (int q = 0; q < nbf; ++q) {
    const int l = g.b_free[q];
    bvd[q] = be[l] - bb[l] + 1;
    bvs[q] = bs[l];
    N *= bvd[q];
    out_begin[p * ro + naf + q] = bb[l];
    out_end[p * ro + naf + q] = be[l];
  }
  bool a_full = true, b_full = true;
  bool a_box = true, b_box = true;  // sub-box of the block panel is plain
  int last_nonunit = -1;
  for (int k = 0; k < nl; ++k) {
    const int la = g.legs.a[k], lb = g.legs.b[k];
    const i64 lo = max(ab[la], bb[lb]);
    const i64 hi = min(ae[la], be[lb]);
    const i64 cd = hi - lo + 1;
    a_off += (lo - ab[la]) * as[la];
    b_off += (lo - bb[lb]) * bs[lb];
    avd[naf + k] = cd;
    avs[naf + k] = as[la];
    bvd[nbf + k] = cd;
    bvs[nbf + k] = bs[lb];
    K *= cd;
    a_full = a_full && (cd == ae[la] - ab[la] + 1);
    b_full = b_full && (cd == be[lb] - bb[lb] + 1);
    if (cd != 1) last_nonunit = k;
  }
  // inside a block panel the sub-box is plain only if every contracted leg
  // before the last non-unit one spans its whole block extent
  for (int k = 0; k < last_nonunit; ++k) {
    const int la = g.legs.a[k], lb = g.legs.b[k];
    const i64 cd = avd[naf + k];
    a_box = a_box && (cd == ae[la] - ab[la] + 1);
    b_box = b_box && (cd == be[lb] - bb[lb] + 1);
  }
  sub_off[2 * p] = a_off;
  sub_off[2 * p + 1] = b_off;
  GemmDesc d;
  d.a_off = a_off;
  d.b_off = b_off;
  d.c_off = 0;
  d.m = (int32_t)M;
  d.n = (int32_t)N;
  d.k = (int32_t)K;
  d.flags = 0;
  int32_t need = 0;
  i64 lda = plain_ld(naf, nl, avd, avs, false, M, K);
  if (lda < 0) {
    lda = plain_ld(naf, nl, avd, avs, true, M, K);
    if (lda >= 0)
      d.flags |= kAKFast;
    else
      need |= kNeedPackA;
  }
  i64 ldb = plain_ld(nbf, nl, bvd, bvs, false, N, K);
  if (ldb < 0) {
    ldb = plain_ld(nbf, nl, bvd, bvs, true, N, K);
    if (ldb >= 0)
      d.flags |= kBKFast;
    else
      need |= kNeedPackB;
  }
  if (lda > 0x7fffffff || ldb > 0x7fffffff) need |= kNeedPackA | kNeedPackB;
  d.lda = (int32_t)(lda < 0 ? M : lda);
  d.ldb = (int32_t)(ldb < 0 ? N : ldb);
  if (!(a_full || a_box)) need |= kNeedPairPanelA;
  if (!(b_full || b_box)) need |= kNeedPairPanelB;
  if (!a_full) need |= kNotFullA;
  if (!b_full) need |= kNotFullB;
  if (M * K > 0x7fffffff || N * K > 0x7fffffff) need |= kTooLarge;
  desc[p] = d;
  mn[p] = M * N;
  cost[p] = pair_order_key(M, N, K);
  if (need) atomicOr(flags_any, need);
}

__global__ void set_c_offsets_kernel(int64_t n, const i64 *__restrict__ c_off,
                                     GemmDesc *__restrict__ desc) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p < n) desc[p].c_off = c_off[p];
}

__device__ inline i64 round8(i64 x) { return (x + 7) & ~(i64)7; }

// Per-block panel sizes (unused blocks get 0). A plain panel is the block
// permuted to [free..., contracted in pair order...]; a split-BF16 panel has
// rows = the free legs and kpad = kdim rounded up to 8 contracted values per
// row, stored k-chunked (CopyJob).
__global__ void block_panel_sizes_kernel(PanelArgs g,
                                         const uint8_t *__restrict__ used,
                                         i64 *__restrict__ sizes) {
  const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= g.n_blocks) return;
  i64 n = 1, kd = 1;
  for (int q = 0; q < g.rank; ++q) {
    const int l = g.order[q];
    const i64 d = g.end[b * g.rank + l] - g.begin[b * g.rank + l] + 1;
    n *= d;
    if (q < g.nk) kd *= d;
  }
  if (g.split) n = (n / kd) * round8(kd);
  sizes[b] = used[b] ? n : 0;
}

__global__ void block_panel_jobs_kernel(PanelArgs g,
                                        const i64 *__restrict__ panel_off,
                                        const uint8_t *__restrict__ used,
                                        CopyJob *__restrict__ jobs,
                                        i64 *__restrict__ wg) {
  const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= g.n_blocks) return;
  CopyJob j{};
  j.src_off = g.offsets[b];
  j.dst_off = panel_off[b];
  j.rank = g.rank;
  i64 n = 1, kd = 1;
  for (int q = 0; q < g.rank; ++q) {
    const int l = g.order[q];
    const i64 d = g.end[b * g.rank + l] - g.begin[b * g.rank + l] + 1;
    j.dims[q] = (int32_t)d;
    j.strides[q] = g.strides[b * g.rank + l];
    n *= d;
    if (q < g.nk) kd *= d;
  }
  if (g.split) {
    j.kdim = (int32_t)kd;
    j.kpad = (int32_t)round8(kd);
    j.nk = g.nk;
  }
  j.n_elem = used[b] ? n : 0;
  finalize_job(j);
  jobs[b] = j;
  wg[b] = pack_job_wgs(j);
}

// Redirect each pair's operand into its block panel: the pair's sub-box of
// the contracted legs starts at (lo - begin) and, being plain, has a single
// leading dimension.
__global__ void block_panel_desc_kernel(BoxArgs g,
                                        const int32_t *__restrict__ pair_blk,
                                        const int32_t *__restrict__ pair_other,
                                        const i64 *__restrict__ panel_off,
                                        GemmDesc *__restrict__ desc) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= g.n_pairs) return;
  const int64_t i = pair_blk[p], j = pair_other[p];
  const i64 *sb = g.self_begin + i * g.self_rank, *se = g.self_end + i * g.self_rank;
  const i64 *ob = g.other_begin + j * g.other_rank;
  i64 rows = 1, kfull = 1;
  for (int q = 0; q < g.n_free; ++q) {
    const int l = g.free_legs[q];
    rows *= se[l] - sb[l] + 1;
  }
  for (int k = 0; k < g.nl; ++k) kfull *= se[g.self_legs[k]] - sb[g.self_legs[k]] + 1;
  // plain panel: column stride of contracted index starts at rows;
  // split panel: the pair reads the block's whole contracted extent (pairs
  // that do not get pair panels), so its offset stays at the panel start
  // (k-fast plain panel: the contracted index has stride 1, rows kfull)
  i64 off = panel_off[i], cstr = (g.split || g.kfast) ? 1 : rows;
  for (int k = 0; k < g.nl; ++k) {
    const int ls = g.self_legs[k], lo_ = g.other_legs[k];
    const i64 lo = max(sb[ls], ob[lo_]);
    off += (lo - sb[ls]) * cstr;
    cstr *= se[ls] - sb[ls] + 1;
  }
  const i64 ld = g.split ? round8(kfull) : g.kfast ? (kfull > 0 ? kfull : 1)
                                                    : (rows > 0 ? rows : 1);
  const int32_t kf = (g.split || g.kfast) ? (g.is_b ? kBKFast : kAKFast) : 0;
  GemmDesc d = desc[p];
  if (g.is_b) {
    d.b_off = off;
    d.ldb = (int32_t)ld;
    d.flags = (d.flags & ~kBKFast) | kBPacked | kf;
  } else {
    d.a_off = off;
    d.lda = (int32_t)ld;
    d.flags = (d.flags & ~kAKFast) | kAPacked | kf;
  }
  desc[p] = d;
}

// Per-pair panels (general overlapping blocks): panel p holds the pair's
// sub-block permuted to [free..., contracted...], or for split-BF16 the
// k-chunked planes of its rows (free legs) x contracted values.
__global__ void pair_panel_sizes_kernel(int64_t n_pairs, int32_t is_b, int32_t split,
                                        const GemmDesc *__restrict__ desc,
                                        i64 *__restrict__ sizes) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n_pairs) return;
  const GemmDesc d = desc[p];
  const i64 rows = is_b ? d.n : d.m;
  sizes[p] = rows * (split ? round8(d.k) : (i64)d.k);
}

__global__ void pair_panel_jobs_kernel(BoxArgs g,
                                       const int32_t *__restrict__ pair_blk,
                                       const int32_t *__restrict__ pair_other,
                                       const i64 *__restrict__ self_str,
                                       const i64 *__restrict__ self_offsets,
                                       const i64 *__restrict__ panel_off,
                                       const i64 *__restrict__ panel_ld,
                                       GemmDesc *__restrict__ desc,
                                       CopyJob *__restrict__ jobs,
                                       i64 *__restrict__ wg) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= g.n_pairs) return;
  const int64_t i = pair_blk[p], j = pair_other[p];
  const int r = g.self_rank;
  const i64 *sb = g.self_begin + i * r, *se = g.self_end + i * r,
            *ss = self_str + i * r;
  const i64 *ob = g.other_begin + j * g.other_rank,
            *oe = g.other_end + j * g.other_rank;
  CopyJob job{};
  job.rank = g.n_free + g.nl;
  const bool kfirst = g.split || g.kfast;  // contracted legs first
  const int f0 = kfirst ? g.nl : 0;          // position of the first free leg
  const int c0 = kfirst ? 0 : g.n_free;      // position of the first contracted leg
  i64 src = self_offsets[i], n = 1, rows = 1, kk = 1;
  for (int q = 0; q < g.n_free; ++q) {
    const int l = g.free_legs[q];
    const i64 d = se[l] - sb[l] + 1;
    job.dims[f0 + q] = (int32_t)d;
    job.strides[f0 + q] = ss[l];
    n *= d;
    rows *= d;
  }
  for (int k = 0; k < g.nl; ++k) {
    const int ls = g.self_legs[k], lo_ = g.other_legs[k];
    const i64 lo = max(sb[ls], ob[lo_]);
    const i64 hi = min(se[ls], oe[lo_]);
    const i64 d = hi - lo + 1;
    src += (lo - sb[ls]) * ss[ls];
    job.dims[c0 + k] = (int32_t)d;
    job.strides[c0 + k] = ss[ls];
    n *= d;
    kk *= d;
  }
  if (g.split) {
    job.kdim = (int32_t)kk;
    job.kpad = (int32_t)round8(kk);
    job.nk = g.nl;
  }
  job.src_off = src;
  job.dst_off = panel_off[p];
  job.n_elem = n;
  i64 ld = g.split ? round8(kk) : g.kfast ? (kk > 0 ? kk : 1) : (rows > 0 ? rows : 1);
  if (g.kfast && !g.split && panel_ld && panel_ld[p] > kk) {
    // the pair's slice of a k-fast panel with longer rows (a group panel):
    // contracted legs compact, each row panel_ld[p] apart
    job.dst_strided = 1;
    i64 st = 1;
    for (int k = 0; k < g.nl; ++k) {
      job.dst_strides[c0 + k] = st;
      st *= job.dims[c0 + k];
    }
    st = panel_ld[p];
    for (int q = 0; q < g.n_free; ++q) {
      job.dst_strides[f0 + q] = st;
      st *= job.dims[f0 + q];
    }
    ld = panel_ld[p];
  }
  finalize_job(job);
  jobs[p] = job;
  wg[p] = pack_job_wgs(job);
  GemmDesc d = desc[p];
  const int32_t kf = kfirst ? (g.is_b ? kBKFast : kAKFast) : 0;
  if (g.is_b) {
    d.b_off = panel_off[p];
    d.ldb = (int32_t)ld;
    d.flags = (d.flags & ~kBKFast) | kBPacked | kf;
  } else {
    d.a_off = panel_off[p];
    d.lda = (int32_t)ld;
    d.flags = (d.flags & ~kAKFast) | kAPacked | kf;
  }
  desc[p] = d;
}

__global__ void i64_to_i32_kernel(const i64 *__restrict__ in,
                                  int32_t *__restrict__ out, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = (int32_t)in[i];
}

// ---- stable counting sort of pairs by cost (descending) ------------------

constexpr int kSortChunk = 1024;
constexpr int kBuckets = 64;

__device__ inline int cost_bucket(i64 c) {
  // larger cost -> smaller key, so an ascending sort puts big pairs first
  const int lg = c <= 0 ? 0 : 63 - __clzll((unsigned long long)c);
  return 63 - lg;
}

__global__ __launch_bounds__(256) void sort_hist_kernel(
    const i64 *__restrict__ cost, int64_t n, int64_t n_chunks,
    i64 *__restrict__ hist) {
  __shared__ int h[kBuckets];
  if (threadIdx.x < kBuckets) h[threadIdx.x] = 0;
  __syncthreads();
  const int64_t base = (int64_t)blockIdx.x * kSortChunk;
  for (int q = threadIdx.x; q < kSortChunk; q += 256) {
    const int64_t idx = base + q;
    if (idx < n) atomicAdd(&h[cost_bucket(cost[idx])], 1);
  }
  __syncthreads();
  if (threadIdx.x < kBuckets)
    hist[(int64_t)threadIdx.x * n_chunks + blockIdx.x] = h[threadIdx.x];
}

// Stable scatter: items of a chunk are ranked in order; within a round of
// 256 items each wavefront finds lanes with equal keys from six ballots (one
// per key bit) and counts the earlier ones.
__global__ __launch_bounds__(256) void sort_scatter_kernel(
    const i64 *__restrict__ cost, int64_t n, int64_t n_chunks,
    const i64 *__restrict__ base_off, i64 *__restrict__ order) {
  __shared__ i64 running[kBuckets];
  __shared__ int wave_cnt[4][kBuckets];
  const int wave = threadIdx.x >> 6;
  if (threadIdx.x < kBuckets)
    running[threadIdx.x] = base_off[(int64_t)threadIdx.x * n_chunks + blockIdx.x];
  const int64_t base = (int64_t)blockIdx.x * kSortChunk;
  for (int round = 0; round < kSortChunk / 256; ++round) {
    for (int q = threadIdx.x; q < 4 * kBuckets; q += 256)
      (&wave_cnt[0][0])[q] = 0;
    __syncthreads();
    const int64_t idx = base + round * 256 + threadIdx.x;
    const bool valid = idx < n;
    const int key = valid ? cost_bucket(cost[idx]) : kBuckets;  // invalid: 64
    uint64_t peers = __ballot(valid);
#pragma unroll
    for (int bit = 0; bit < 7; ++bit) {
      const uint64_t bm = __ballot((key >> bit) & 1);
      peers &= ((key >> bit) & 1) ? bm : ~bm;
    }
    const int rank_in_wave = __popcll(peers & lanemask_lt());
    const int wave_count = __popcll(peers);
    if (valid && rank_in_wave == wave_count - 1) wave_cnt[wave][key] = wave_count;
    __syncthreads();
    if (valid) {
      i64 off = running[key] + rank_in_wave;
      for (int w = 0; w < wave; ++w) off += wave_cnt[w][key];
      order[off] = idx;
    }
    __syncthreads();
    if (threadIdx.x < kBuckets) {
      int t = 0;
      for (int w = 0; w < 4; ++w) t += wave_cnt[w][threadIdx.x];
      running[threadIdx.x] += t;
    }
    __syncthreads();
  }
}

// Work items per position of the work list: streaming row segments of
// pairs with K <= max_k (cls = -1; bm = segment height), or MFMA tiles of the
// pairs in tile class cls (0: min(M, N) < large_min, 1: otherwise).
__global__ void tiles_per_position_kernel(int64_t n, const i64 *__restrict__ order,
                                          const GemmDesc *__restrict__ desc,
                                          int32_t bm, int32_t bn, int32_t max_k,
                                          int32_t large_min, int32_t cls,
                                          i64 *__restrict__ tiles) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  const GemmDesc d = desc[order[q]];
  const int c = d.k <= max_k ? -1 : (d.m >= large_min && d.n >= large_min) ? 1 : 0;
  tiles[q] = c != cls ? 0 : ceil_div(d.m, bm) * ceil_div(d.n, bn);
}

__global__ void expand_tiles_kernel(int64_t n, const i64 *__restrict__ order,
                                    const i64 *__restrict__ tile_start,
                                    int32_t *__restrict__ tile_pair,
                                    int32_t *__restrict__ tile_first) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  const int64_t p = order[q];
  const i64 t0 = tile_start[q], t1 = tile_start[q + 1];
  tile_first[p] = (int32_t)t0;
  for (i64 t = t0; t < t1; ++t) tile_pair[t] = (int32_t)p;
}

// Streaming small-K work (pairs with K <= max_k, or every unit of a plan
// whose groups stream): column slices of kStreamCols per pair at work-list
// position q, each cut into segments of kStreamRows rows.
__global__ void stream_slices_kernel(int64_t n, const i64 *__restrict__ order,
                                     const GemmDesc *__restrict__ desc, int32_t max_k,
                                     i64 *__restrict__ slices) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  const GemmDesc d = desc[order[q]];
  slices[q] = d.k <= max_k ? ceil_div(d.n, kStreamCols) : 0;
}

// The slices of the pair at position q (B and C offsets moved to the slice's
// first column; for streamed groups the group's member range instead) and
// each slice's segment count.
__global__ void stream_slice_desc_kernel(int64_t n, const i64 *__restrict__ order,
                                         const GemmDesc *__restrict__ desc,
                                         const i64 *__restrict__ slice_start, int32_t split,
                                         const i64 *__restrict__ member_first,
                                         GemmDesc *__restrict__ slices,
                                         GroupSlice *__restrict__ gslices,
                                         i64 *__restrict__ segs) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  const int64_t p = order[q];
  const GemmDesc d = desc[p];
  const i64 s0 = slice_start[q], s1 = slice_start[q + 1];
  const i64 row_segs = ceil_div(d.m, kStreamRows);
  // element distance between consecutive B rows (output columns): a split
  // panel with one chunk holds 8 values per row
  const i64 b_row = split ? 8 : (d.flags & kBKFast) ? d.ldb : 1;
  for (i64 s = s0; s < s1; ++s) {
    const int32_t n0 = (int32_t)(s - s0) * kStreamCols;
    const int32_t nc = min(kStreamCols, d.n - n0);
    if (gslices) {
      gslices[s] = GroupSlice{d.c_off + (i64)d.m * n0, d.m, nc, n0, (int32_t)member_first[p],
                              (int32_t)(member_first[p + 1] - member_first[p]), 0};
    } else {
      GemmDesc sd = d;
      sd.b_off += n0 * b_row;
      sd.c_off += (i64)d.m * n0;
      sd.n = nc;
      slices[s] = sd;
    }
    segs[s] = row_segs;
  }
}

// Segment t of slice s for seg_start[s] <= t < seg_start[s + 1].
__global__ void stream_segments_kernel(int64_t n_slices, const i64 *__restrict__ seg_start,
                                       int32_t *__restrict__ seg_slice,
                                       int32_t *__restrict__ slice_first) {
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n_slices) return;
  const i64 t0 = seg_start[s], t1 = seg_start[s + 1];
  slice_first[s] = (int32_t)t0;
  for (i64 t = t0; t < t1; ++t) seg_slice[t] = (int32_t)s;
}

unsigned grid_for(int64_t n, int threads = 256) {
  return (unsigned)std::max<int64_t>(1, ceil_div(n, threads));
}

bool valid_legs(int ra, int rb, int nl, const int32_t *legs) {
  bool sa[TQ_MAX_RANK] = {false}, sb[TQ_MAX_RANK] = {false};
  for (int k = 0; k < nl; ++k) {
    const int a = legs[2 * k], b = legs[2 * k + 1];
    if (a < 0 || a >= ra || b < 0 || b >= rb || sa[a] || sb[b]) return false;
    sa[a] = sb[b] = true;
  }
  return true;
}

bool valid_desc(const tq_bs_desc *d) {
  if (!d || d->rank < 0 || d->rank > TQ_MAX_RANK || d->n_blocks < 0) return false;
  if (d->rank > 0 && !d->dims) return false;
  if (d->n_blocks > 0 && d->rank > 0 &&
      (!d->begin || !d->end || !d->strides || !d->offsets))
    return false;
  if (d->n_blocks > 0x7fffffff) return false;
  for (int l = 0; l < d->rank; ++l)
    if (d->dims[l] < 0 || d->dims[l] > 0x7fffffff) return false;
  for (int64_t b = 0; b < d->n_blocks; ++b)
    for (int l = 0; l < d->rank; ++l) {
      const i64 lo = d->begin[b * d->rank + l], hi = d->end[b * d->rank + l];
      if (lo < 0 || hi < lo || hi >= d->dims[l]) return false;
    }
  return true;
}

}  // namespace

// Stable device counting sort of n items by cost bucket, largest cost first:
// order receives item indices.
tq_status tq_sort_by_cost(tq_handle_t h, const i64 *cost, int64_t n, i64 *order) {
  if (n <= 0) return TQ_SUCCESS;
  const int64_t n_chunks = ceil_div(n, kSortChunk);
  DevBuf hist, stmp;
  TQ_HIP(h, hist.alloc(sizeof(i64) * (kBuckets * n_chunks + 1)));
  TQ_HIP(h, stmp.alloc(tq_scan_tmp_bytes(kBuckets * n_chunks + 1)));
  hipLaunchKernelGGL(sort_hist_kernel, dim3((unsigned)n_chunks), dim3(256), 0, h->stream,
                     cost, n, n_chunks, hist.as<i64>());
  TQ_CHECK(tq_scan_i64(h, hist.as<i64>(), hist.as<i64>(), kBuckets * n_chunks, nullptr,
                       stmp.p));
  hipLaunchKernelGGL(sort_scatter_kernel, dim3((unsigned)n_chunks), dim3(256), 0, h->stream,
                     cost, n, n_chunks, hist.as<i64>(), order);
  TQ_HIP(h, hipGetLastError());
  // the temporaries are released on return
  TQ_HIP(h, hipStreamSynchronize(h->stream));
  return TQ_SUCCESS;
}

// ---- plan object ----------------------------------------------------------

// Expands the work list into per-tile pair indices (tile_pair) and each
// pair's first tile; `total` is known on the host from the shapes.
static tq_status expand_work(tq_plan_t pl, int32_t bm, int32_t bn, int32_t cls,
                             int64_t total, DevBuf &item_pair, DevBuf &item_first) {
  tq_handle_t h = pl->h;
  const int64_t np = pl->n_pairs;
  DevBuf items, tmp;
  TQ_HIP(h, items.alloc(sizeof(i64) * (np + 1)));
  TQ_HIP(h, tmp.alloc(tq_scan_tmp_bytes(np + 1)));
  TQ_HIP(h, item_pair.alloc(sizeof(int32_t) * std::max<int64_t>(total, 1)));
  TQ_HIP(h, item_first.alloc(sizeof(int32_t) * std::max<int64_t>(np, 1)));
  hipLaunchKernelGGL(tiles_per_position_kernel, dim3(grid_for(np)), dim3(256), 0,
                     h->stream, np, pl->order.as<i64>(), pl->desc.as<GemmDesc>(), bm, bn,
                     (int32_t)pl->stream_max_k, (int32_t)pl->large_min, cls,
                     items.as<i64>());
  TQ_CHECK(tq_scan_i64(h, items.as<i64>(), items.as<i64>(), np, items.as<i64>() + np,
                       tmp.p));
  hipLaunchKernelGGL(expand_tiles_kernel, dim3(grid_for(np)), dim3(256), 0, h->stream,
                     np, pl->order.as<i64>(), items.as<i64>(), item_pair.as<int32_t>(),
                     item_first.as<int32_t>());
  TQ_HIP(h, hipGetLastError());
  TQ_HIP(h, hipStreamSynchronize(h->stream));
  return TQ_SUCCESS;
}

// Runs: maximal stretches of consecutive pairs (pair order = output order)
// of one MFMA tile class that share their A operand and whose output blocks
// follow each other, each with its tile class; then listed by K, largest
// first (stable), as the pairs of the work list are.
static tq_status build_runs(tq_plan_t pl) {
  tq_handle_t h = pl->h;
  const int64_t np = pl->n_pairs;
  if ((int64_t)pl->h_desc.size() != np)
    return tq_fail(h, TQ_ERROR_INVALID_VALUE, "contract: host descriptors missing");
  std::vector<RunDesc> runs;
  int prev_cls = -2;
  for (int64_t p = 0; p < np; ++p) {
    const GemmDesc &d = pl->h_desc[p];
    const int cls = pair_class(*pl, d.m, d.n, d.k);
    if (cls < 0) {
      prev_cls = -2;
      continue;
    }
    const int32_t af = d.flags & (kAKFast | kAPacked);
    if (!runs.empty() && cls == prev_cls) {
      RunDesc &r = runs.back();
      const GemmDesc &last = pl->h_desc[r.p0 + r.np - 1];
      if (r.p0 + r.np == p && r.a_off == d.a_off && r.m == d.m && r.k == d.k &&
          r.lda == d.lda && r.flags == af && last.c_off + (i64)last.m * last.n == d.c_off &&
          (i64)r.n + d.n <= 0x7fffffff) {
        r.n += d.n;
        ++r.np;
        continue;
      }
    }
    runs.push_back(RunDesc{d.a_off, d.c_off, d.m, d.n, d.k, d.lda, af, (int32_t)p, 1});
    prev_cls = cls;
  }
  std::stable_sort(runs.begin(), runs.end(), [](const RunDesc &x, const RunDesc &y) {
    return 63 - __builtin_clzll((unsigned long long)pair_order_key(x.m, x.n, x.k)) >
           63 - __builtin_clzll((unsigned long long)pair_order_key(y.m, y.n, y.k));
  });
  pl->h_runs.swap(runs);
  TQ_HIP(h, pl->run_desc.alloc(sizeof(RunDesc) * std::max<size_t>(pl->h_runs.size(), 1)));
  if (!pl->h_runs.empty())
    TQ_HIP(h, hipMemcpyAsync(pl->run_desc.p, pl->h_runs.data(),
                             sizeof(RunDesc) * pl->h_runs.size(), hipMemcpyHostToDevice,
                             h->stream));
  TQ_HIP(h, hipStreamSynchronize(h->stream));
  pl->runs_built = true;
  return TQ_SUCCESS;
}

// Tile list of the runs of class cls under a bm x bn tile: run index per tile
// and each run's first tile.
static tq_status build_run_tiles(tq_plan_t pl, int cls, int32_t bm, int32_t bn,
                                 TileClass &tc) {
  tq_handle_t h = pl->h;
  const int64_t nr = (int64_t)pl->h_runs.size();
  std::vector<int32_t> first(std::max<int64_t>(nr, 1), 0), tiles;
  for (int64_t u = 0; u < nr; ++u) {
    const RunDesc &r = pl->h_runs[u];
    first[u] = (int32_t)tiles.size();
    // every pair of a run is in one class
    if (pair_class(*pl, r.m, pl->h_desc[r.p0].n, r.k) != cls) continue;
    const int64_t nt = ceil_div(r.m, bm) * ceil_div(r.n, bn);
    if ((int64_t)tiles.size() + nt > 0x7fffffff)
      return tq_fail(h, TQ_ERROR_UNSUPPORTED, "contract: too many GEMM tiles");
    tiles.insert(tiles.end(), (size_t)nt, (int32_t)u);
  }
  TQ_HIP(h, tc.tile_pair.alloc(sizeof(int32_t) * std::max<size_t>(tiles.size(), 1)));
  TQ_HIP(h, tc.tile_first.alloc(sizeof(int32_t) * first.size()));
  if (!tiles.empty())
    TQ_HIP(h, hipMemcpyAsync(tc.tile_pair.p, tiles.data(), sizeof(int32_t) * tiles.size(),
                             hipMemcpyHostToDevice, h->stream));
  TQ_HIP(h, hipMemcpyAsync(tc.tile_first.p, first.data(), sizeof(int32_t) * first.size(),
                           hipMemcpyHostToDevice, h->stream));
  TQ_HIP(h, hipStreamSynchronize(h->stream));
  tc.n_tiles = (int64_t)tiles.size();
  return TQ_SUCCESS;
}

// Work items of the streaming small-K kernel, in work-list order: each
// streamed pair is cut into column slices of at most kStreamCols columns (a
// slice is a pair of its own for the kernel: B and C offsets moved to its
// first column), and each slice into kStreamRows-row segments. Slicing wide
// blocks keeps the items small enough to balance over the wavefronts, which
// take them round-robin.
static tq_status build_stream_segments(tq_plan_t pl) {
  tq_handle_t h = pl->h;
  const int64_t np = pl->n_pairs;
  pl->n_chunks = 0;
  pl->chunks_ready = true;
  if (np == 0) return TQ_SUCCESS;
  const bool split = pl->prec == TQ_PREC_BF16X3 && !pl->split_direct;
  // slices per work-list position, their scan, then the slices and their
  // segment counts, that scan, then the segment list: all on the device,
  // two 8-byte totals read back
  DevBuf slice_start, seg_start, tmp;
  TQ_HIP(h, slice_start.alloc(sizeof(i64) * (np + 1)));
  TQ_HIP(h, tmp.alloc(tq_scan_tmp_bytes(np + 1)));
  hipLaunchKernelGGL(stream_slices_kernel, dim3(grid_for(np)), dim3(256), 0, h->stream, np,
                     pl->order.as<i64>(), pl->desc.as<GemmDesc>(),
                     (int32_t)(pl->member_stream ? 0x7fffffff : pl->stream_max_k),
                     slice_start.as<i64>());
  TQ_CHECK(tq_scan_i64(h, slice_start.as<i64>(), slice_start.as<i64>(), np,
                       slice_start.as<i64>() + np, tmp.p));
  i64 n_slices = 0;
  TQ_HIP(h, hipMemcpyAsync(&n_slices, slice_start.as<i64>() + np, sizeof(i64),
                           hipMemcpyDeviceToHost, h->stream));
  TQ_HIP(h, hipStreamSynchronize(h->stream));
  if (n_slices == 0) return TQ_SUCCESS;
  if (n_slices > 0x7fffffff)
    return tq_fail(h, TQ_ERROR_UNSUPPORTED, "contract: too many output segments");
  GemmDesc *slices = nullptr;
  GroupSlice *gslices = nullptr;
  if (pl->member_stream) {
    TQ_HIP(h, pl->group_slices.alloc(sizeof(GroupSlice) * n_slices));
    gslices = pl->group_slices.as<GroupSlice>();
  } else {
    TQ_HIP(h, pl->stream_desc.alloc(sizeof(GemmDesc) * n_slices));
    slices = pl->stream_desc.as<GemmDesc>();
  }
  TQ_HIP(h, seg_start.alloc(sizeof(i64) * (n_slices + 1)));
  TQ_HIP(h, tmp.alloc(tq_scan_tmp_bytes(n_slices + 1)));
  hipLaunchKernelGGL(stream_slice_desc_kernel, dim3(grid_for(np)), dim3(256), 0, h->stream,
                     np, pl->order.as<i64>(), pl->desc.as<GemmDesc>(), slice_start.as<i64>(),
                     (int32_t)split, pl->member_stream ? pl->member_first.as<i64>() : nullptr,
                     slices, gslices, seg_start.as<i64>());
  TQ_CHECK(tq_scan_i64(h, seg_start.as<i64>(), seg_start.as<i64>(), n_slices,
                       seg_start.as<i64>() + n_slices, tmp.p));
  i64 n_segs = 0;
  TQ_HIP(h, hipMemcpyAsync(&n_segs, seg_start.as<i64>() + n_slices, sizeof(i64),
                           hipMemcpyDeviceToHost, h->stream));
  TQ_HIP(h, hipStreamSynchronize(h->stream));
  if (n_segs > 0x7fffffff)
    return tq_fail(h, TQ_ERROR_UNSUPPORTED, "contract: too many output segments");
  TQ_HIP(h, pl->chunk_pair.alloc(sizeof(int32_t) * std::max<i64>(n_segs, 1)));
  TQ_HIP(h, pl->chunk_first.alloc(sizeof(int32_t) * n_slices));
  hipLaunchKernelGGL(stream_segments_kernel, dim3(grid_for(n_slices)), dim3(256), 0,
                     h->stream, n_slices, seg_start.as<i64>(), pl->chunk_pair.as<int32_t>(),
                     pl->chunk_first.as<int32_t>());
  TQ_HIP(h, hipGetLastError());
  // the temporaries are freed on return
  TQ_HIP(h, hipStreamSynchronize(h->stream));
  pl->n_chunks = n_segs;
  return TQ_SUCCESS;
}

// Tile list of a split-K class: every pair of the class (in work-list
// order) becomes up to `split` GEMMs over slices of its contracted range,
// each writing its own block of the workspace; the reduction jobs then add a
// pair's slices into its output block.
static tq_status build_split_tiles(tq_plan_t pl, int cls, const GemmConfig &g,
                                   TileClass &tc) {
  tq_handle_t h = pl->h;
  const int64_t np = pl->n_pairs;
  std::vector<i64> order(np);
  TQ_HIP(h, hipMemcpyAsync(order.data(), pl->order.p, sizeof(i64) * np, hipMemcpyDeviceToHost,
                           h->stream));
  TQ_HIP(h, hipStreamSynchronize(h->stream));
  // split-BF16 panels are k-chunked: a slice starting at k0 (a multiple of
  // 32) begins k0 * rows values into the panel and holds round8(slice) k
  const bool chunked = pl->prec == TQ_PREC_BF16X3 && !pl->split_direct;
  std::vector<GemmDesc> vd;
  std::vector<SplitReduceJob> jobs;
  std::vector<int32_t> tiles, first;
  i64 ws_off = 0;
  for (int64_t q = 0; q < np; ++q) {
    const GemmDesc &d = pl->h_desc[order[q]];
    if (pair_class(*pl, d.m, d.n, d.k) != cls) continue;
    const int64_t ks = split_slice_k(d.k, tc.split);
    const int64_t ns = ceil_div(d.k, ks);
    const i64 mn = (i64)d.m * d.n;
    for (int64_t sl = 0; sl < ns; ++sl) {
      const i64 k0 = sl * ks;
      const int32_t kl = (int32_t)std::min<i64>(ks, d.k - k0);
      GemmDesc v = d;
      if (chunked && (d.flags & kAPacked)) {
        v.a_off += k0 * d.m;
        v.lda = (kl + 7) / 8 * 8;
      } else {
        v.a_off += (d.flags & kAKFast) ? k0 : k0 * d.lda;
      }
      if (chunked && (d.flags & kBPacked)) {
        v.b_off += k0 * d.n;
        v.ldb = (kl + 7) / 8 * 8;
      } else {
        v.b_off += (d.flags & kBKFast) ? k0 : k0 * d.ldb;
      }
      v.k = kl;
      v.c_off = ws_off + sl * mn;
      const int64_t nt = ceil_div(d.m, g.bm) * ceil_div(d.n, g.bn);
      if ((int64_t)tiles.size() + nt > 0x7fffffff || vd.size() >= 0x7fffffff)
        return tq_fail(h, TQ_ERROR_UNSUPPORTED, "contract: too many split-K tiles");
      first.push_back((int32_t)tiles.size());
      tiles.insert(tiles.end(), (size_t)nt, (int32_t)vd.size());
      vd.push_back(v);
    }
    for (i64 e0 = 0; e0 < mn; e0 += kSplitReduceElems)
      jobs.push_back(SplitReduceJob{d.c_off + e0, ws_off + e0,
                                    (int32_t)std::min<i64>(kSplitReduceElems, mn - e0),
                                    (int32_t)mn, (int32_t)ns, 0});
    ws_off += ns * mn;
  }
  const size_t es = tq_elem_size(pl->dtype);
  TQ_HIP(h, tc.vdesc.alloc(sizeof(GemmDesc) * std::max<size_t>(vd.size(), 1)));
  TQ_HIP(h, tc.tile_pair.alloc(sizeof(int32_t) * std::max<size_t>(tiles.size(), 1)));
  TQ_HIP(h, tc.tile_first.alloc(sizeof(int32_t) * std::max<size_t>(first.size(), 1)));
  TQ_HIP(h, tc.red_jobs.alloc(sizeof(SplitReduceJob) * std::max<size_t>(jobs.size(), 1)));
  TQ_HIP(h, tc.ws.alloc(es * std::max<i64>(ws_off, 1)));
  if (!vd.empty()) {
    TQ_HIP(h, hipMemcpyAsync(tc.vdesc.p, vd.data(), sizeof(GemmDesc) * vd.size(),
                             hipMemcpyHostToDevice, h->stream));
    TQ_HIP(h, hipMemcpyAsync(tc.tile_pair.p, tiles.data(), sizeof(int32_t) * tiles.size(),
                             hipMemcpyHostToDevice, h->stream));
    TQ_HIP(h, hipMemcpyAsync(tc.tile_first.p, first.data(), sizeof(int32_t) * first.size(),
                             hipMemcpyHostToDevice, h->stream));
    TQ_HIP(h, hipMemcpyAsync(tc.red_jobs.p, jobs.data(), sizeof(SplitReduceJob) * jobs.size(),
                             hipMemcpyHostToDevice, h->stream));
  }
  TQ_HIP(h, hipStreamSynchronize(h->stream));
  tc.n_tiles = (int64_t)tiles.size();
  tc.n_red_jobs = (int64_t)jobs.size();
  return TQ_SUCCESS;
}

tq_status plan_build_tiles(tq_plan_t pl) {
  tq_handle_t h = pl->h;
  const int64_t np = pl->n_pairs;
  if (!pl->chunks_ready) TQ_CHECK(build_stream_segments(pl));
  pl->n_tiles = 0;
  for (int c = 0; c < 2; ++c) {
    TileClass &tc = pl->cls[c];
    const int cfg = tc.current();
    TQ_HIP(h, tc.tile_ctr.alloc(sizeof(uint32_t) * kTileCounters));
    if (!(tc.ready && tc.tiles_cfg == cfg && tc.tiles_split == tc.split)) {
      const GemmConfig g = tq_gemm_config(pl->prec, cfg);
      if (tc.split > 1) {
        TQ_CHECK(build_split_tiles(pl, c, g, tc));
        tc.runs = false;
        tc.tiles_cfg = cfg;
        tc.tiles_split = tc.split;
        tc.ready = true;
        pl->n_tiles += tc.n_tiles;
        continue;
      }
      tc.tiles_split = 1;
      // the wave-tile configurations (FP64 and FP32 tiers) and every
      // split-BF16 configuration tile runs of pairs when runs join pairs of
      // this class (otherwise pair tiles are the same work with cheaper B
      // addressing)
      if (g.wave_tile || pl->prec == TQ_PREC_BF16X3) {
        if (!pl->runs_built) TQ_CHECK(build_runs(pl));
        int64_t n_runs = 0, n_cls = 0;
        for (const RunDesc &r : pl->h_runs)
          if (pair_class(*pl, r.m, pl->h_desc[r.p0].n, r.k) == c) {
            ++n_runs;
            n_cls += r.np;
          }
        if (n_runs < n_cls) {
          TQ_CHECK(build_run_tiles(pl, c, g.bm, g.bn, tc));
          tc.runs = true;
          tc.tiles_cfg = cfg;
          tc.ready = true;
          pl->n_tiles += tc.n_tiles;
          continue;
        }
      }
      tc.runs = false;
      int64_t total = 0;
      for (int64_t p = 0; p < np; ++p)
        if (pair_class(*pl, pl->h_m[p], pl->h_n[p], pl->h_k[p]) == c)
          total += ceil_div(pl->h_m[p], g.bm) * ceil_div(pl->h_n[p], g.bn);
      if (total > 0x7fffffff)
        return tq_fail(h, TQ_ERROR_UNSUPPORTED, "contract: too many GEMM tiles");
      if (total > 0) TQ_CHECK(expand_work(pl, g.bm, g.bn, c, total, tc.tile_pair, tc.tile_first));
      tc.n_tiles = total;
      tc.tiles_cfg = cfg;
      tc.ready = true;
    }
    pl->n_tiles += tc.n_tiles;
  }
  return TQ_SUCCESS;
}

// pair_dst (per-pair panels only, may be null): the element offset of each
// pair's panel and, last, the panel's total size (n_pairs + 1 entries)
// instead of consecutive panels -- the K-concatenated group panels of a
// merged-output plan.
static tq_status build_pack(tq_plan_t pl, bool is_b, int mode, bool split,
                            const PlanUpload &up,
                            const std::vector<i64> *pair_dst = nullptr,
                            const std::vector<i64> *pair_ld = nullptr) {
  tq_handle_t h = pl->h;
  const int64_t np = pl->n_pairs;
  PackSide &side = is_b ? pl->pack_b : pl->pack_a;
  side.mode = mode;
  side.built_mode = mode;
  side.split = split;
  if (mode == kModeDirect) return TQ_SUCCESS;
  const int self_rank = is_b ? pl->rb : pl->ra;
  const int other_rank = is_b ? pl->ra : pl->rb;
  const int nl = pl->nl;
  BoxArgs box{};
  box.n_pairs = np;
  box.is_b = is_b;
  box.split = split;
  box.kfast = !split && (is_b ? pl->kfast_panel_b : pl->kfast_panel_a);
  box.self_rank = self_rank;
  box.other_rank = other_rank;
  box.nl = nl;
  box.self_begin = is_b ? up.b_begin : up.a_begin;
  box.self_end = is_b ? up.b_end : up.a_end;
  box.other_begin = is_b ? up.a_begin : up.b_begin;
  box.other_end = is_b ? up.a_end : up.b_end;
  box.n_free = self_rank - nl;
  for (int q = 0; q < box.n_free; ++q)
    box.free_legs[q] = is_b ? pl->b_free[q] : pl->a_free[q];
  for (int k = 0; k < nl; ++k) {
    box.self_legs[k] = is_b ? pl->legs_b[k] : pl->legs_a[k];
    box.other_legs[k] = is_b ? pl->legs_a[k] : pl->legs_b[k];
  }
  const int32_t *pair_self = is_b ? pl->pair_b.as<int32_t>() : pl->pair_a.as<int32_t>();
  const int32_t *pair_other = is_b ? pl->pair_a.as<int32_t>() : pl->pair_b.as<int32_t>();
  const i64 *self_str = is_b ? up.b_str : up.a_str;
  const i64 *self_off = is_b ? up.b_off : up.a_off;
  DevBuf sizes, tmp;
  int64_t n_jobs;
  if (mode == kModeBlockPanel) {
    const int64_t nb = is_b ? pl->nb : pl->na;
    n_jobs = nb;
    TQ_HIP(h, sizes.alloc(sizeof(i64) * (nb + 1)));
    TQ_HIP(h, tmp.alloc(tq_scan_tmp_bytes(nb + 1)));
    TQ_HIP(h, side.panel_off.alloc(sizeof(i64) * (nb + 1)));
    TQ_HIP(h, side.jobs.alloc(sizeof(CopyJob) * std::max<int64_t>(nb, 1)));
    TQ_HIP(h, side.wg.alloc(sizeof(i64) * (nb + 1)));
    TQ_HIP(h, side.wg32.alloc(sizeof(int32_t) * (nb + 1)));
    const uint8_t *used = is_b ? pl->b_used.as<uint8_t>() : pl->a_used.as<uint8_t>();
    PanelArgs pa{};
    pa.n_blocks = nb;
    pa.rank = self_rank;
    pa.nk = nl;
    pa.split = split;
    pa.begin = box.self_begin;
    pa.end = box.self_end;
    pa.strides = self_str;
    pa.offsets = self_off;
    if (split || box.kfast) {
      for (int k = 0; k < nl; ++k) pa.order[k] = box.self_legs[k];
      for (int q = 0; q < box.n_free; ++q) pa.order[nl + q] = box.free_legs[q];
      if (!split) pa.nk = 0;  // a plain gather in that leg order
    } else {
      for (int q = 0; q < box.n_free; ++q) pa.order[q] = box.free_legs[q];
      for (int k = 0; k < nl; ++k) pa.order[box.n_free + k] = box.self_legs[k];
      pa.nk = 0;
    }
    hipLaunchKernelGGL(block_panel_sizes_kernel, dim3(grid_for(nb)), dim3(256), 0,
                       h->stream, pa, used, sizes.as<i64>());
    TQ_CHECK(tq_scan_i64(h, sizes.as<i64>(), side.panel_off.as<i64>(), nb,
                         side.panel_off.as<i64>() + nb, tmp.p));
    hipLaunchKernelGGL(block_panel_jobs_kernel, dim3(grid_for(nb)), dim3(256), 0,
                       h->stream, pa, side.panel_off.as<i64>(), used,
                       side.jobs.as<CopyJob>(), side.wg.as<i64>());
    hipLaunchKernelGGL(block_panel_desc_kernel, dim3(grid_for(np)), dim3(256), 0,
                       h->stream, box, pair_self, pair_other,
                       side.panel_off.as<i64>(), pl->desc.as<GemmDesc>());
  } else {
    n_jobs = np;
    TQ_HIP(h, sizes.alloc(sizeof(i64) * (np + 1)));
    TQ_HIP(h, tmp.alloc(tq_scan_tmp_bytes(np + 1)));
    TQ_HIP(h, side.panel_off.alloc(sizeof(i64) * (np + 1)));
    TQ_HIP(h, side.jobs.alloc(sizeof(CopyJob) * std::max<int64_t>(np, 1)));
    TQ_HIP(h, side.wg.alloc(sizeof(i64) * (np + 1)));
    TQ_HIP(h, side.wg32.alloc(sizeof(int32_t) * (np + 1)));
    if (pair_dst) {
      TQ_HIP(h, hipMemcpyAsync(side.panel_off.p, pair_dst->data(), sizeof(i64) * (np + 1),
                               hipMemcpyHostToDevice, h->stream));
    } else {
      hipLaunchKernelGGL(pair_panel_sizes_kernel, dim3(grid_for(np)), dim3(256), 0,
                         h->stream, np, (int32_t)is_b, (int32_t)split,
                         pl->desc.as<GemmDesc>(), sizes.as<i64>());
      TQ_CHECK(tq_scan_i64(h, sizes.as<i64>(), side.panel_off.as<i64>(), np,
                           side.panel_off.as<i64>() + np, tmp.p));
    }
    DevBuf ld_dev;
    if (pair_ld) {
      TQ_HIP(h, ld_dev.alloc(sizeof(i64) * std::max<int64_t>(np, 1)));
      TQ_HIP(h, hipMemcpyAsync(ld_dev.p, pair_ld->data(), sizeof(i64) * np,
                               hipMemcpyHostToDevice, h->stream));
    }
    hipLaunchKernelGGL(pair_panel_jobs_kernel, dim3(grid_for(np)), dim3(256), 0,
                       h->stream, box, pair_self, pair_other, self_str, self_off,
                       side.panel_off.as<i64>(), pair_ld ? ld_dev.as<i64>() : nullptr,
                       pl->desc.as<GemmDesc>(), side.jobs.as<CopyJob>(), side.wg.as<i64>());
    // (the host vectors and ld_dev outlive the kernel: the sizes read below
    // synchronise the stream before they go)
    TQ_HIP(h, hipStreamSynchronize(h->stream));
  }
  TQ_HIP(h, hipGetLastError());
  TQ_CHECK(tq_scan_i64(h, side.wg.as<i64>(), side.wg.as<i64>(), n_jobs,
                       side.wg.as<i64>() + n_jobs, tmp.p));
  hipLaunchKernelGGL(i64_to_i32_kernel, dim3(grid_for(n_jobs + 1)), dim3(256), 0,
                     h->stream, side.wg.as<i64>(), side.wg32.as<int32_t>(),
                     n_jobs + 1);
  TQ_HIP(h, hipGetLastError());
  i64 sizes_h[2];
  TQ_HIP(h, hipMemcpyAsync(&sizes_h[0], side.panel_off.as<i64>() + n_jobs,
                           sizeof(i64), hipMemcpyDeviceToHost, h->stream));
  TQ_HIP(h, hipMemcpyAsync(&sizes_h[1], side.wg.as<i64>() + n_jobs, sizeof(i64),
                           hipMemcpyDeviceToHost, h->stream));
  TQ_HIP(h, hipStreamSynchronize(h->stream));
  if (sizes_h[1] > 0x7fffffff)
    return tq_fail(h, TQ_ERROR_UNSUPPORTED, "contract: pack too large");
  side.n_jobs = n_jobs;
  side.total_wg = sizes_h[1];
  side.panel_elems = sizes_h[0];
  if (split) {
    // three BF16 planes; row padding stays zero (packs never write it)
    const size_t bytes = 3 * sizeof(uint16_t) * std::max<int64_t>(sizes_h[0], 1);
    TQ_HIP(h, side.panel.alloc(bytes));
    TQ_HIP(h, hipMemsetAsync(side.panel.p, 0, bytes, h->stream));
    TQ_HIP(h, hipStreamSynchronize(h->stream));
  } else {
    TQ_HIP(h, side.panel.alloc(tq_elem_size(pl->dtype) * std::max<int64_t>(sizes_h[0], 1)));
  }
  return TQ_SUCCESS;
}

// Pack jobs of both operands for the geometry flags of the plan: operands
// that are not plain matrices (or every operand, for split panels) are
// permuted into block panels, or per-pair panels where blocks are used only
// in part and a block panel would not serve.
static tq_status build_operand_packs(tq_plan_t pl, int64_t flags, bool split) {
  // split panels are read with 16-byte loads: a pair that starts inside a
  // block panel row would be misaligned, so such pairs get their own panels
  if (split && (flags & kNotFullA)) flags |= kNeedPairPanelA;
  if (split && (flags & kNotFullB)) flags |= kNeedPairPanelB;
  const bool pack_a = split || (flags & kNeedPackA);
  const bool pack_b = split || (flags & kNeedPackB);
  const int mode_a = !pack_a ? kModeDirect
                     : (flags & kNeedPairPanelA) ? kModePairPanel
                                                 : kModeBlockPanel;
  const int mode_b = !pack_b ? kModeDirect
                     : (flags & kNeedPairPanelB) ? kModePairPanel
                                                 : kModeBlockPanel;
  TQ_CHECK(build_pack(pl, false, mode_a, split, pl->up));
  return build_pack(pl, true, mode_b, split, pl->up);
}

static void set_layouts(tq_plan_t pl) {
  pl->a_layout = pl->pack_a.mode == kModeDirect ? 0 : 2;
  pl->b_layout = pl->pack_b.mode == kModeDirect ? 0 : 2;
  if (pl->n_pairs > 0 && !pl->h_desc.empty()) {
    if (pl->pack_a.mode == kModeDirect && (pl->h_desc[0].flags & kAKFast)) pl->a_layout = 1;
    if (pl->pack_b.mode == kModeDirect && (pl->h_desc[0].flags & kBKFast)) pl->b_layout = 1;
  }
}

tq_status plan_set_split_form(tq_plan_t pl, bool in_place) {
  tq_handle_t h = pl->h;
  if (!pl->can_split_direct || in_place == pl->split_direct) return TQ_SUCCESS;
  const int64_t np = pl->n_pairs;
  const size_t bytes = sizeof(GemmDesc) * np;
  pl->drop_graph();
  if (in_place) {
    TQ_HIP(h, hipMemcpyAsync(pl->desc.p, pl->desc_inplace.p, bytes, hipMemcpyDeviceToDevice,
                             h->stream));
    pl->pack_a.mode = pl->pack_b.mode = kModeDirect;
    pl->h_desc = pl->h_desc_inplace;
  } else if (!pl->panels_built) {
    // desc holds the in-place descriptors the panel jobs are built from
    TQ_CHECK(build_operand_packs(pl, pl->geom_flags, true));
    TQ_HIP(h, pl->desc_panels.alloc(bytes));
    TQ_HIP(h, hipMemcpyAsync(pl->desc_panels.p, pl->desc.p, bytes, hipMemcpyDeviceToDevice,
                             h->stream));
    pl->h_desc_panels.resize(np);
    TQ_HIP(h, hipMemcpyAsync(pl->h_desc_panels.data(), pl->desc.p, bytes,
                             hipMemcpyDeviceToHost, h->stream));
    pl->panels_built = true;
  } else {
    TQ_HIP(h, hipMemcpyAsync(pl->desc.p, pl->desc_panels.p, bytes, hipMemcpyDeviceToDevice,
                             h->stream));
    pl->pack_a.mode = pl->pack_a.built_mode;
    pl->pack_b.mode = pl->pack_b.built_mode;
  }
  TQ_HIP(h, hipStreamSynchronize(h->stream));
  if (!in_place) pl->h_desc = pl->h_desc_panels;
  pl->split_direct = in_place;
  pl->max_ld = 0;
  for (const GemmDesc &d : pl->h_desc)
    pl->max_ld = std::max<int64_t>(pl->max_ld, std::max(d.lda, d.ldb));
  // runs, segments and tiles address the operands through the descriptors
  pl->runs_built = false;
  pl->chunks_ready = false;
  for (int c = 0; c < 2; ++c) pl->cls[c].ready = false;
  plan_pick_defaults(pl);
  set_layouts(pl);
  return TQ_SUCCESS;
}

// Merged output. The reference writes one output block per matched pair
// (src/tensor/block_sparse.h:548-551; src/gpu/block_sparse.cuh:972-975); a
// merged plan writes one block per distinct output block instead, holding
// the sum of the products of its pairs. Pairs with equal output begin and
// end points form a group, listed in the order of their first pair (A block,
// then B block). Each group is one GEMM over the concatenation of its pairs'
// contracted ranges: both operands are packed into per-group panels in which
// pair p's slice starts at K offset kofs_p (split-BF16 slices padded to 8
// values, the padding zero), so the grouped GEMM and the streaming kernel run
// unchanged with one unit per group. With member_stream the operands keep the
// unmerged plan's packing (hd holds its descriptors) and the streaming
// kernel reads each group's pairs in place of the group panels (a group's
// sum runs over its pairs' contracted values in the same order as over the
// concatenated range).
static tq_status merge_groups(tq_plan_t pl, const std::vector<GemmDesc> &hd,
                              double operand_elems, bool member_stream) {
  tq_handle_t h = pl->h;
  const int64_t np = pl->n_pairs;
  const int ro = pl->ro;
  const bool split = pl->prec == TQ_PREC_BF16X3;
  std::map<std::vector<i64>, int64_t> key_group;
  std::vector<int64_t> grp(np), first;
  std::vector<i64> key(2 * ro);
  for (int64_t p = 0; p < np; ++p) {
    std::copy_n(pl->h_out_begin.begin() + p * ro, ro, key.begin());
    std::copy_n(pl->h_out_end.begin() + p * ro, ro, key.begin() + ro);
    auto it = key_group.emplace(key, (int64_t)first.size());
    if (it.second) first.push_back(p);
    grp[p] = it.first->second;
  }
  const int64_t ng = (int64_t)first.size();
  std::vector<i64> kofs(np), ktot(ng, 0);
  for (int64_t p = 0; p < np; ++p) {
    const i64 kp = split ? (hd[p].k + 7) / 8 * 8 : hd[p].k;
    kofs[p] = ktot[grp[p]];
    ktot[grp[p]] += kp;
  }
  std::vector<i64> a_base(ng), b_base(ng);
  i64 a_tot = 0, b_tot = 0;
  if (member_stream) {
    // each group's pairs in pair order
    pl->h_member_first.assign(ng + 1, 0);
    for (int64_t p = 0; p < np; ++p) ++pl->h_member_first[grp[p] + 1];
    for (int64_t g = 0; g < ng; ++g) pl->h_member_first[g + 1] += pl->h_member_first[g];
    std::vector<int64_t> fill(pl->h_member_first.begin(), pl->h_member_first.end() - 1);
    pl->h_members.resize(np);
    for (int64_t p = 0; p < np; ++p) pl->h_members[fill[grp[p]]++] = hd[p];
    TQ_HIP(h, pl->member_desc.alloc(sizeof(GemmDesc) * std::max<int64_t>(np, 1)));
    TQ_HIP(h, hipMemcpyAsync(pl->member_desc.p, pl->h_members.data(), sizeof(GemmDesc) * np,
                             hipMemcpyHostToDevice, h->stream));
    TQ_HIP(h, pl->member_first.alloc(sizeof(i64) * (ng + 1)));
    TQ_HIP(h, hipMemcpyAsync(pl->member_first.p, pl->h_member_first.data(),
                             sizeof(i64) * (ng + 1), hipMemcpyHostToDevice, h->stream));
    pl->member_stream = true;
  }
  for (int64_t g = 0; g < ng && !member_stream; ++g) {
    const GemmDesc &d = hd[first[g]];
    if (ktot[g] > 0x7fffffff || d.m * ktot[g] > 0x7fffffff || d.n * ktot[g] > 0x7fffffff)
      return tq_fail(h, TQ_ERROR_UNSUPPORTED, "contract: merged block GEMM above 2^31 elements");
    a_base[g] = a_tot;
    b_base[g] = b_tot;
    a_tot += d.m * ktot[g];
    b_tot += d.n * ktot[g];
  }
  // a pair's slice of its group panel: split and m-fast panels hold it as
  // kofs_p whole columns (chunks); a k-fast plain panel as a column band of
  // rows ktot long
  const bool kfa = !split && pl->kfast_panel_a, kfb = !split && pl->kfast_panel_b;
  std::vector<i64> dst_a(np + 1), dst_b(np + 1), ld_a(np), ld_b(np);
  for (int64_t p = 0; p < np; ++p) {
    dst_a[p] = a_base[grp[p]] + (kfa ? kofs[p] : kofs[p] * hd[p].m);
    dst_b[p] = b_base[grp[p]] + (kfb ? kofs[p] : kofs[p] * hd[p].n);
    ld_a[p] = ld_b[p] = ktot[grp[p]];
  }
  dst_a[np] = a_tot;
  dst_b[np] = b_tot;
  if (!member_stream) {
    TQ_CHECK(build_pack(pl, false, kModePairPanel, split, pl->up, &dst_a, &ld_a));
    TQ_CHECK(build_pack(pl, true, kModePairPanel, split, pl->up, &dst_b, &ld_b));
  }

  // one GEMM unit per group; the matched pairs keep their own maps
  pl->merged = true;
  pl->n_matched = np;
  pl->h_pair_m.swap(pl->h_m);
  pl->h_pair_n.swap(pl->h_n);
  pl->h_pair_k.swap(pl->h_k);
  std::vector<GemmDesc> gd(ng);
  std::vector<i64> cost(ng), ob((size_t)ng * ro), oe((size_t)ng * ro), coff(ng + 1);
  pl->h_m.resize(ng);
  pl->h_n.resize(ng);
  pl->h_k.resize(ng);
  // (streamed groups keep the pairs' leading dimensions: their members are
  // the GEMM operands)
  if (!member_stream) pl->max_ld = 0;
  i64 c = 0;
  const bool ka = split || kfa, kb = split || kfb;
  const int32_t kf = (ka ? kAKFast : 0) | (kb ? kBKFast : 0);
  for (int64_t g = 0; g < ng; ++g) {
    const int64_t p0 = first[g];
    const GemmDesc &d = hd[p0];
    const int32_t kt = (int32_t)ktot[g];
    gd[g] = GemmDesc{a_base[g], b_base[g], c, d.m, d.n, kt, ka ? kt : d.m,
                     kb ? kt : d.n, kAPacked | kBPacked | kf};
    std::copy_n(pl->h_out_begin.begin() + p0 * ro, ro, ob.begin() + g * ro);
    std::copy_n(pl->h_out_end.begin() + p0 * ro, ro, oe.begin() + g * ro);
    coff[g] = c;
    c += (i64)d.m * d.n;
    pl->h_m[g] = d.m;
    pl->h_n[g] = d.n;
    pl->h_k[g] = kt;
    // (a streamed group costs its output writes and K-long sums, whatever K)
    cost[g] = member_stream ? (i64)d.m * d.n * (kt + 4) : pair_order_key(d.m, d.n, kt);
    pl->max_ld = std::max<int64_t>(pl->max_ld, std::max(gd[g].lda, gd[g].ldb));
  }
  coff[ng] = c;
  pl->n_pairs = ng;
  pl->h_desc = gd;
  pl->h_out_begin.swap(ob);
  pl->h_out_end.swap(oe);
  pl->h_c_off.swap(coff);
  pl->out_n_elem = c;
  pl->bytes = (double)tq_elem_size(pl->dtype) * (operand_elems + (double)c);
  TQ_HIP(h, hipMemcpyAsync(pl->desc.p, gd.data(), sizeof(GemmDesc) * ng, hipMemcpyHostToDevice,
                           h->stream));
  DevBuf dcost;
  TQ_HIP(h, dcost.alloc(sizeof(i64) * ng));
  TQ_HIP(h, hipMemcpyAsync(dcost.p, cost.data(), sizeof(i64) * ng, hipMemcpyHostToDevice,
                           h->stream));
  TQ_HIP(h, pl->order.alloc(sizeof(i64) * ng));
  TQ_CHECK(tq_sort_by_cost(h, dcost.as<i64>(), ng, pl->order.as<i64>()));
  return TQ_SUCCESS;
}

extern "C" {

tq_status tq_plan_set_split_operands(tq_plan_t pl, int32_t form) {
  if (!pl) return TQ_ERROR_NOT_INITIALIZED;
  if (form < 0 || form > 2) return TQ_ERROR_INVALID_VALUE;
  if (pl->prec != TQ_PREC_BF16X3)
    return tq_fail(pl->h, TQ_ERROR_UNSUPPORTED, "split operands: not a split-BF16 plan");
  if (form == 2 && !pl->can_split_direct)
    return tq_fail(pl->h, TQ_ERROR_UNSUPPORTED,
                   "split operands: in place needs both operands to be plain matrices");
  pl->split_form = form;
  if (form == 0) return TQ_SUCCESS;
  return plan_set_split_form(pl, form == 2);
}

}  // extern "C"

static tq_status plan_contract(tq_handle_t h, const tq_bs_desc *a, const tq_bs_desc *b,
                              int32_t n_legs, const int32_t *legs, tq_dtype dtype,
                              tq_precision precision, bool merged, tq_plan_t *plan_out) {
  if (!h) return TQ_ERROR_NOT_INITIALIZED;
  if (!plan_out) return TQ_ERROR_INVALID_VALUE;
  *plan_out = nullptr;
  if (!valid_desc(a) || !valid_desc(b))
    return tq_fail(h, TQ_ERROR_INVALID_VALUE, "contract: invalid tensor metadata");
  if (!tq_valid_dtype(dtype)) return TQ_ERROR_UNSUPPORTED;
  if (precision != TQ_PREC_NATIVE && dtype != TQ_F64)
    return tq_fail(h, TQ_ERROR_UNSUPPORTED,
                   "contract: reduced-precision tiers take FP64 data");
  if (precision < TQ_PREC_NATIVE || precision > TQ_PREC_BF16X3)
    return TQ_ERROR_UNSUPPORTED;
  if (n_legs < 0 || n_legs > std::min(a->rank, b->rank) || (n_legs > 0 && !legs) ||
      !valid_legs(a->rank, b->rank, n_legs, legs))
    return tq_fail(h, TQ_ERROR_INVALID_VALUE, "contract: invalid contracting indices");
  for (int k = 0; k < n_legs; ++k)
    if (a->dims[legs[2 * k]] != b->dims[legs[2 * k + 1]])
      return tq_fail(h, TQ_ERROR_DIMENSION_MISMATCH,
                     "The dimensions from two tensors to be contracted do not match");

  std::unique_ptr<tq_plan_s> pl(new tq_plan_s());
  pl->h = h;
  pl->dtype = dtype;
  pl->prec = precision;
  pl->ra = a->rank;
  pl->rb = b->rank;
  pl->nl = n_legs;
  pl->na = a->n_blocks;
  pl->nb = b->n_blocks;
  for (int k = 0; k < n_legs; ++k) {
    pl->legs_a[k] = legs[2 * k];
    pl->legs_b[k] = legs[2 * k + 1];
  }
  // free legs in ascending order (arma shed_rows, block_sparse.cuh:611-616)
  {
    int q = 0;
    for (int l = 0; l < a->rank; ++l) {
      bool c = false;
      for (int k = 0; k < n_legs; ++k) c |= legs[2 * k] == l;
      if (!c) {
        pl->a_free[q++] = l;
        pl->out_dims.push_back(a->dims[l]);
      }
    }
    q = 0;
    for (int l = 0; l < b->rank; ++l) {
      bool c = false;
      for (int k = 0; k < n_legs; ++k) c |= legs[2 * k + 1] == l;
      if (!c) {
        pl->b_free[q++] = l;
        pl->out_dims.push_back(b->dims[l]);
      }
    }
  }
  pl->ro = (int)pl->out_dims.size();
  // panel orientation: k-fast when the operand's fastest leg (smallest
  // stride of a non-unit leg of its first block) is contracted
  {
    auto fastest_contracted = [&](const tq_bs_desc *t, int side) {
      if (t->n_blocks <= 0 || t->rank <= 0) return false;
      int best = -1;
      i64 bs = 0;
      for (int l = 0; l < t->rank; ++l) {
        if (t->end[l] - t->begin[l] < 1) continue;
        if (best < 0 || t->strides[l] < bs) {
          best = l;
          bs = t->strides[l];
        }
      }
      for (int k = 0; k < n_legs; ++k)
        if (legs[2 * k + side] == best) return true;
      return false;
    };
    pl->kfast_panel_a = fastest_contracted(a, 0);
    pl->kfast_panel_b = fastest_contracted(b, 1);
  }
  // a rank-0 result (the network scalar) is always reduced in FP64
  if (pl->ro == 0) pl->prec = TQ_PREC_NATIVE;
  if (pl->prec == TQ_PREC_BF16X3) {
    TQ_HIP(h, pl->range_flag.alloc(sizeof(int32_t)));
    TQ_HIP(h, hipMemsetAsync(pl->range_flag.p, 0, sizeof(int32_t), h->stream));
  }
  const int ra = a->rank, rb = b->rank, ro = pl->ro, nl = n_legs;
  const int64_t na = a->n_blocks, nb = b->n_blocks;

  // ---- upload A and B metadata in one transfer ----
  const size_t n_a_meta = (size_t)na * ra * 3 + na;
  const size_t n_b_meta = (size_t)nb * rb * 3 + nb;
  const size_t up_bytes = (n_a_meta + n_b_meta) * 8 + 16;
  void *host = nullptr;
  TQ_CHECK(tq_stage(h, up_bytes, &host));
  i64 *hp = reinterpret_cast<i64 *>(host);
  auto put = [&](const int64_t *src, size_t n) {
    if (n) std::memcpy(hp, src, n * 8);
    hp += n;
  };
  put(a->begin, (size_t)na * ra);
  put(a->end, (size_t)na * ra);
  put(a->strides, (size_t)na * ra);
  put(a->offsets, (size_t)na);
  put(b->begin, (size_t)nb * rb);
  put(b->end, (size_t)nb * rb);
  put(b->strides, (size_t)nb * rb);
  put(b->offsets, (size_t)nb);
  TQ_HIP(h, pl->meta.alloc(up_bytes));
  TQ_HIP(h, hipMemcpyAsync(pl->meta.p, host, up_bytes, hipMemcpyHostToDevice,
                           h->stream));
  PlanUpload up;
  {
    i64 *dp = pl->meta.as<i64>();
    up.a_begin = dp; dp += (size_t)na * ra;
    up.a_end = dp; dp += (size_t)na * ra;
    up.a_str = dp; dp += (size_t)na * ra;
    up.a_off = dp; dp += na;
    up.b_begin = dp; dp += (size_t)nb * rb;
    up.b_end = dp; dp += (size_t)nb * rb;
    up.b_str = dp; dp += (size_t)nb * rb;
    up.b_off = dp; dp += nb;
  }
  pl->up = up;

  // ---- 1. match: count, scan, fill ----
  int64_t n_pairs = 0;
  DevBuf b_soa, counts, scan_tmp;
  const int64_t b_chunks = std::max<int64_t>(1, ceil_div(nb, kWave));
  int64_t n_split = std::min<int64_t>(b_chunks, std::max<int64_t>(1, ceil_div(2048, std::max<int64_t>(na, 1))));
  const int32_t chunks_per_split = (int32_t)ceil_div(b_chunks, n_split);
  n_split = ceil_div(b_chunks, chunks_per_split);
  const int64_t n_items = na * n_split;
  TQ_HIP(h, pl->a_used.alloc(std::max<int64_t>(na, 1)));
  TQ_HIP(h, pl->b_used.alloc(std::max<int64_t>(nb, 1)));
  TQ_HIP(h, hipMemsetAsync(pl->b_used.p, 0, std::max<int64_t>(nb, 1), h->stream));
  TQ_HIP(h, hipMemsetAsync(pl->a_used.p, 0, std::max<int64_t>(na, 1), h->stream));
  if (na > 0 && nb > 0) {
    TQ_HIP(h, b_soa.alloc(sizeof(int32_t) * 2 * std::max<int64_t>(1, nl * nb)));
    int32_t *b_lo = b_soa.as<int32_t>();
    int32_t *b_hi = b_lo + std::max<int64_t>(1, nl * nb);
    LegList ll{};
    for (int k = 0; k < nl; ++k) ll.b[k] = pl->legs_b[k];
    if (nl > 0)
      hipLaunchKernelGGL(soa_legs_kernel, dim3(grid_for(nb)), dim3(256), 0,
                         h->stream, nb, rb, nl, up.b_begin, up.b_end, ll, b_lo,
                         b_hi);
    MatchArgs mg{};
    mg.nl = nl;
    mg.a_rank = ra;
    mg.n_a = na;
    mg.n_b = nb;
    mg.chunks_per_split = chunks_per_split;
    mg.n_split = n_split;
    mg.a_begin = up.a_begin;
    mg.a_end = up.a_end;
    mg.b_lo = b_lo;
    mg.b_hi = b_hi;
    for (int k = 0; k < nl; ++k) mg.a_legs[k] = pl->legs_a[k];
    TQ_HIP(h, counts.alloc(sizeof(i64) * (n_items + 1)));
    TQ_HIP(h, scan_tmp.alloc(tq_scan_tmp_bytes(n_items + 1)));
    const unsigned mgrid = (unsigned)ceil_div(n_items, 4);
    hipLaunchKernelGGL(match_kernel<false>, dim3(mgrid), dim3(256), 0, h->stream,
                       mg, counts.as<i64>(), nullptr, nullptr, nullptr, nullptr, nullptr);
    TQ_HIP(h, hipGetLastError());
    TQ_CHECK(tq_scan_i64(h, counts.as<i64>(), counts.as<i64>(), n_items,
                         counts.as<i64>() + n_items, scan_tmp.p));
    i64 np_h = 0;
    TQ_HIP(h, hipMemcpyAsync(&np_h, counts.as<i64>() + n_items, sizeof(i64),
                             hipMemcpyDeviceToHost, h->stream));
    TQ_HIP(h, hipStreamSynchronize(h->stream));
    n_pairs = np_h;
    if (n_pairs > 0x7fffffff)
      return tq_fail(h, TQ_ERROR_UNSUPPORTED, "contract: too many block pairs");
    TQ_HIP(h, pl->pair_a.alloc(sizeof(int32_t) * std::max<int64_t>(n_pairs, 1)));
    TQ_HIP(h, pl->pair_b.alloc(sizeof(int32_t) * std::max<int64_t>(n_pairs, 1)));
    if (n_pairs > 0) {
      hipLaunchKernelGGL(match_kernel<true>, dim3(mgrid), dim3(256), 0, h->stream,
                         mg, nullptr, counts.as<i64>(), pl->pair_a.as<int32_t>(),
                         pl->pair_b.as<int32_t>(), pl->a_used.as<uint8_t>(),
                         pl->b_used.as<uint8_t>());
      TQ_HIP(h, hipGetLastError());
    }
  }
  pl->n_pairs = n_pairs;

  // ---- 2. geometry ----
  const int64_t np = n_pairs;
  DevBuf mn, cost, flags_any, ctmp;
  TQ_HIP(h, pl->desc.alloc(sizeof(GemmDesc) * std::max<int64_t>(np, 1)));
  TQ_HIP(h, pl->sub_off.alloc(sizeof(i64) * 2 * std::max<int64_t>(np, 1)));
  TQ_HIP(h, pl->out_be.alloc(sizeof(i64) * 2 * std::max<int64_t>(np * ro, 1)));
  TQ_HIP(h, mn.alloc(sizeof(i64) * (np + 1)));
  TQ_HIP(h, cost.alloc(sizeof(i64) * (np + 1)));
  TQ_HIP(h, flags_any.alloc(sizeof(int32_t)));
  TQ_HIP(h, ctmp.alloc(tq_scan_tmp_bytes(np + 1)));
  TQ_HIP(h, hipMemsetAsync(flags_any.p, 0, sizeof(int32_t), h->stream));
  i64 *d_out_begin = pl->out_be.as<i64>();
  i64 *d_out_end = d_out_begin + std::max<int64_t>(np * ro, 1);
  if (np > 0) {
    GeomArgs gg{};
    gg.n_pairs = np;
    gg.a_rank = ra;
    gg.b_rank = rb;
    gg.nl = nl;
    gg.a_begin = up.a_begin;
    gg.a_end = up.a_end;
    gg.a_str = up.a_str;
    gg.a_off = up.a_off;
    gg.b_begin = up.b_begin;
    gg.b_end = up.b_end;
    gg.b_str = up.b_str;
    gg.b_off = up.b_off;
    for (int k = 0; k < nl; ++k) {
      gg.legs.a[k] = pl->legs_a[k];
      gg.legs.b[k] = pl->legs_b[k];
    }
    for (int q = 0; q < ra - nl; ++q) gg.a_free[q] = pl->a_free[q];
    for (int q = 0; q < rb - nl; ++q) gg.b_free[q] = pl->b_free[q];
    hipLaunchKernelGGL(geometry_kernel, dim3(grid_for(np)), dim3(256), 0,
                       h->stream, gg, pl->pair_a.as<int32_t>(),
                       pl->pair_b.as<int32_t>(), pl->desc.as<GemmDesc>(),
                       pl->sub_off.as<i64>(), d_out_begin, d_out_end, mn.as<i64>(), cost.as<i64>(),
                       flags_any.as<int32_t>());
    TQ_HIP(h, hipGetLastError());
  }
  TQ_CHECK(tq_scan_i64(h, mn.as<i64>(), mn.as<i64>(), np, mn.as<i64>() + np,
                       ctmp.p));
  TQ_HIP(h, pl->c_off.alloc(sizeof(i64) * (np + 1)));
  TQ_HIP(h, hipMemcpyAsync(pl->c_off.p, mn.p, sizeof(i64) * (np + 1),
                           hipMemcpyDeviceToDevice, h->stream));
  if (np > 0)
    hipLaunchKernelGGL(set_c_offsets_kernel, dim3(grid_for(np)), dim3(256), 0,
                       h->stream, np, pl->c_off.as<i64>(), pl->desc.as<GemmDesc>());

  // ---- 3. results the host object needs (one batch of transfers) ----
  i64 flags_h = 0;
  int32_t fl32 = 0;
  i64 out_total = 0;
  pl->h_m.resize(np);
  pl->h_n.resize(np);
  pl->h_k.resize(np);
  std::vector<GemmDesc> hd(np);
  if (np > 0)
    TQ_HIP(h, hipMemcpyAsync(hd.data(), pl->desc.p, sizeof(GemmDesc) * np,
                             hipMemcpyDeviceToHost, h->stream));
  TQ_HIP(h, hipMemcpyAsync(&fl32, flags_any.p, sizeof(int32_t),
                           hipMemcpyDeviceToHost, h->stream));
  TQ_HIP(h, hipMemcpyAsync(&out_total, pl->c_off.as<i64>() + np, sizeof(i64),
                           hipMemcpyDeviceToHost, h->stream));
  TQ_HIP(h, hipStreamSynchronize(h->stream));
  flags_h = fl32;
  if (flags_h & kTooLarge)
    return tq_fail(h, TQ_ERROR_UNSUPPORTED, "contract: block GEMM above 2^31 elements");
  double flops = 0.0;
  for (int64_t p = 0; p < np; ++p) {
    pl->h_m[p] = hd[p].m;
    pl->h_n[p] = hd[p].n;
    pl->h_k[p] = hd[p].k;
    flops += 2.0 * hd[p].m * (double)hd[p].n * hd[p].k;
  }
  pl->flops = flops;
  pl->out_n_elem = ro == 0 ? 1 : out_total;
  pl->n_matched = np;
  // a rank-0 result is one scalar whatever the layout
  merged = merged && ro > 0 && np > 0;
  double operand_elems = 0;
  {
    double na_elem = 0, nb_elem = 0;
    for (int64_t i = 0; i < na; ++i) {
      double n = 1;
      for (int l = 0; l < ra; ++l) n *= a->end[i * ra + l] - a->begin[i * ra + l] + 1;
      na_elem += n;
    }
    for (int64_t j = 0; j < nb; ++j) {
      double n = 1;
      for (int l = 0; l < rb; ++l) n *= b->end[j * rb + l] - b->begin[j * rb + l] + 1;
      nb_elem += n;
    }
    operand_elems = na_elem + nb_elem;
    pl->bytes = (double)tq_elem_size(dtype) * (na_elem + nb_elem + (double)pl->out_n_elem);
  }

  // ---- output metadata on the host (the drop-in tensor keeps it there) ----
  pl->h_out_begin.resize((size_t)np * ro);
  pl->h_out_end.resize((size_t)np * ro);
  pl->h_c_off.resize(np + 1);
  if (np > 0 && ro > 0) {
    TQ_HIP(h, hipMemcpyAsync(pl->h_out_begin.data(), d_out_begin,
                             sizeof(i64) * np * ro, hipMemcpyDeviceToHost, h->stream));
    TQ_HIP(h, hipMemcpyAsync(pl->h_out_end.data(), d_out_end, sizeof(i64) * np * ro,
                             hipMemcpyDeviceToHost, h->stream));
  }
  TQ_HIP(h, hipMemcpyAsync(pl->h_c_off.data(), pl->c_off.p, sizeof(i64) * (np + 1),
                           hipMemcpyDeviceToHost, h->stream));
  TQ_HIP(h, hipStreamSynchronize(h->stream));
  // a merged plan whose pairs all write distinct output blocks is the
  // unmerged plan (one pair per block, in the same order): it keeps the
  // unmerged operand layouts (block-diagonal products read their blocks in
  // place) instead of copying every pair into a group panel of its own
  if (merged) {
    std::map<std::vector<i64>, int> seen;
    std::vector<i64> key(2 * ro);
    bool distinct = true;
    for (int64_t p = 0; p < np && distinct; ++p) {
      std::copy_n(pl->h_out_begin.begin() + p * ro, ro, key.begin());
      std::copy_n(pl->h_out_end.begin() + p * ro, ro, key.begin() + ro);
      distinct = seen.emplace(key, 0).second;
    }
    if (distinct) merged = false;
  }
  // a merged plan streams its groups when every pair is short and no group
  // holds more than kGroupMaxMembers pairs (not for the split-BF16 tier,
  // which streams from split panels)
  bool member_stream = false;
  if (merged && pl->prec != TQ_PREC_BF16X3) {
    member_stream = true;
    for (int64_t p = 0; p < np && member_stream; ++p) member_stream = hd[p].k <= kStreamMaxK;
    if (member_stream) {
      std::map<std::vector<i64>, int> count;
      std::vector<i64> key(2 * ro);
      for (int64_t p = 0; p < np && member_stream; ++p) {
        std::copy_n(pl->h_out_begin.begin() + p * ro, ro, key.begin());
        std::copy_n(pl->h_out_end.begin() + p * ro, ro, key.begin() + ro);
        member_stream = ++count[key] <= kGroupMaxMembers;
      }
    }
  }

  // ---- 4. operand layouts and pack jobs ----
  if (np > 0 && (!merged || member_stream)) {
    // The split-BF16 tier reads split panels, or, when both operands are
    // plain matrices, may read them in place and split FP64 values in
    // registers (no pack pass); it starts in place and plan_set_split_form
    // builds the panels when asked.
    pl->geom_flags = flags_h;
    pl->can_split_direct =
        pl->prec == TQ_PREC_BF16X3 && !(flags_h & (kNeedPackA | kNeedPackB));
    pl->split_direct = pl->can_split_direct;
    const bool split = pl->prec == TQ_PREC_BF16X3 && !pl->split_direct;
    if (pl->can_split_direct) {
      TQ_HIP(h, pl->desc_inplace.alloc(sizeof(GemmDesc) * np));
      TQ_HIP(h, hipMemcpyAsync(pl->desc_inplace.p, pl->desc.p, sizeof(GemmDesc) * np,
                               hipMemcpyDeviceToDevice, h->stream));
      pl->h_desc_inplace = hd;
    }
    TQ_CHECK(build_operand_packs(pl.get(), flags_h, split));

    // packing rewrote descriptors: the largest leading dimension gates the
    // wave-tile kernel's 32-bit offsets
    TQ_HIP(h, hipMemcpyAsync(hd.data(), pl->desc.p, sizeof(GemmDesc) * np,
                             hipMemcpyDeviceToHost, h->stream));
    TQ_HIP(h, hipStreamSynchronize(h->stream));
    for (int64_t p = 0; p < np; ++p)
      pl->max_ld = std::max<int64_t>(pl->max_ld, std::max(hd[p].lda, hd[p].ldb));
    pl->h_desc = hd;

    // ---- 5. work list: stable device sort by cost, largest first ----
    TQ_HIP(h, pl->order.alloc(sizeof(i64) * np));
    TQ_CHECK(tq_sort_by_cost(h, cost.as<i64>(), np, pl->order.as<i64>()));
  }
  set_layouts(pl.get());

  if (merged) {
    TQ_CHECK(merge_groups(pl.get(), hd, operand_elems, member_stream));
    set_layouts(pl.get());
  }

  // short contractions are written as streams (the split-BF16 tier streams
  // them from its split panels)
  pl->stream_max_k = kStreamMaxK;
  plan_pick_defaults(pl.get());
  if (np > 0 && ro > 0) TQ_CHECK(plan_build_tiles(pl.get()));
  *plan_out = pl.release();
  return TQ_SUCCESS;
}

extern "C" {

tq_status tq_bs_contract_plan(tq_handle_t h, const tq_bs_desc *a, const tq_bs_desc *b,
                              int32_t n_legs, const int32_t *legs, tq_dtype dtype,
                              tq_precision precision, tq_plan_t *plan_out) {
  return plan_contract(h, a, b, n_legs, legs, dtype, precision, false, plan_out);
}

tq_status tq_bs_contract_plan_merged(tq_handle_t h, const tq_bs_desc *a, const tq_bs_desc *b,
                                     int32_t n_legs, const int32_t *legs, tq_dtype dtype,
                                     tq_precision precision, tq_plan_t *plan_out) {
  return plan_contract(h, a, b, n_legs, legs, dtype, precision, true, plan_out);
}

tq_status tq_bs_contract_plan_cached(tq_handle_t h, const tq_bs_desc *a, const tq_bs_desc *b,
                                     int32_t n_legs, const int32_t *legs, tq_dtype dtype,
                                     tq_precision precision, int32_t merged,
                                     tq_plan_t *plan_out) {
  if (!h) return TQ_ERROR_NOT_INITIALIZED;
  if (!plan_out) return TQ_ERROR_INVALID_VALUE;
  *plan_out = nullptr;
  if (!valid_desc(a) || !valid_desc(b) || n_legs < 0 || n_legs > TQ_MAX_RANK ||
      (n_legs > 0 && !legs))
    return tq_fail(h, TQ_ERROR_INVALID_VALUE, "contract: invalid tensor metadata");
  // the key: every input the plan depends on, byte for byte
  std::string key;
  auto put = [&key](const void *p, size_t n) {
    key.append(reinterpret_cast<const char *>(p), n);
  };
  const int32_t head[4] = {(int32_t)dtype, (int32_t)precision, merged ? 1 : 0, n_legs};
  key.reserve(64 + 8 * (size_t)(a->n_blocks * (3 * a->rank + 1) +
                                b->n_blocks * (3 * b->rank + 1) + a->rank + b->rank));
  put(head, sizeof(head));
  put(legs, sizeof(int32_t) * 2 * n_legs);
  for (const tq_bs_desc *d : {a, b}) {
    const int64_t r = d->rank, nb = d->n_blocks;
    put(&r, 8);
    put(&nb, 8);
    put(d->dims, 8 * r);
    if (r > 0 && nb > 0) {
      put(d->begin, 8 * r * nb);
      put(d->end, 8 * r * nb);
      put(d->strides, 8 * r * nb);
      put(d->offsets, 8 * nb);
    }
  }
  for (PlanCacheEntry &e : h->plan_cache)
    if (e.key == key) {
      e.last_use = ++h->plan_clock;
      *plan_out = e.plan;
      return TQ_SUCCESS;
    }
  tq_plan_t pl = nullptr;
  TQ_CHECK(plan_contract(h, a, b, n_legs, legs, dtype, precision, merged != 0, &pl));
  constexpr size_t kCachePlans = 32;
  if (h->plan_cache.size() >= kCachePlans) {
    auto lru = std::min_element(h->plan_cache.begin(), h->plan_cache.end(),
                                [](const PlanCacheEntry &x, const PlanCacheEntry &y) {
                                  return x.last_use < y.last_use;
                                });
    lru->plan->cached = false;
    (void)tq_plan_destroy(lru->plan);
    h->plan_cache.erase(lru);
  }
  pl->cached = true;
  h->plan_cache.push_back(PlanCacheEntry{std::move(key), pl, ++h->plan_clock});
  *plan_out = pl;
  return TQ_SUCCESS;
}

tq_status tq_plan_get_info(tq_plan_t pl, tq_plan_info *info) {
  if (!pl) return TQ_ERROR_NOT_INITIALIZED;
  if (!info) return TQ_ERROR_INVALID_VALUE;
  info->out_rank = pl->ro;
  info->n_legs = pl->nl;
  info->n_pairs = pl->n_pairs;
  info->out_n_elem = pl->out_n_elem;
  info->flops = pl->flops;
  info->bytes = pl->bytes;
  info->a_layout = pl->a_layout;
  info->b_layout = pl->b_layout;
  if (pl->n_pairs > 0 && pl->ro > 0) TQ_CHECK(plan_build_tiles(pl));
  info->n_tiles = pl->n_tiles;
  info->n_segments = pl->n_chunks;
  info->n_matched = pl->merged ? pl->n_matched : pl->n_pairs;
  return TQ_SUCCESS;
}

tq_status tq_plan_get_output(tq_plan_t pl, int64_t *dims, int64_t *begin,
                             int64_t *end, int64_t *strides, int64_t *offsets) {
  if (!pl) return TQ_ERROR_NOT_INITIALIZED;
  if (pl->is_batch)
    return tq_fail(pl->h, TQ_ERROR_UNSUPPORTED, "batch: ask the member plans");
  const int ro = pl->ro;
  const int64_t np = pl->n_pairs;
  if (dims) std::copy(pl->out_dims.begin(), pl->out_dims.end(), dims);
  if (begin) std::copy(pl->h_out_begin.begin(), pl->h_out_begin.end(), begin);
  if (end) std::copy(pl->h_out_end.begin(), pl->h_out_end.end(), end);
  if (strides) {
    // compact column-major index tables (generate_index_table,
    // src/gpu/block_sparse.cuh:658-664)
    for (int64_t p = 0; p < np; ++p) {
      i64 t = 1;
      for (int l = 0; l < ro; ++l) {
        strides[p * ro + l] = t;
        t *= pl->h_out_end[p * ro + l] - pl->h_out_begin[p * ro + l] + 1;
      }
    }
  }
  if (offsets && ro > 0) std::copy(pl->h_c_off.begin(), pl->h_c_off.begin() + np, offsets);
  return TQ_SUCCESS;
}

tq_status tq_plan_get_pairs(tq_plan_t pl, int64_t *a_block, int64_t *b_block,
                            int64_t *a_offset, int64_t *b_offset, int64_t *m,
                            int64_t *n, int64_t *k) {
  if (!pl) return TQ_ERROR_NOT_INITIALIZED;
  tq_handle_t h = pl->h;
  if (pl->is_batch) return tq_fail(h, TQ_ERROR_UNSUPPORTED, "batch: ask the member plans");
  const int64_t np = pl->merged ? pl->n_matched : pl->n_pairs;
  const std::vector<i64> &hm = pl->merged ? pl->h_pair_m : pl->h_m;
  const std::vector<i64> &hn = pl->merged ? pl->h_pair_n : pl->h_n;
  const std::vector<i64> &hk = pl->merged ? pl->h_pair_k : pl->h_k;
  if (np == 0) return TQ_SUCCESS;
  std::vector<int32_t> pa(np), pb(np);
  TQ_HIP(h, hipMemcpyAsync(pa.data(), pl->pair_a.p, sizeof(int32_t) * np,
                           hipMemcpyDeviceToHost, h->stream));
  TQ_HIP(h, hipMemcpyAsync(pb.data(), pl->pair_b.p, sizeof(int32_t) * np,
                           hipMemcpyDeviceToHost, h->stream));
  TQ_HIP(h, hipStreamSynchronize(h->stream));
  for (int64_t p = 0; p < np; ++p) {
    if (a_block) a_block[p] = pa[p];
    if (b_block) b_block[p] = pb[p];
    if (m) m[p] = hm[p];
    if (n) n[p] = hn[p];
    if (k) k[p] = hk[p];
  }
  if (a_offset || b_offset) {
    // sub-block offsets as the reference computes them
    // (src/gpu/block_sparse.cuh:700-732)
    std::vector<i64> sub(np * 2);
    TQ_HIP(h, hipMemcpyAsync(sub.data(), pl->sub_off.p, sizeof(i64) * 2 * np,
                             hipMemcpyDeviceToHost, h->stream));
    TQ_HIP(h, hipStreamSynchronize(h->stream));
    for (int64_t p = 0; p < np; ++p) {
      if (a_offset) a_offset[p] = sub[2 * p];
      if (b_offset) b_offset[p] = sub[2 * p + 1];
    }
  }
  return TQ_SUCCESS;
}

tq_status tq_plan_get_worklist(tq_plan_t pl, int64_t *order) {
  if (!pl) return TQ_ERROR_NOT_INITIALIZED;
  if (!order) return TQ_ERROR_INVALID_VALUE;
  if (pl->n_pairs == 0) return TQ_SUCCESS;
  TQ_HIP(pl->h, hipMemcpyAsync(order, pl->order.p, sizeof(i64) * pl->n_pairs,
                               hipMemcpyDeviceToHost, pl->h->stream));
  TQ_HIP(pl->h, hipStreamSynchronize(pl->h->stream));
  return TQ_SUCCESS;
}

tq_status tq_plan_destroy(tq_plan_t pl) {
  if (!pl) return TQ_ERROR_NOT_INITIALIZED;
  if (pl->cached)
    return tq_fail(pl->h, TQ_ERROR_INVALID_VALUE,
                   "plan destroy: the plan belongs to its handle's cache");
  (void)hipStreamSynchronize(pl->h->stream);
  delete pl;
  return TQ_SUCCESS;
}

}  // extern "C"
