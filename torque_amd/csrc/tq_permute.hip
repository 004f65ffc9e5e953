// tq_permute.hip -- tensor permutation kernels (replacing cuTT and the
// reference's reshape_kernel gather/scatter, src/gpu/block_sparse.cu:9-64).
//
// Two kernels, both pure copies of the raw bits (bit-exact):
//  * gather: each work item covers kGatherPerWG output elements of one job;
//    output writes are coalesced, each source address comes from a
//    column-major decode of the output index. Used for blocks and for
//    permutations that keep the fastest leg in place.
//  * tiled: 64x64 tiles through LDS for permutations that move the input's
//    fastest leg. The tile's "output group" is the leading output legs and
//    its "input group" the legs with the smallest input strides; offset
//    tables for both groups are built per tile in LDS, so reads are
//    coalesced along the input and writes along the output.
// and a box copy (tq_copy_plan_create / tq_copy_execute) that moves boxes
// between strided blocks of two buffers: the charge-aware front end's leg
// fusion and its inverse.
// Semantics: out dims[k] = in dims[perm[k]]; out[n] = in[sum_k n_k *
// in_stride[perm[k]]] (src/tensor/dense.h:325-366,
// src/tensor/block_sparse.h:608-651).
#include <algorithm>
#include <memory>
#include <numeric>

#include "tq_common.h"
#include "tq_split.h"

namespace {

constexpr int kGatherThreads = 256;
constexpr int kPerThread = kGatherPerWG / 64;  // 8 elements per lane
// elements of one workgroup of the single-job gather
constexpr int kSingleChunk = kGatherThreads * kPerThread;

// Elements e0 + q * stride + t, q < 8, of a job (t = the thread's index in a
// group of `stride` threads): all source offsets, then all loads in flight
// together, then the stores. Elements past the job skip the index decode (a
// small block leaves most of a chunk, and whole wavefronts, with nothing to
// do).
template <class W>
__device__ inline void gather_range(const CopyJob &job, int64_t e0, int stride, int t,
                                    const W *__restrict__ src,
                                    W *__restrict__ dst) {
  const int r = job.rank;
  auto decode = [&](int64_t e) {
    uint32_t rem = (uint32_t)e;
    i64 o = job.src_off;
    for (int l = 0; l < r; ++l) {
      const uint32_t d = (uint32_t)job.dims[l];
      const uint32_t qd = job_div(rem, d, job.magic[l]);
      o += (i64)(rem - qd * d) * job.strides[l];
      rem = qd;
    }
    return o;
  };
  i64 off[kPerThread];
  if (e0 + (int64_t)kPerThread * stride <= job.n_elem) {
    // a whole chunk: no per-element test between the decodes
#pragma unroll
    for (int q = 0; q < kPerThread; ++q) off[q] = decode(e0 + (int64_t)q * stride + t);
  } else {
#pragma unroll
    for (int q = 0; q < kPerThread; ++q) {
      const int64_t e = e0 + (int64_t)q * stride + t;
      off[q] = e < job.n_elem ? decode(e) : -1;
    }
  }
  W v[kPerThread];
#pragma unroll
  for (int q = 0; q < kPerThread; ++q) v[q] = off[q] >= 0 ? src[off[q]] : W(0);
#pragma unroll
  for (int q = 0; q < kPerThread; ++q) {
    const int64_t e = e0 + (int64_t)q * stride + t;
    if (off[q] >= 0) dst[job.dst_off + e] = v[q];
  }
}

// gather_range for a job with a strided destination: both offsets of each
// element come from one decode of its index.
template <class W>
__device__ inline void gather_range_strided(const CopyJob &job, int64_t e0, int stride, int t,
                                            const W *__restrict__ src,
                                            W *__restrict__ dst) {
  const int r = job.rank;
  i64 so[kPerThread], dof[kPerThread];
#pragma unroll
  for (int q = 0; q < kPerThread; ++q) {
    const int64_t e = e0 + (int64_t)q * stride + t;
    so[q] = -1;
    dof[q] = 0;
    if (e < job.n_elem) {
      uint32_t rem = (uint32_t)e;
      i64 s = job.src_off, d = job.dst_off;
      for (int l = 0; l < r; ++l) {
        const uint32_t n = (uint32_t)job.dims[l];
        const uint32_t qd = job_div(rem, n, job.magic[l]);
        const uint32_t i = rem - qd * n;
        s += (i64)i * job.strides[l];
        d += (i64)i * job.dst_strides[l];
        rem = qd;
      }
      so[q] = s;
      dof[q] = d;
    }
  }
  W v[kPerThread];
#pragma unroll
  for (int q = 0; q < kPerThread; ++q) v[q] = so[q] >= 0 ? src[so[q]] : W(0);
#pragma unroll
  for (int q = 0; q < kPerThread; ++q)
    if (so[q] >= 0) dst[dof[q]] = v[q];
}

// Job of workgroup chunk c: the last j < n with start[j] <= c (start[0] = 0).
// A 64-ary search: each round every lane of the wavefront probes one
// candidate and a ballot keeps the last one at or below c, so a list of n
// jobs costs ceil(log64 n) dependent loads (three for 2^18 jobs) where a
// binary search costs log2 n. Called by whole wavefronts.
__device__ inline int64_t find_job(const int32_t *__restrict__ start, int64_t n, int32_t c) {
  const int lane = threadIdx.x & 63;
  int64_t lo = 0, hi = n;  // the answer lies in [lo, hi); start[lo] <= c
  while (hi - lo > 1) {
    const int64_t step = (hi - lo + 63) >> 6;
    const int64_t j = lo + lane * step;
    const uint64_t m = __ballot(j < hi && start[j] <= c);
    const int64_t nlo = lo + (int64_t)(63 - __clzll(m)) * step;
    hi = min(hi, nlo + step);
    lo = nlo;
  }
  return lo;
}

// Persistent gather over a job list: chunk c (kGatherPerWG elements, one
// wavefront's) of the whole list belongs to job j with job_wg_start[j] <= c <
// job_wg_start[j+1]; the chunk count is read from job_wg_start[n_jobs] so no
// host round trip is needed. A workgroup covers up to four small blocks at
// once, and a large block spreads over many wavefronts.
template <class W>
__global__ __launch_bounds__(kGatherThreads) void gather_jobs_kernel(
    const CopyJob *__restrict__ jobs, const int32_t *__restrict__ job_wg_start,
    int64_t n_jobs, const W *__restrict__ src, W *__restrict__ dst) {
  constexpr int kWaves = kGatherThreads / 64;
  const int32_t total = job_wg_start[n_jobs];
  const int lane = threadIdx.x & 63;
  const int32_t wave = __builtin_amdgcn_readfirstlane((int32_t)(threadIdx.x >> 6));
  for (int32_t c = (int32_t)blockIdx.x * kWaves + wave; c < total;
       c += (int32_t)gridDim.x * kWaves) {
    const int64_t lo = find_job(job_wg_start, n_jobs, c);
    const CopyJob &job = jobs[lo];
    const int64_t e0 = (int64_t)(c - job_wg_start[lo]) * kGatherPerWG;
    if (job.dst_strided)
      gather_range_strided(job, e0, 64, lane, src, dst);
    else
      gather_range(job, e0, 64, lane, src, dst);
  }
}

// ---- box copies ----------------------------------------------------------

// One box of a block copy after the host merged its contiguous legs: element
// e of the box (decoded leg 0 fastest over dims) moves from
// src_off + sum_l i_l * sstr[l] to dst_off + sum_l i_l * dstr[l].
struct BoxJob {
  i64 src_off, dst_off, n_elem;
  int32_t rank, pad;
  uint32_t dims[TQ_MAX_RANK];
  i64 sstr[TQ_MAX_RANK];
  i64 dstr[TQ_MAX_RANK];
};

// Persistent copy over a box list, one kGatherPerWG-element work item per
// wavefront (work item c belongs to box j with job_wg_start[j] <= c <
// job_wg_start[j+1]): both offsets of each element are decoded, then all
// loads are issued, then the stores. The first merged leg is the one that is
// contiguous in the destination (or the source), so a wavefront moves runs
// of consecutive elements.
template <class W>
__global__ __launch_bounds__(kGatherThreads) void box_jobs_kernel(
    const BoxJob *__restrict__ jobs, const int32_t *__restrict__ job_wg_start,
    int64_t n_jobs, const W *__restrict__ src, W *__restrict__ dst) {
  constexpr int kWaves = kGatherThreads / 64;
  const int32_t total = job_wg_start[n_jobs];
  const int lane = threadIdx.x & 63;
  const int32_t wave = __builtin_amdgcn_readfirstlane((int32_t)(threadIdx.x >> 6));
  for (int32_t c = (int32_t)blockIdx.x * kWaves + wave; c < total;
       c += (int32_t)gridDim.x * kWaves) {
    const int64_t j = find_job(job_wg_start, n_jobs, c);
    const BoxJob &job = jobs[j];
    const int64_t e0 = (int64_t)(c - job_wg_start[j]) * kGatherPerWG;
    const int r = job.rank;
    i64 so[kPerThread], dof[kPerThread];
#pragma unroll
    for (int q = 0; q < kPerThread; ++q) {
      const int64_t e = e0 + (int64_t)q * 64 + lane;
      so[q] = -1;
      dof[q] = 0;
      if (e < job.n_elem) {
        uint32_t rem = (uint32_t)e;
        i64 s = job.src_off, d = job.dst_off;
        for (int l = 0; l < r; ++l) {
          const uint32_t n = job.dims[l];
          const uint32_t qd = rem / n;
          const uint32_t i = rem - qd * n;
          s += (i64)i * job.sstr[l];
          d += (i64)i * job.dstr[l];
          rem = qd;
        }
        so[q] = s;
        dof[q] = d;
      }
    }
    W v[kPerThread];
#pragma unroll
    for (int q = 0; q < kPerThread; ++q) v[q] = so[q] >= 0 ? src[so[q]] : W(0);
#pragma unroll
    for (int q = 0; q < kPerThread; ++q)
      if (so[q] >= 0) dst[dof[q]] = v[q];
  }
}

template <class W>
__global__ __launch_bounds__(kGatherThreads) void gather_single_kernel(
    CopyJob job, const W *__restrict__ src, W *__restrict__ dst) {
  const int64_t e0 = (int64_t)blockIdx.x * kSingleChunk;
  gather_range(job, e0, kGatherThreads, (int)threadIdx.x, src, dst);
}

// ---- tiled transpose -----------------------------------------------------

constexpr int kT = 64;  // tile edge

struct TileDesc {
  i64 src_off, dst_off;
  int32_t n_og, n_ig, n_b;       // legs in each group
  int32_t og_dims[TQ_MAX_RANK];  // output legs 0..n_og-1 (contiguous output)
  i64 og_in_str[TQ_MAX_RANK];
  int32_t ig_dims[TQ_MAX_RANK];
  i64 ig_in_str[TQ_MAX_RANK];
  i64 ig_out_str[TQ_MAX_RANK];
  int32_t b_dims[TQ_MAX_RANK];
  i64 b_in_str[TQ_MAX_RANK];
  i64 b_out_str[TQ_MAX_RANK];
  i64 p_og, p_ig;  // extents of the two groups
  int32_t t_og, t_ig;  // tile extents (t_og <= 64, t_ig <= 512, product <= 4096)
  uint32_t m_og, m_ig;  // ceil(2^32 / t): e / t == umulhi(e, m) for e < 2^32 / t
  i64 n_og_tiles, n_ig_tiles;
  i64 n_tiles;
};

constexpr int kTileElems = 4096;  // elements of one tile
constexpr int kMaxTig = 512;

__device__ inline int udiv_magic(uint32_t e, uint32_t m, int t) {
  return t == 1 ? (int)e : (int)__umulhi(e, m);
}

// Tiles of t_og output-group x t_ig input-group elements through LDS. Each
// thread owns 16 elements of the tile, numbered linearly: for the read the
// input-group index runs fastest (lanes read along the input), for the write
// the output-group index does (lanes write along the output), so edge tiles
// and short groups keep every lane busy. All loads of a thread are issued
// before its LDS stores.
template <class W>
__global__ __launch_bounds__(256) void tiled_permute_kernel(
    TileDesc td, const W *__restrict__ src, W *__restrict__ dst) {
  // row r of the tile at r * ld, ld = t_ig + 1 or + 2 (odd: rows on distinct
  // bank pairs for the column-wise reads of the write phase)
  __shared__ W tile[kTileElems + 2 * kT];
  __shared__ i64 in_og[kT];
  __shared__ i64 in_ig[kMaxTig];
  __shared__ i64 out_ig[kMaxTig];
  constexpr int kPer = kTileElems / 256;
  const int tid = threadIdx.x;
  const int ld = td.t_ig + ((td.t_ig & 1) ? 2 : 1);
  for (i64 t = blockIdx.x; t < td.n_tiles; t += gridDim.x) {
    i64 rest = t;
    const i64 c_og = rest % td.n_og_tiles;
    rest /= td.n_og_tiles;
    const i64 c_ig = rest % td.n_ig_tiles;
    i64 bidx = rest / td.n_ig_tiles;
    i64 in_base = td.src_off, out_base = td.dst_off;
    for (int l = 0; l < td.n_b; ++l) {
      const i64 d = td.b_dims[l];
      const i64 q = bidx / d;
      const i64 dg = bidx - q * d;
      in_base += dg * td.b_in_str[l];
      out_base += dg * td.b_out_str[l];
      bidx = q;
    }
    const i64 og0 = c_og * td.t_og, ig0 = c_ig * td.t_ig;
    const int n_og = (int)min<i64>(td.t_og, td.p_og - og0);
    const int n_ig = (int)min<i64>(td.t_ig, td.p_ig - ig0);
    if (tid < n_og) {
      uint32_t og = (uint32_t)(og0 + tid);
      i64 off = 0;
      for (int l = 0; l < td.n_og; ++l) {
        const uint32_t d = (uint32_t)td.og_dims[l];
        const uint32_t q = og / d;
        off += (i64)(og - q * d) * td.og_in_str[l];
        og = q;
      }
      in_og[tid] = off;
    }
    for (int u = tid; u < n_ig; u += 256) {
      // group indices stay below 2^31 (p_ig <= 2^31 is checked on the host)
      uint32_t ig = (uint32_t)(ig0 + u);
      i64 oi = 0, oo = 0;
      for (int l = 0; l < td.n_ig; ++l) {
        const uint32_t d = (uint32_t)td.ig_dims[l];
        const uint32_t q = ig / d;
        const uint32_t dg = ig - q * d;
        oi += (i64)dg * td.ig_in_str[l];
        oo += (i64)dg * td.ig_out_str[l];
        ig = q;
      }
      in_ig[u] = oi;
      out_ig[u] = oo;
    }
    __syncthreads();
    // two halves of 8 loads each in flight (fewer live registers)
#pragma unroll
    for (int half = 0; half < 2; ++half) {
      constexpr int kH = kPer / 2;
      W v[kH];
#pragma unroll
      for (int q = 0; q < kH; ++q) {
        const uint32_t e = (uint32_t)(tid + 256 * (half * kH + q));
        const int r = udiv_magic(e, td.m_ig, td.t_ig);
        const int u = (int)e - r * td.t_ig;
        v[q] = (r < n_og && u < n_ig) ? src[in_base + in_og[r] + in_ig[u]] : W(0);
      }
#pragma unroll
      for (int q = 0; q < kH; ++q) {
        const uint32_t e = (uint32_t)(tid + 256 * (half * kH + q));
        const int r = udiv_magic(e, td.m_ig, td.t_ig);
        const int u = (int)e - r * td.t_ig;
        if (r < n_og && u < n_ig) tile[r * ld + u] = v[q];
      }
    }
    __syncthreads();
#pragma unroll 4
    for (int q = 0; q < kPer; ++q) {
      const uint32_t e = (uint32_t)(tid + 256 * q);
      const int u = udiv_magic(e, td.m_og, td.t_og);
      const int r = (int)e - u * td.t_og;
      if (r < n_og && u < n_ig) dst[out_base + og0 + r + out_ig[u]] = tile[r * ld + u];
    }
    __syncthreads();
  }
}

template <class W>
tq_status launch_gather_single(tq_handle_t h, const CopyJob &job,
                               const void *src, void *dst) {
  const int64_t nwg = ceil_div(job.n_elem, kSingleChunk);
  if (nwg == 0) return TQ_SUCCESS;
  hipLaunchKernelGGL(gather_single_kernel<W>, dim3((unsigned)nwg),
                     dim3(kGatherThreads), 0, h->stream, job,
                     reinterpret_cast<const W *>(src),
                     reinterpret_cast<W *>(dst));
  TQ_HIP(h, hipGetLastError());
  return TQ_SUCCESS;
}


// Effective legs after dropping unit dims and fusing output-adjacent legs
// that are also adjacent (same order, contiguous) in the input.
struct Legs {
  int r = 0;
  i64 d[TQ_MAX_RANK];
  i64 s[TQ_MAX_RANK];  // input stride of each output leg
};

Legs simplify(int rank, const i64 *dims, const i64 *in_str,
              const int32_t *perm) {
  Legs L;
  for (int k = 0; k < rank; ++k) {
    const i64 d = dims[perm[k]];
    if (d == 1) continue;
    const i64 s = in_str[perm[k]];
    if (L.r > 0 && L.s[L.r - 1] * L.d[L.r - 1] == s) {
      L.d[L.r - 1] *= d;
    } else {
      L.d[L.r] = d;
      L.s[L.r] = s;
      ++L.r;
    }
  }
  return L;
}

template <class W>
tq_status dense_permute_t(tq_handle_t h, const Legs &L, i64 n_total,
                          const void *in, void *out) {
  if (n_total == 0) return TQ_SUCCESS;
  if (L.r == 0 || (L.r == 1 && L.s[0] == 1)) {
    TQ_HIP(h, hipMemcpyAsync(out, in, (size_t)n_total * sizeof(W),
                             hipMemcpyDeviceToDevice, h->stream));
    return TQ_SUCCESS;
  }
  // input-fastest output leg
  int f = 0;
  for (int k = 1; k < L.r; ++k)
    if (L.s[k] < L.s[f]) f = k;
  if (f == 0 || n_total < 4096) {
    if (n_total >= (i64(1) << 32))
      return tq_fail(h, TQ_ERROR_UNSUPPORTED, "permute: tensor above 2^32 elements");
    CopyJob job{};
    job.src_off = 0;
    job.dst_off = 0;
    job.n_elem = n_total;
    job.rank = L.r;
    for (int k = 0; k < L.r; ++k) {
      job.dims[k] = (int32_t)L.d[k];
      job.strides[k] = L.s[k];
    }
    finalize_job(job);
    return launch_gather_single<W>(h, job, in, out);
  }
  TileDesc td{};
  i64 ostr[TQ_MAX_RANK];
  ostr[0] = 1;
  for (int k = 1; k < L.r; ++k) ostr[k] = ostr[k - 1] * L.d[k - 1];
  bool used[TQ_MAX_RANK] = {false};
  // output group: leading output legs (stop before the input-fastest leg)
  i64 p_og = 1;
  int a = 0;
  while (a < f && p_og < kT) {
    td.og_dims[td.n_og] = (int32_t)L.d[a];
    td.og_in_str[td.n_og] = L.s[a];
    ++td.n_og;
    p_og *= L.d[a];
    used[a] = true;
    ++a;
  }
  // tile extents: up to 64 along the output group; along the input group as
  // many as fill a tile of kTileElems (at most kMaxTig)
  const int t_og = (int)std::min<i64>(kT, p_og);
  const i64 want_ig = std::min<i64>(kMaxTig, std::max<i64>(kT, kTileElems / t_og));
  // input group: legs by increasing input stride, starting at f
  i64 p_ig = 1;
  while (p_ig < want_ig) {
    int best = -1;
    for (int k = 0; k < L.r; ++k)
      if (!used[k] && (best < 0 || L.s[k] < L.s[best])) best = k;
    if (best < 0) break;
    used[best] = true;
    td.ig_dims[td.n_ig] = (int32_t)L.d[best];
    td.ig_in_str[td.n_ig] = L.s[best];
    td.ig_out_str[td.n_ig] = ostr[best];
    ++td.n_ig;
    p_ig *= L.d[best];
  }
  i64 n_b = 1;
  for (int k = 0; k < L.r; ++k) {
    if (used[k]) continue;
    td.b_dims[td.n_b] = (int32_t)L.d[k];
    td.b_in_str[td.n_b] = L.s[k];
    td.b_out_str[td.n_b] = ostr[k];
    ++td.n_b;
    n_b *= L.d[k];
  }
  if (p_og >= (i64(1) << 31) || p_ig >= (i64(1) << 31))
    return tq_fail(h, TQ_ERROR_UNSUPPORTED, "permute: leg group above 2^31 elements");
  td.p_og = p_og;
  td.p_ig = p_ig;
  td.t_og = t_og;
  td.t_ig = (int)std::min<i64>(p_ig, std::min<i64>(kMaxTig, kTileElems / t_og));
  td.m_og = (uint32_t)(((uint64_t)1 << 32) / (uint64_t)td.t_og +
                       ((((uint64_t)1 << 32) % (uint64_t)td.t_og) ? 1 : 0));
  td.m_ig = (uint32_t)(((uint64_t)1 << 32) / (uint64_t)td.t_ig +
                       ((((uint64_t)1 << 32) % (uint64_t)td.t_ig) ? 1 : 0));
  td.n_og_tiles = ceil_div(p_og, td.t_og);
  td.n_ig_tiles = ceil_div(p_ig, td.t_ig);
  td.n_tiles = td.n_og_tiles * td.n_ig_tiles * n_b;
  const i64 grid = std::min<i64>(td.n_tiles, (i64)h->n_cu * 16);
  hipLaunchKernelGGL(tiled_permute_kernel<W>, dim3((unsigned)grid), dim3(256),
                     0, h->stream, td, reinterpret_cast<const W *>(in),
                     reinterpret_cast<W *>(out));
  TQ_HIP(h, hipGetLastError());
  return TQ_SUCCESS;
}

bool valid_perm(int rank, const int32_t *perm) {
  bool seen[TQ_MAX_RANK] = {false};
  for (int k = 0; k < rank; ++k) {
    if (perm[k] < 0 || perm[k] >= rank || seen[perm[k]]) return false;
    seen[perm[k]] = true;
  }
  return true;
}

// Builds one gather job per block of a block-sparse tensor (device side):
// dims/strides permuted by perm, destination offsets from the scan of sizes.
__global__ void bs_permute_jobs_kernel(int32_t rank, int64_t n_blocks,
                                       const i64 *__restrict__ begin,
                                       const i64 *__restrict__ end,
                                       const i64 *__restrict__ strides,
                                       const i64 *__restrict__ offsets,
                                       const i64 *__restrict__ dst_offsets,
                                       const int32_t *__restrict__ perm,
                                       CopyJob *__restrict__ jobs,
                                       i64 *__restrict__ wg_count) {
  const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= n_blocks) return;
  CopyJob j{};
  j.src_off = offsets[b];
  j.dst_off = dst_offsets[b];
  j.rank = rank;
  i64 n = 1;
  for (int k = 0; k < rank; ++k) {
    const int l = perm[k];
    const i64 d = end[b * rank + l] - begin[b * rank + l] + 1;
    j.dims[k] = (int32_t)d;
    j.strides[k] = strides[b * rank + l];
    n *= d;
  }
  j.n_elem = n;
  finalize_job(j);
  jobs[b] = j;
  wg_count[b] = ceil_div(n, kGatherPerWG);
}

__global__ void block_sizes_kernel(int32_t rank, int64_t n_blocks,
                                   const i64 *__restrict__ begin,
                                   const i64 *__restrict__ end,
                                   i64 *__restrict__ sizes) {
  const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= n_blocks) return;
  i64 n = 1;
  for (int l = 0; l < rank; ++l)
    n *= end[b * rank + l] - begin[b * rank + l] + 1;
  sizes[b] = n;
}

__global__ void i64_to_i32_kernel(const i64 *__restrict__ in,
                                  int32_t *__restrict__ out, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = (int32_t)in[i];
}

}  // namespace

namespace {

// e / t for any 32-bit e, from m = ceil(2^32 / t): umulhi(e, m) exceeds the
// quotient by at most one.
struct Div32 {
  uint32_t t, m;
  __device__ explicit Div32(uint32_t d)
      : t(d), m((uint32_t)((((uint64_t)1 << 32) + d - 1) / d)) {}
  __device__ inline uint32_t div(uint32_t e) const {
    if (t == 1) return e;
    const uint32_t q = __umulhi(e, m);
    return (uint64_t)q * t > e ? q - 1 : q;
  }
};


// Source offset of row `row` over the row legs (nk..rank-1) of a job.
__device__ inline i64 split_row_offset(const CopyJob &job, uint32_t row) {
  i64 o = job.src_off;
#pragma unroll
  for (int l = 0; l < TQ_MAX_RANK; ++l) {
    if (l < job.nk || l >= job.rank) continue;
    const uint32_t d = (uint32_t)job.dims[l];
    const uint32_t q = job_div(row, d, job.magic[l]);
    o += (i64)(row - q * d) * job.strides[l];
    row = q;
  }
  return o;
}

// Split tile of a job whose first contracted leg is contiguous in memory
// (k fastest): two lanes share a row's 8-value chunk, four values each, so
// a wavefront reads 32 rows x 64 contiguous bytes (16-byte loads when the
// four values are adjacent) and writes 512 contiguous bytes per plane (an
// 8-byte word of four BF16 pieces per lane).
__device__ inline void split_kfast_tile(const CopyJob &job, int64_t rows, int64_t row0,
                                        int64_t chunk, const double *__restrict__ src,
                                        uint16_t *__restrict__ dst, int64_t plane,
                                        bool &bad) {
  if (chunk * 8 >= job.kpad) return;
  const int lane = threadIdx.x & 63;
  const int nk = job.nk;
  const int64_t k0 = chunk * 8 + 4 * (lane & 1);
  // offsets of k0 .. k0 + 3 over the contracted legs (-1: past kdim)
  i64 koff[4];
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    koff[e] = -1;
    if (k0 + e < job.kdim) {
      uint32_t rem = (uint32_t)(k0 + e);
      i64 o = 0;
#pragma unroll
      for (int l = 0; l < TQ_MAX_RANK; ++l) {
        if (l >= nk) continue;
        const uint32_t d = (uint32_t)job.dims[l];
        const uint32_t q = job_div(rem, d, job.magic[l]);
        o += (i64)(rem - q * d) * job.strides[l];
        rem = q;
      }
      koff[e] = o;
    }
  }
  // the row offsets of the two row groups, then all loads in flight
  // together, then the splits and stores
  constexpr int kG = kSplitRows / 32;
  i64 roff[kG];
#pragma unroll
  for (int rg = 0; rg < kG; ++rg) {
    const int64_t row = row0 + rg * 32 + (lane >> 1);
    i64 o = job.src_off;
    uint32_t rem = (uint32_t)row;
#pragma unroll
    for (int l = 0; l < TQ_MAX_RANK; ++l) {
      if (l < nk || l >= job.rank) continue;
      const uint32_t d = (uint32_t)job.dims[l];
      const uint32_t q = job_div(rem, d, job.magic[l]);
      o += (i64)(rem - q * d) * job.strides[l];
      rem = q;
    }
    roff[rg] = row < rows ? o : -1;
  }
  double xv[kG][4];
  if (koff[3] == koff[0] + 3 && koff[1] == koff[0] + 1 && koff[2] == koff[0] + 2) {
    // adjacent in memory: two 16-byte loads per row (element aligned only,
    // which the global loads accept)
    typedef double d2u __attribute__((ext_vector_type(2), aligned(8)));
#pragma unroll
    for (int rg = 0; rg < kG; ++rg) {
      d2u v0 = d2u{0.0, 0.0}, v1 = d2u{0.0, 0.0};
      if (roff[rg] >= 0) {
        const d2u *p = reinterpret_cast<const d2u *>(src + roff[rg] + koff[0]);
        v0 = p[0];
        v1 = p[1];
      }
      xv[rg][0] = v0.x;
      xv[rg][1] = v0.y;
      xv[rg][2] = v1.x;
      xv[rg][3] = v1.y;
    }
  } else {
#pragma unroll
    for (int rg = 0; rg < kG; ++rg)
#pragma unroll
      for (int e = 0; e < 4; ++e)
        xv[rg][e] = (roff[rg] >= 0 && koff[e] >= 0) ? src[roff[rg] + koff[e]] : 0.0;
  }
#pragma unroll
  for (int rg = 0; rg < kG; ++rg) {
    const int64_t row = row0 + rg * 32 + (lane >> 1);
    if (row >= rows) continue;
    uint2 wh, wm, wl;
#pragma unroll
    for (int e = 0; e < 4; ++e) bad = bad || tqs::split_out_of_range(xv[rg][e]);
    tqs::split3x2(xv[rg][0], xv[rg][1], wh.x, wm.x, wl.x);
    tqs::split3x2(xv[rg][2], xv[rg][3], wh.y, wm.y, wl.y);
    const int64_t o = job.dst_off + chunk * rows * 8 + row * 8 + 4 * (lane & 1);
    *reinterpret_cast<uint2 *>(dst + o) = wh;
    *reinterpret_cast<uint2 *>(dst + plane + o) = wm;
    *reinterpret_cast<uint2 *>(dst + 2 * plane + o) = wl;
  }
}

// Workgroup index for a grid-stride loop in which each XCD walks a
// contiguous stretch of the work items: workgroups are dealt to the 8 XCDs
// round-robin, so neighbouring items (which share cache lines) would
// otherwise land in different L2s.
__device__ inline int32_t xcd_local_id() {
  const uint32_t g = gridDim.x, b = blockIdx.x;
  if (g < 8 || g % 8 != 0) return (int32_t)b;
  return (int32_t)((b % 8) * (g / 8) + b / 8);
}

// Split-BF16 pack, one work item per wavefront. The items of a job are its
// (row, 8-value chunk) cells, numbered rows fastest (cell = chunk * rows +
// row), kSplitCells per item: one per lane, which reads the cell's 8 values
// (lanes run along the rows, so the reads coalesce when a row leg is fastest
// in memory), splits them and writes one 16-byte vector to each plane (a
// wavefront writes 1 KiB of contiguous output per plane). When the first
// contracted leg is contiguous an item is a unit of kSplitRows rows x one
// chunk instead (split_kfast_tile), units numbered chunks fastest, so that
// the chunks of a row band share workgroups and XCDs. Values past kdim are
// zeros. Numbering the work across the whole job keeps every lane busy for
// jobs of few rows or few chunks.
__global__ __launch_bounds__(kGatherThreads) void split_jobs_kernel(
    const CopyJob *__restrict__ jobs, const int32_t *__restrict__ job_wg_start,
    int64_t n_jobs, const double *__restrict__ src, uint16_t *__restrict__ dst,
    int64_t plane, int32_t *__restrict__ range_flag) {
  constexpr int kWaves = kGatherThreads / 64;
  const int32_t total = job_wg_start[n_jobs];
  const int lane = threadIdx.x & 63;
  bool bad = false;  // a value outside the split's domain (flagged at the end)
  const int32_t wave = __builtin_amdgcn_readfirstlane((int32_t)(threadIdx.x >> 6));
  for (int32_t c = xcd_local_id() * kWaves + wave; c < total;
       c += (int32_t)gridDim.x * kWaves) {
    const int64_t lo = find_job(job_wg_start, n_jobs, c);
    const CopyJob &job = jobs[lo];
    const int nk = job.nk;
    const int64_t kd = job.kdim;
    // a split job holds at most 2^31 values (the planner packs no larger
    // panel), so its cells index in 32 bits
    const uint32_t rows = (uint32_t)(job.n_elem / kd);
    const uint32_t cells = rows * (uint32_t)(job.kpad >> 3);
    const uint32_t tile = (uint32_t)(c - job_wg_start[lo]);
    if (split_kfast(job)) {
      // unit = (row band, chunk), chunks fastest
      const uint32_t chunks = (uint32_t)(job.kpad >> 3);
      const uint32_t band = tile / chunks;
      split_kfast_tile(job, rows, (int64_t)band * kSplitRows, tile - band * chunks, src, dst,
                       plane, bad);
      continue;
    }
    const Div32 rdiv(rows);
    const uint32_t cell = tile * kSplitCells + lane;
    if (cell >= cells) continue;  // no barrier in this loop
    const uint32_t chunk = rdiv.div(cell), row = cell - chunk * rows;
    // source offset of (row, k = 8 * chunk): rows decode over legs nk.., k
    // over legs 0..nk-1 (leg 0 fastest); k digits advance with carries
    i64 off = split_row_offset(job, row);
    int32_t dig[TQ_MAX_RANK];
    uint32_t rem = chunk * 8;
#pragma unroll
    for (int l = 0; l < TQ_MAX_RANK; ++l) {
      dig[l] = 0;
      if (l >= nk) continue;
      const uint32_t d = (uint32_t)job.dims[l];
      const uint32_t q = job_div(rem, d, job.magic[l]);
      dig[l] = (int32_t)(rem - q * d);
      off += (i64)dig[l] * job.strides[l];
      rem = q;
    }
    // the 8 source offsets first (carrying through the contracted legs), then
    // all 8 loads in flight together, then the splits
    i64 offs[8];
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      offs[e] = off;
      bool carry = true;
#pragma unroll
      for (int l = 0; l < TQ_MAX_RANK; ++l) {
        if (!carry || l >= nk) continue;
        off += job.strides[l];
        if (++dig[l] < job.dims[l]) {
          carry = false;
        } else {
          off -= (i64)job.dims[l] * job.strides[l];
          dig[l] = 0;
        }
      }
    }
    double xv[8];
#pragma unroll
    for (int e = 0; e < 8; ++e) xv[e] = ((int64_t)chunk * 8 + e < kd) ? src[offs[e]] : 0.0;
    const int64_t o = job.dst_off + ((int64_t)chunk * rows + row) * 8;
#pragma unroll
    for (int e = 0; e < 8; ++e) bad = bad || tqs::split_out_of_range(xv[e]);
    uint4 vh, vm, vw;
    tqs::split3x2(xv[0], xv[1], vh.x, vm.x, vw.x);
    tqs::split3x2(xv[2], xv[3], vh.y, vm.y, vw.y);
    tqs::split3x2(xv[4], xv[5], vh.z, vm.z, vw.z);
    tqs::split3x2(xv[6], xv[7], vh.w, vm.w, vw.w);
    *reinterpret_cast<uint4 *>(dst + o) = vh;
    *reinterpret_cast<uint4 *>(dst + plane + o) = vm;
    *reinterpret_cast<uint4 *>(dst + 2 * plane + o) = vw;
  }
  if (bad && range_flag) atomicOr(range_flag, 1);
}

}  // namespace

tq_status tq_launch_split_jobs(tq_handle_t h, const CopyJob *jobs,
                               const int32_t *job_wg_start, int64_t n_jobs,
                               int64_t total_wg, const double *src,
                               uint16_t *dst, int64_t plane, int32_t *range_flag) {
  if (n_jobs <= 0) return TQ_SUCCESS;
  // one work item per wavefront, four wavefronts per workgroup
  i64 grid = std::max<i64>(1, std::min<i64>(ceil_div(total_wg, kGatherThreads / 64),
                                            (i64)h->n_cu * 8));
  hipLaunchKernelGGL(split_jobs_kernel, dim3((unsigned)grid), dim3(kGatherThreads), 0,
                     h->stream, jobs, job_wg_start, n_jobs, src, dst, plane, range_flag);
  TQ_HIP(h, hipGetLastError());
  return TQ_SUCCESS;
}

tq_status tq_launch_gather_jobs(tq_handle_t h, tq_dtype dt, const CopyJob *jobs,
                                const int32_t *job_wg_start, int64_t n_jobs,
                                int64_t total_wg, const void *src, void *dst) {
  if (n_jobs <= 0) return TQ_SUCCESS;
  // one chunk per wavefront, four wavefronts per workgroup
  i64 grid = std::max<i64>(1, std::min<i64>(ceil_div(total_wg, kGatherThreads / 64),
                                            (i64)h->n_cu * 8));
  if (dt == TQ_F64)
    hipLaunchKernelGGL(gather_jobs_kernel<uint64_t>, dim3((unsigned)grid),
                       dim3(kGatherThreads), 0, h->stream, jobs, job_wg_start,
                       n_jobs, reinterpret_cast<const uint64_t *>(src),
                       reinterpret_cast<uint64_t *>(dst));
  else if (dt == TQ_F16)
    hipLaunchKernelGGL(gather_jobs_kernel<uint16_t>, dim3((unsigned)grid),
                       dim3(kGatherThreads), 0, h->stream, jobs, job_wg_start,
                       n_jobs, reinterpret_cast<const uint16_t *>(src),
                       reinterpret_cast<uint16_t *>(dst));
  else
    hipLaunchKernelGGL(gather_jobs_kernel<uint32_t>, dim3((unsigned)grid),
                       dim3(kGatherThreads), 0, h->stream, jobs, job_wg_start,
                       n_jobs, reinterpret_cast<const uint32_t *>(src),
                       reinterpret_cast<uint32_t *>(dst));
  TQ_HIP(h, hipGetLastError());
  return TQ_SUCCESS;
}

extern "C" {

tq_status tq_permute(tq_handle_t h, tq_dtype dtype, int32_t rank,
                     const int64_t *dims, const int64_t *in_strides,
                     const int32_t *perm, const void *in, void *out) {
  if (!h) return TQ_ERROR_NOT_INITIALIZED;
  if (rank < 0 || rank > TQ_MAX_RANK)
    return tq_fail(h, TQ_ERROR_UNSUPPORTED, "permute: rank above TQ_MAX_RANK");
  if (!tq_valid_dtype(dtype)) return TQ_ERROR_UNSUPPORTED;
  if ((rank > 0 && (!dims || !perm)) || !in || !out)
    return TQ_ERROR_INVALID_VALUE;
  if (!valid_perm(rank, perm))
    return tq_fail(h, TQ_ERROR_RANK_MISMATCH,
                   "The number of permutation does not match the rank of tensor");
  i64 table[TQ_MAX_RANK];
  i64 n_total = 1;
  for (int l = 0; l < rank; ++l) {
    if (dims[l] < 0) return TQ_ERROR_INVALID_VALUE;
    table[l] = n_total;
    n_total *= dims[l];
  }
  const i64 *s = in_strides ? in_strides : table;
  Legs L = simplify(rank, dims, s, perm);
  if (dtype == TQ_F64) return dense_permute_t<uint64_t>(h, L, n_total, in, out);
  if (dtype == TQ_F16) return dense_permute_t<uint16_t>(h, L, n_total, in, out);
  return dense_permute_t<uint32_t>(h, L, n_total, in, out);
}

tq_status tq_bs_permute(tq_handle_t h, tq_dtype dtype, const tq_bs_desc *d,
                        const int32_t *perm, const void *in, void *out,
                        int64_t *out_offsets) {
  if (!h) return TQ_ERROR_NOT_INITIALIZED;
  if (!d) return TQ_ERROR_INVALID_VALUE;
  const int r = d->rank;
  if (r < 0 || r > TQ_MAX_RANK) return TQ_ERROR_UNSUPPORTED;
  if (!tq_valid_dtype(dtype)) return TQ_ERROR_UNSUPPORTED;
  if (r > 0 && !perm) return TQ_ERROR_INVALID_VALUE;
  if (!valid_perm(r, perm))
    return tq_fail(h, TQ_ERROR_RANK_MISMATCH,
                   "The number of permutation does not match the rank of tensor");
  const i64 nb = d->n_blocks;
  // host-side offsets of the output blocks (nest_sum of the block sizes,
  // src/gpu/block_sparse.cuh:1113-1117 via the constructor at :189-230)
  if (out_offsets) {
    i64 acc = 0;
    for (i64 b = 0; b < nb; ++b) {
      i64 n = 1;
      for (int l = 0; l < r; ++l) n *= d->end[b * r + l] - d->begin[b * r + l] + 1;
      out_offsets[b] = acc;
      acc += n;
    }
  }
  if (nb == 0) return TQ_SUCCESS;
  if (!in || !out || !d->begin || !d->end || !d->strides || !d->offsets)
    return TQ_ERROR_INVALID_VALUE;
  if (r == 0) {
    // rank-0 blocks hold one element each
    TQ_HIP(h, hipMemcpyAsync(out, in, (size_t)nb * tq_elem_size(dtype),
                             hipMemcpyDeviceToDevice, h->stream));
    return TQ_SUCCESS;
  }
  // sizes are checked before anything is uploaded or launched; the work-item
  // total also bounds the grid and must fit the gather's int32 prefix
  i64 total_elem = 0;
  for (i64 b = 0; b < nb; ++b) {
    i64 n = 1;
    for (int l = 0; l < r; ++l) n *= d->end[b * r + l] - d->begin[b * r + l] + 1;
    if (n >= (i64(1) << 32))
      return tq_fail(h, TQ_ERROR_UNSUPPORTED, "permute: block above 2^32 elements");
    total_elem += ceil_div(n, kGatherPerWG);
  }
  if (total_elem > 0x7fffffff)
    return tq_fail(h, TQ_ERROR_UNSUPPORTED, "permute: too many work items");
  // one upload: begin, end, strides (nb*r each), offsets (nb), perm (r)
  const size_t meta = (size_t)nb * r * 3 + nb;
  const size_t bytes = meta * 8 + 64;
  void *host = nullptr;
  TQ_CHECK(tq_stage(h, bytes, &host));
  i64 *hp = reinterpret_cast<i64 *>(host);
  std::copy(d->begin, d->begin + nb * r, hp);
  std::copy(d->end, d->end + nb * r, hp + nb * r);
  std::copy(d->strides, d->strides + nb * r, hp + 2 * nb * r);
  std::copy(d->offsets, d->offsets + nb, hp + 3 * nb * r);
  int32_t *hperm = reinterpret_cast<int32_t *>(hp + meta);
  std::copy(perm, perm + r, hperm);
  DevBuf dmeta, djobs, dsz, dwg, dwg32, dtmp;
  TQ_HIP(h, dmeta.alloc(bytes));
  TQ_HIP(h, hipMemcpyAsync(dmeta.p, host, bytes, hipMemcpyHostToDevice, h->stream));
  i64 *dbeg = dmeta.as<i64>(), *dend = dbeg + nb * r, *dstr = dend + nb * r,
      *doff = dstr + nb * r;
  const int32_t *dperm = reinterpret_cast<const int32_t *>(doff + nb);
  TQ_HIP(h, djobs.alloc(sizeof(CopyJob) * nb));
  TQ_HIP(h, dsz.alloc(sizeof(i64) * (nb + 1)));
  TQ_HIP(h, dwg.alloc(sizeof(i64) * (nb + 1)));
  TQ_HIP(h, dwg32.alloc(sizeof(int32_t) * (nb + 1)));
  TQ_HIP(h, dtmp.alloc(tq_scan_tmp_bytes(nb + 1)));
  const unsigned g = (unsigned)ceil_div(nb, 256);
  hipLaunchKernelGGL(block_sizes_kernel, dim3(g), dim3(256), 0, h->stream, r, nb,
                     dbeg, dend, dsz.as<i64>());
  TQ_CHECK(tq_scan_i64(h, dsz.as<i64>(), dsz.as<i64>(), nb, nullptr, dtmp.p));
  hipLaunchKernelGGL(bs_permute_jobs_kernel, dim3(g), dim3(256), 0, h->stream, r,
                     nb, dbeg, dend, dstr, doff, dsz.as<i64>(), dperm,
                     djobs.as<CopyJob>(), dwg.as<i64>());
  TQ_CHECK(tq_scan_i64(h, dwg.as<i64>(), dwg.as<i64>(), nb,
                       dwg.as<i64>() + nb, dtmp.p));
  hipLaunchKernelGGL(i64_to_i32_kernel, dim3((unsigned)ceil_div(nb + 1, 256)),
                     dim3(256), 0, h->stream, dwg.as<i64>(), dwg32.as<int32_t>(),
                     nb + 1);
  TQ_HIP(h, hipGetLastError());
  TQ_CHECK(tq_launch_gather_jobs(h, dtype, djobs.as<CopyJob>(),
                                 dwg32.as<int32_t>(), nb, total_elem, in, out));
  // the temporaries are freed on return: wait for the kernels using them
  TQ_HIP(h, hipStreamSynchronize(h->stream));
  return TQ_SUCCESS;
}

tq_status tq_bs_slice(tq_handle_t h, tq_dtype dtype, const tq_bs_desc *d,
                      const int64_t *divisor, const void *in, void *out,
                      int64_t *out_begin, int64_t *out_end) {
  if (!h) return TQ_ERROR_NOT_INITIALIZED;
  if (!d || !divisor) return TQ_ERROR_INVALID_VALUE;
  const int r = d->rank;
  if (r < 1 || r > TQ_MAX_RANK) return TQ_ERROR_UNSUPPORTED;
  if (!tq_valid_dtype(dtype)) return TQ_ERROR_UNSUPPORTED;
  const i64 nb = d->n_blocks;
  if (nb > 0 && (!d->begin || !d->end || !d->strides || !d->offsets)) return TQ_ERROR_INVALID_VALUE;
  i64 n_sub = 1;
  for (int l = 0; l < r; ++l) {
    if (divisor[l] < 1) return TQ_ERROR_INVALID_VALUE;
    n_sub *= divisor[l];
  }
  // sub-block extents; every block extent must divide evenly
  // (src/gpu/block_sparse.cuh:1126-1141)
  for (i64 b = 0; b < nb; ++b)
    for (int l = 0; l < r; ++l)
      if ((d->end[b * r + l] - d->begin[b * r + l] + 1) % divisor[l] != 0)
        return tq_fail(h, TQ_ERROR_INVALID_VALUE, "current slice does not support residues");
  const i64 n_jobs = nb * n_sub;
  if (n_jobs == 0) return TQ_SUCCESS;
  if (!in || !out) return TQ_ERROR_INVALID_VALUE;
  // one job per sub-block: sub-block j of block b (j decoded leg 0 fastest
  // over the divisor, :1155-1163) is copied compact, in order, from its
  // region of the parent block (read with the parent's index table)
  const size_t job_bytes = sizeof(CopyJob) * n_jobs;
  const size_t wg_off = (job_bytes + 15) / 16 * 16;
  const size_t bytes = wg_off + sizeof(int32_t) * (n_jobs + 1);
  void *host = nullptr;
  TQ_CHECK(tq_stage(h, bytes, &host));
  CopyJob *jobs = reinterpret_cast<CopyJob *>(host);
  int32_t *wg = reinterpret_cast<int32_t *>(reinterpret_cast<char *>(host) + wg_off);
  i64 dst = 0, wg_total = 0;
  for (i64 b = 0; b < nb; ++b) {
    i64 sub[TQ_MAX_RANK];
    for (int l = 0; l < r; ++l) sub[l] = (d->end[b * r + l] - d->begin[b * r + l] + 1) / divisor[l];
    for (i64 j = 0; j < n_sub; ++j) {
      CopyJob &jb = jobs[b * n_sub + j];
      jb = CopyJob{};
      i64 rem = j, src = d->offsets[b], n = 1;
      for (int l = 0; l < r; ++l) {
        const i64 q = rem % divisor[l];
        rem /= divisor[l];
        const i64 lo = q * sub[l];
        src += lo * d->strides[b * r + l];
        if (out_begin) out_begin[(b * n_sub + j) * r + l] = d->begin[b * r + l] + lo;
        if (out_end) out_end[(b * n_sub + j) * r + l] = d->begin[b * r + l] + lo + sub[l] - 1;
        jb.dims[l] = (int32_t)sub[l];
        jb.strides[l] = d->strides[b * r + l];
        n *= sub[l];
      }
      if (n >= (i64(1) << 31))
        return tq_fail(h, TQ_ERROR_UNSUPPORTED, "slice: sub-block above 2^31 elements");
      jb.rank = r;
      jb.src_off = src;
      jb.dst_off = dst;
      jb.n_elem = n;
      finalize_job(jb);
      wg[b * n_sub + j] = (int32_t)wg_total;
      wg_total += ceil_div(n, kGatherPerWG);
      dst += n;
    }
  }
  if (wg_total > 0x7fffffff) return tq_fail(h, TQ_ERROR_UNSUPPORTED, "slice: too many elements");
  wg[n_jobs] = (int32_t)wg_total;
  DevBuf dmeta;
  TQ_HIP(h, dmeta.alloc(bytes));
  TQ_HIP(h, hipMemcpyAsync(dmeta.p, host, bytes, hipMemcpyHostToDevice, h->stream));
  TQ_CHECK(tq_launch_gather_jobs(h, dtype, dmeta.as<CopyJob>(),
                                 reinterpret_cast<const int32_t *>(dmeta.as<char>() + wg_off),
                                 n_jobs, wg_total, in, out));
  TQ_HIP(h, hipStreamSynchronize(h->stream));
  return TQ_SUCCESS;
}

}  // extern "C"

// ---- box copy plans ------------------------------------------------------

struct tq_copy_plan_s {
  tq_handle_t h = nullptr;
  tq_dtype dtype = TQ_F64;
  int64_t n_jobs = 0;
  int64_t total_wg = 0;
  size_t wg_off = 0;  // byte offset of the int32 work-item starts in meta
  DevBuf meta;        // BoxJob[n_jobs], then job_wg_start[n_jobs + 1]
};

namespace {

// Largest offset a box reaches from `off` (its extents all >= 1), or -1
// when it does not fit in int64.
i64 box_last(i64 off, int r, const int64_t *dims, const int64_t *str) {
  __int128 last = off;
  for (int l = 0; l < r; ++l) last += (__int128)(dims[l] - 1) * str[l];
  return last > (__int128)INT64_MAX ? -1 : (i64)last;
}

}  // namespace

extern "C" {

tq_status tq_copy_plan_create(tq_handle_t h, tq_dtype dtype, int64_t n, const tq_box *boxes,
                              int64_t in_elems, int64_t out_elems, tq_copy_plan_t *plan) {
  if (!h) return TQ_ERROR_NOT_INITIALIZED;
  if (!plan || n < 0 || (n > 0 && !boxes) || in_elems < 0 || out_elems < 0)
    return TQ_ERROR_INVALID_VALUE;
  if (!tq_valid_dtype(dtype)) return TQ_ERROR_UNSUPPORTED;
  *plan = nullptr;
  // every box is checked and its legs merged before anything is uploaded
  std::vector<BoxJob> jobs;
  jobs.reserve((size_t)n);
  std::vector<int32_t> wg;
  wg.reserve((size_t)n + 1);
  i64 wg_total = 0;
  for (int64_t b = 0; b < n; ++b) {
    const tq_box &x = boxes[b];
    if (x.rank < 0 || x.rank > TQ_MAX_RANK)
      return tq_fail(h, TQ_ERROR_UNSUPPORTED, "copy: box rank above TQ_MAX_RANK");
    if (x.src_off < 0 || x.dst_off < 0)
      return tq_fail(h, TQ_ERROR_INVALID_VALUE, "copy: negative box offset");
    bool empty = false;
    i64 n_elem = 1;
    for (int l = 0; l < x.rank; ++l) {
      if (x.dims[l] < 0 || x.src_strides[l] < 0 || x.dst_strides[l] < 0)
        return tq_fail(h, TQ_ERROR_INVALID_VALUE, "copy: negative box extent or stride");
      if (x.dims[l] == 0) empty = true;
      else if (n_elem < (i64(1) << 40)) n_elem *= x.dims[l];
    }
    if (empty) continue;
    if (n_elem >= (i64(1) << 32))
      return tq_fail(h, TQ_ERROR_UNSUPPORTED, "copy: box above 2^32 elements");
    const i64 ls = box_last(x.src_off, x.rank, x.dims, x.src_strides);
    const i64 ld = box_last(x.dst_off, x.rank, x.dims, x.dst_strides);
    if (ls < 0 || ls >= in_elems || ld < 0 || ld >= out_elems)
      return tq_fail(h, TQ_ERROR_OUT_OF_RANGE, "copy: box outside its buffer");
    // merge leg l into the previous kept leg when it continues it in both
    // buffers; unit legs drop out
    BoxJob j{};
    j.src_off = x.src_off;
    j.dst_off = x.dst_off;
    j.n_elem = n_elem;
    int r = 0;
    for (int l = 0; l < x.rank; ++l) {
      if (x.dims[l] == 1) continue;
      if (r > 0 && x.src_strides[l] == j.sstr[r - 1] * (i64)j.dims[r - 1] &&
          x.dst_strides[l] == j.dstr[r - 1] * (i64)j.dims[r - 1]) {
        j.dims[r - 1] *= (uint32_t)x.dims[l];
        continue;
      }
      j.dims[r] = (uint32_t)x.dims[l];
      j.sstr[r] = x.src_strides[l];
      j.dstr[r] = x.dst_strides[l];
      ++r;
    }
    j.rank = r;
    jobs.push_back(j);
    wg.push_back((int32_t)wg_total);
    wg_total += ceil_div(n_elem, kGatherPerWG);
    if (wg_total > 0x7fffffff)
      return tq_fail(h, TQ_ERROR_UNSUPPORTED, "copy: too many elements");
  }
  wg.push_back((int32_t)wg_total);
  std::unique_ptr<tq_copy_plan_s> p(new tq_copy_plan_s);
  p->h = h;
  p->dtype = dtype;
  p->n_jobs = (int64_t)jobs.size();
  p->total_wg = wg_total;
  if (p->n_jobs > 0) {
    const size_t job_bytes = sizeof(BoxJob) * jobs.size();
    p->wg_off = (job_bytes + 15) / 16 * 16;
    const size_t bytes = p->wg_off + sizeof(int32_t) * wg.size();
    void *host = nullptr;
    TQ_CHECK(tq_stage(h, bytes, &host));
    std::copy(jobs.begin(), jobs.end(), reinterpret_cast<BoxJob *>(host));
    std::copy(wg.begin(), wg.end(),
              reinterpret_cast<int32_t *>(reinterpret_cast<char *>(host) + p->wg_off));
    TQ_HIP(h, p->meta.alloc(bytes));
    TQ_HIP(h, hipMemcpyAsync(p->meta.p, host, bytes, hipMemcpyHostToDevice, h->stream));
    // the staging buffer is reused by the next upload
    TQ_HIP(h, hipStreamSynchronize(h->stream));
  }
  *plan = p.release();
  return TQ_SUCCESS;
}

tq_status tq_copy_execute(tq_copy_plan_t p, const void *in, void *out) {
  if (!p) return TQ_ERROR_NOT_INITIALIZED;
  if (p->n_jobs == 0) return TQ_SUCCESS;
  tq_handle_t h = p->h;
  if (!in || !out) return TQ_ERROR_INVALID_VALUE;
  const BoxJob *jobs = p->meta.as<BoxJob>();
  const int32_t *wg = reinterpret_cast<const int32_t *>(p->meta.as<char>() + p->wg_off);
  // one work item per wavefront, four wavefronts per workgroup
  const i64 grid = std::max<i64>(
      1, std::min<i64>(ceil_div(p->total_wg, kGatherThreads / 64), (i64)h->n_cu * 8));
  if (p->dtype == TQ_F64)
    hipLaunchKernelGGL(box_jobs_kernel<uint64_t>, dim3((unsigned)grid), dim3(kGatherThreads), 0,
                       h->stream, jobs, wg, p->n_jobs, static_cast<const uint64_t *>(in),
                       static_cast<uint64_t *>(out));
  else if (p->dtype == TQ_F16)
    hipLaunchKernelGGL(box_jobs_kernel<uint16_t>, dim3((unsigned)grid), dim3(kGatherThreads), 0,
                       h->stream, jobs, wg, p->n_jobs, static_cast<const uint16_t *>(in),
                       static_cast<uint16_t *>(out));
  else
    hipLaunchKernelGGL(box_jobs_kernel<uint32_t>, dim3((unsigned)grid), dim3(kGatherThreads), 0,
                       h->stream, jobs, wg, p->n_jobs, static_cast<const uint32_t *>(in),
                       static_cast<uint32_t *>(out));
  TQ_HIP(h, hipGetLastError());
  return TQ_SUCCESS;
}

tq_status tq_copy_plan_destroy(tq_copy_plan_t p) {
  delete p;
  return TQ_SUCCESS;
}

}  // extern "C"
