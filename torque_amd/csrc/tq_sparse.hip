// tq_sparse.hip -- element-sparse tensors on the device: contraction
// (replaces SparseTensor<T>::contract, src/gpu/sparse.cuh:270-375, with its
// handle_indices kernel src/gpu/sparse.cu:88-194 and the thrust sort_by_key /
// reduce_by_key at src/gpu/sparse.cuh:356-371) and the index re-encoding of
// hard_transpose (src/gpu/sparse.cuh:398-432, src/tensor/sparse.h:371-405).
//
// A stored element is (linear index under the tensor's index table, value).
// The reference forms all nnz_A x nnz_B products (a k = 1 GEMM), tags the
// pairs whose contracted coordinates differ with key -1, then sorts and
// reduces every product. Here only matching products are formed:
//   1. decode: every element's coordinates give a contracted key (linear
//      index over the contracted legs) and a free key (its share of the
//      output linear index);
//   2. B is sorted by contracted key (stable LSD radix sort, 8-bit digits,
//      equal digits ranked inside a wavefront by ballots);
//   3. each A element finds its run of equal contracted keys in sorted B by
//      binary search; an exclusive scan of the run lengths places its
//      products;
//   4. products are formed (key = free key A + free key B), sorted by output
//      key (stable, so equal keys keep A-major order) and each run of equal
//      keys is summed in that order by one thread, in FP64.
// The output lists every output index that receives a product, ascending
// (the reference's reduce_by_key output without the non-matching key, which
// the reference drops by position, src/gpu/sparse.cuh:370-371). A rank-0
// result is the one element {0: sum}, summed by the fixed-order FP64
// reduction of tq_reduce.hip.
#include <algorithm>
#include <vector>

#include "tq_common.h"

namespace {

constexpr int kRadix = 256;
constexpr int kChunk = 2048;  // items per workgroup of a radix pass

__device__ inline uint64_t lanes_below(int lane) {
  return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}

__global__ __launch_bounds__(256) void radix_hist_kernel(const u64 *__restrict__ keys,
                                                         int64_t n, int shift,
                                                         int64_t n_chunks,
                                                         i64 *__restrict__ hist) {
  __shared__ int h[kRadix];
  h[threadIdx.x] = 0;
  __syncthreads();
  const int64_t base = (int64_t)blockIdx.x * kChunk;
  for (int q = threadIdx.x; q < kChunk; q += 256) {
    const int64_t i = base + q;
    if (i < n) atomicAdd(&h[(keys[i] >> shift) & (kRadix - 1)], 1);
  }
  __syncthreads();
  hist[(int64_t)threadIdx.x * n_chunks + blockIdx.x] = h[threadIdx.x];
}

// Stable scatter of one chunk: items are placed in chunk order; within each
// round of 256 items a wavefront finds the lanes sharing its digit from eight
// ballots and counts the earlier ones.
__global__ __launch_bounds__(256) void radix_scatter_kernel(
    const u64 *__restrict__ kin, const i64 *__restrict__ vin, int64_t n, int shift,
    int64_t n_chunks, const i64 *__restrict__ offs, u64 *__restrict__ kout,
    i64 *__restrict__ vout) {
  __shared__ i64 running[kRadix];
  __shared__ int wave_cnt[4][kRadix];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  running[threadIdx.x] = offs[(int64_t)threadIdx.x * n_chunks + blockIdx.x];
  const int64_t base = (int64_t)blockIdx.x * kChunk;
  for (int round = 0; round < kChunk / 256; ++round) {
    for (int q = threadIdx.x; q < 4 * kRadix; q += 256) (&wave_cnt[0][0])[q] = 0;
    __syncthreads();
    const int64_t idx = base + round * 256 + threadIdx.x;
    const bool valid = idx < n;
    const u64 k = valid ? kin[idx] : 0;
    const i64 v = valid ? vin[idx] : 0;
    const int digit = (int)((k >> shift) & (kRadix - 1));
    uint64_t peers = __ballot(valid);
#pragma unroll
    for (int bit = 0; bit < 8; ++bit) {
      const uint64_t bm = __ballot((digit >> bit) & 1);
      peers &= ((digit >> bit) & 1) ? bm : ~bm;
    }
    const int rank_in_wave = __popcll(peers & lanes_below(lane));
    const int wave_count = __popcll(peers);
    if (valid && rank_in_wave == wave_count - 1) wave_cnt[wave][digit] = wave_count;
    __syncthreads();
    if (valid) {
      i64 off = running[digit] + rank_in_wave;
      for (int w = 0; w < wave; ++w) off += wave_cnt[w][digit];
      kout[off] = k;
      vout[off] = v;
    }
    __syncthreads();
    int t = 0;
#pragma unroll
    for (int w = 0; w < 4; ++w) t += wave_cnt[w][threadIdx.x];
    running[threadIdx.x] += t;
    __syncthreads();
  }
}

__global__ void iota_kernel(i64 *__restrict__ v, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) v[i] = i;
}

unsigned grid_for(int64_t n) { return (unsigned)std::max<int64_t>(1, ceil_div(n, 256)); }

// Scratch of one radix sort of n items.
struct SortTmp {
  DevBuf k2, v2, hist, scan;
  tq_status alloc(tq_handle_t h, int64_t n) {
    const int64_t chunks = std::max<int64_t>(1, ceil_div(n, kChunk));
    TQ_HIP(h, k2.alloc(sizeof(u64) * std::max<int64_t>(n, 1)));
    TQ_HIP(h, v2.alloc(sizeof(i64) * std::max<int64_t>(n, 1)));
    TQ_HIP(h, hist.alloc(sizeof(i64) * (kRadix * chunks + 1)));
    TQ_HIP(h, scan.alloc(tq_scan_tmp_bytes(kRadix * chunks + 1)));
    return TQ_SUCCESS;
  }
};

int bits_for(u64 max_key) {
  int b = 0;
  while (b < 64 && (max_key >> b) != 0) ++b;
  return b;
}

// Stable sort of (keys, vals) by the low `bits` bits of the keys; the result
// ends in (*keys, *vals) (the pointers are swapped with the scratch buffers
// after an odd number of passes).
tq_status radix_sort(tq_handle_t h, u64 **keys, i64 **vals, int64_t n, int bits,
                     SortTmp &tmp) {
  if (n <= 1 || bits <= 0) return TQ_SUCCESS;
  const int64_t chunks = ceil_div(n, kChunk);
  if (chunks > 0x7fffffff) return tq_fail(h, TQ_ERROR_UNSUPPORTED, "sort: too many items");
  u64 *ka = *keys, *kb = tmp.k2.as<u64>();
  i64 *va = *vals, *vb = tmp.v2.as<i64>();
  for (int shift = 0; shift < bits; shift += 8) {
    hipLaunchKernelGGL(radix_hist_kernel, dim3((unsigned)chunks), dim3(256), 0, h->stream,
                       ka, n, shift, chunks, tmp.hist.as<i64>());
    TQ_CHECK(tq_scan_i64(h, tmp.hist.as<i64>(), tmp.hist.as<i64>(), kRadix * chunks, nullptr,
                         tmp.scan.p));
    hipLaunchKernelGGL(radix_scatter_kernel, dim3((unsigned)chunks), dim3(256), 0, h->stream,
                       ka, va, n, shift, chunks, tmp.hist.as<i64>(), kb, vb);
    TQ_HIP(h, hipGetLastError());
    std::swap(ka, kb);
    std::swap(va, vb);
  }
  if (ka != *keys) {
    // odd pass count: copy back so the caller's buffers hold the result
    TQ_HIP(h, hipMemcpyAsync(*keys, ka, sizeof(u64) * n, hipMemcpyDeviceToDevice, h->stream));
    TQ_HIP(h, hipMemcpyAsync(*vals, va, sizeof(i64) * n, hipMemcpyDeviceToDevice, h->stream));
  }
  return TQ_SUCCESS;
}

// Decode of linear indices under a (possibly soft-transposed) index table:
// coordinates are peeled off from the largest stride down (the order of
// util::index_to_indices with a sort index, src/util/space.h:38-71); legs of
// extent 1 have coordinate 0 and are skipped. ck = sum coord * cmul,
// fk = sum coord * fmul.
struct SpCode {
  int32_t n;                      // legs taking part in the decode
  int32_t leg[TQ_MAX_RANK];       // by decreasing stride
  i64 table[TQ_MAX_RANK];
  i64 cmul[TQ_MAX_RANK];
  i64 fmul[TQ_MAX_RANK];
};

__global__ void sp_decode_kernel(int64_t n, const i64 *__restrict__ idx, SpCode c,
                                 u64 *__restrict__ ck, u64 *__restrict__ fk) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  i64 rem = idx[i];
  u64 a = 0, b = 0;
  for (int q = 0; q < c.n; ++q) {
    const i64 t = c.table[q];
    const i64 x = rem / t;
    rem -= x * t;
    a += (u64)(x * c.cmul[q]);
    b += (u64)(x * c.fmul[q]);
  }
  if (ck) ck[i] = a;
  fk[i] = b;
}

// Run of A element i in B sorted by contracted key: [lo, lo + cnt).
__global__ void sp_range_kernel(int64_t na, const u64 *__restrict__ cka, int64_t nb,
                                const u64 *__restrict__ ckb_sorted, i64 *__restrict__ lo,
                                i64 *__restrict__ cnt) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= na) return;
  const u64 key = cka[i];
  int64_t l = 0, r = nb;  // first >= key
  while (l < r) {
    const int64_t m = (l + r) >> 1;
    if (ckb_sorted[m] < key)
      l = m + 1;
    else
      r = m;
  }
  int64_t u = l, r2 = nb;  // first > key
  while (u < r2) {
    const int64_t m = (u + r2) >> 1;
    if (ckb_sorted[m] <= key)
      u = m + 1;
    else
      r2 = m;
  }
  lo[i] = l;
  cnt[i] = u - l;
}

// Product p: A element i with off[i] <= p < off[i + 1], B element at sorted
// position lo[i] + p - off[i].
template <class T>
__global__ void sp_expand_kernel(int64_t total, int64_t na, const i64 *__restrict__ off,
                                 const i64 *__restrict__ lo, const i64 *__restrict__ b_pos,
                                 const u64 *__restrict__ fka, const u64 *__restrict__ fkb,
                                 const T *__restrict__ va, const T *__restrict__ vb,
                                 u64 *__restrict__ keys, i64 *__restrict__ pay,
                                 double *__restrict__ prod) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= total) return;
  int64_t l = 0, r = na - 1;  // last i with off[i] <= p
  while (l < r) {
    const int64_t m = (l + r + 1) >> 1;
    if (off[m] <= p)
      l = m;
    else
      r = m - 1;
  }
  const int64_t j = b_pos[lo[l] + (p - off[l])];
  prod[p] = (double)va[l] * (double)vb[j];
  if (keys) {
    keys[p] = fka[l] + fkb[j];
    pay[p] = p;
  }
}

__global__ void sp_head_kernel(int64_t n, const u64 *__restrict__ keys, i64 *__restrict__ head) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  head[q] = (q == 0 || keys[q] != keys[q - 1]) ? 1 : 0;
}

__global__ void sp_seg_start_kernel(int64_t n, const u64 *__restrict__ keys,
                                    const i64 *__restrict__ seg, i64 *__restrict__ start) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  if (q == 0 || keys[q] != keys[q - 1]) start[seg[q]] = q;
}

// Run u of equal keys [start[u], start[u + 1]) summed in order in FP64.
template <class T>
__global__ void sp_segment_sum_kernel(int64_t n_seg, int64_t n, const i64 *__restrict__ start,
                                      const u64 *__restrict__ keys, const i64 *__restrict__ pay,
                                      const double *__restrict__ prod,
                                      i64 *__restrict__ out_idx, T *__restrict__ out_val) {
  const int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (u >= n_seg) return;
  const int64_t s = start[u], e = (u + 1 < n_seg) ? start[u + 1] : n;
  double acc = 0.0;
  for (int64_t q = s; q < e; ++q) acc += prod[pay[q]];
  out_idx[u] = (i64)keys[s];
  out_val[u] = (T)acc;
}

template <class T>
__global__ void sp_scalar_out_kernel(const double *__restrict__ sum, int64_t has,
                                     i64 *__restrict__ out_idx, T *__restrict__ out_val) {
  out_idx[0] = 0;
  out_val[0] = has ? (T)sum[0] : (T)0;
}

// Host-side description of one operand after validation.
struct SpIn {
  int rank;
  std::vector<i64> dims, table;
};

tq_status read_desc(tq_handle_t h, const tq_sp_desc *d, SpIn &s) {
  if (!d) return TQ_ERROR_INVALID_VALUE;
  if (d->rank < 1 || d->rank > TQ_MAX_RANK)
    return tq_fail(h, TQ_ERROR_UNSUPPORTED, "sparse: rank must be 1..TQ_MAX_RANK");
  if (!d->dims || d->nnz < 0 || (d->nnz > 0 && (!d->indices || !d->values)))
    return TQ_ERROR_INVALID_VALUE;
  s.rank = d->rank;
  s.dims.assign(d->dims, d->dims + d->rank);
  s.table.resize(d->rank);
  i64 acc = 1;
  for (int l = 0; l < d->rank; ++l) {
    if (s.dims[l] < 1) return TQ_ERROR_INVALID_VALUE;
    s.table[l] = d->table ? d->table[l] : acc;
    if (s.table[l] < 1) return TQ_ERROR_INVALID_VALUE;
    if (acc > (i64(1) << 62) / s.dims[l])
      return tq_fail(h, TQ_ERROR_UNSUPPORTED, "sparse: tensor above 2^62 elements");
    acc *= s.dims[l];
  }
  return TQ_SUCCESS;
}

// Decode multipliers of one operand; cm/fm give each leg's weight in the
// contracted and free keys.
SpCode make_code(const SpIn &s, const std::vector<i64> &cm, const std::vector<i64> &fm) {
  SpCode c{};
  std::vector<int> legs;
  for (int l = 0; l < s.rank; ++l)
    if (s.dims[l] > 1) legs.push_back(l);
  std::stable_sort(legs.begin(), legs.end(),
                   [&](int x, int y) { return s.table[x] > s.table[y]; });
  c.n = (int32_t)legs.size();
  for (int q = 0; q < c.n; ++q) {
    const int l = legs[q];
    c.leg[q] = l;
    c.table[q] = s.table[l];
    c.cmul[q] = cm[l];
    c.fmul[q] = fm[l];
  }
  return c;
}

template <class T>
tq_status sparse_contract_t(tq_handle_t h, const tq_sp_desc *da, const tq_sp_desc *db,
                            const SpIn &A, const SpIn &B, int n_legs, const int32_t *legs,
                            int64_t *out_dims, int64_t *out_nnz, int64_t **out_indices,
                            void **out_values) {
  // legs and output shape (src/gpu/sparse.cuh:277-303)
  std::vector<bool> ac(A.rank, false), bc(B.rank, false);
  std::vector<i64> ctab(n_legs);
  i64 cacc = 1;
  for (int k = 0; k < n_legs; ++k) {
    const int la = legs[2 * k], lb = legs[2 * k + 1];
    if (la < 0 || la >= A.rank || lb < 0 || lb >= B.rank || ac[la] || bc[lb])
      return tq_fail(h, TQ_ERROR_INVALID_VALUE, "sparse: invalid contracted legs");
    if (A.dims[la] != B.dims[lb])
      return tq_fail(h, TQ_ERROR_DIMENSION_MISMATCH,
                     "The dimensions from two tensors to be contracted do not match");
    ac[la] = bc[lb] = true;
    ctab[k] = cacc;
    cacc *= A.dims[la];
  }
  std::vector<i64> od;
  for (int l = 0; l < A.rank; ++l)
    if (!ac[l]) od.push_back(A.dims[l]);
  const int nfa = (int)od.size();
  for (int l = 0; l < B.rank; ++l)
    if (!bc[l]) od.push_back(B.dims[l]);
  const int out_rank = (int)od.size();
  std::vector<i64> otab(out_rank);
  i64 oacc = 1;
  for (int p = 0; p < out_rank; ++p) {
    otab[p] = oacc;
    if (oacc > (i64(1) << 62) / od[p])
      return tq_fail(h, TQ_ERROR_UNSUPPORTED, "sparse: output above 2^62 elements");
    oacc *= od[p];
  }
  if (out_dims)
    for (int p = 0; p < out_rank; ++p) out_dims[p] = od[p];

  std::vector<i64> cma(A.rank, 0), fma(A.rank, 0), cmb(B.rank, 0), fmb(B.rank, 0);
  for (int k = 0; k < n_legs; ++k) {
    cma[legs[2 * k]] = ctab[k];
    cmb[legs[2 * k + 1]] = ctab[k];
  }
  for (int l = 0, p = 0; l < A.rank; ++l)
    if (!ac[l]) fma[l] = otab[p++];
  for (int l = 0, p = nfa; l < B.rank; ++l)
    if (!bc[l]) fmb[l] = otab[p++];

  const int64_t na = da->nnz, nb = db->nnz;
  DevBuf cka, fka, ckb, fkb, bpos, lo, cnt, scan_tmp, total_dev;
  TQ_HIP(h, cka.alloc(sizeof(u64) * std::max<int64_t>(na, 1)));
  TQ_HIP(h, fka.alloc(sizeof(u64) * std::max<int64_t>(na, 1)));
  TQ_HIP(h, ckb.alloc(sizeof(u64) * std::max<int64_t>(nb, 1)));
  TQ_HIP(h, fkb.alloc(sizeof(u64) * std::max<int64_t>(nb, 1)));
  TQ_HIP(h, bpos.alloc(sizeof(i64) * std::max<int64_t>(nb, 1)));
  TQ_HIP(h, lo.alloc(sizeof(i64) * std::max<int64_t>(na, 1)));
  TQ_HIP(h, cnt.alloc(sizeof(i64) * (na + 1)));
  TQ_HIP(h, scan_tmp.alloc(tq_scan_tmp_bytes(na + 1)));
  int64_t total = 0;
  SortTmp st;
  if (na > 0 && nb > 0) {
    const SpCode code_a = make_code(A, cma, fma), code_b = make_code(B, cmb, fmb);
    hipLaunchKernelGGL(sp_decode_kernel, dim3(grid_for(na)), dim3(256), 0, h->stream, na,
                       da->indices, code_a, cka.as<u64>(), fka.as<u64>());
    hipLaunchKernelGGL(sp_decode_kernel, dim3(grid_for(nb)), dim3(256), 0, h->stream, nb,
                       db->indices, code_b, ckb.as<u64>(), fkb.as<u64>());
    hipLaunchKernelGGL(iota_kernel, dim3(grid_for(nb)), dim3(256), 0, h->stream, bpos.as<i64>(),
                       nb);
    TQ_HIP(h, hipGetLastError());
    TQ_CHECK(st.alloc(h, nb));
    u64 *kb = ckb.as<u64>();
    i64 *vb = bpos.as<i64>();
    TQ_CHECK(radix_sort(h, &kb, &vb, nb, bits_for((u64)(cacc - 1)), st));
    hipLaunchKernelGGL(sp_range_kernel, dim3(grid_for(na)), dim3(256), 0, h->stream, na,
                       cka.as<u64>(), nb, kb, lo.as<i64>(), cnt.as<i64>());
    TQ_CHECK(tq_scan_i64(h, cnt.as<i64>(), cnt.as<i64>(), na, cnt.as<i64>() + na, scan_tmp.p));
    TQ_HIP(h, hipMemcpyAsync(&total, cnt.as<i64>() + na, sizeof(i64), hipMemcpyDeviceToHost,
                             h->stream));
    TQ_HIP(h, hipStreamSynchronize(h->stream));
  }

  const T *va = reinterpret_cast<const T *>(da->values);
  const T *vb = reinterpret_cast<const T *>(db->values);
  DevBuf keys, pay, prod;
  TQ_HIP(h, prod.alloc(sizeof(double) * std::max<int64_t>(total, 1)));
  void *oi = nullptr, *ov = nullptr;
  if (out_rank == 0) {
    // full contraction: one element {0: sum of the products}
    DevBuf red;
    TQ_HIP(h, red.alloc(tq_sum_tmp_bytes(total) + sizeof(double)));
    double *sum = red.as<double>() + tq_sum_tmp_bytes(total) / sizeof(double);
    if (total > 0) {
      hipLaunchKernelGGL(sp_expand_kernel<T>, dim3(grid_for(total)), dim3(256), 0, h->stream,
                         total, na, cnt.as<i64>(), lo.as<i64>(), bpos.as<i64>(),
                         fka.as<u64>(), fkb.as<u64>(), va, vb, nullptr, nullptr,
                         prod.as<double>());
      TQ_CHECK(tq_sum_f64_dev(h, prod.as<double>(), total, sum, red.p));
    }
    TQ_HIP(h, hipMalloc(&oi, sizeof(i64)));
    if (hipMalloc(&ov, sizeof(T)) != hipSuccess) {
      (void)hipFree(oi);
      return tq_fail(h, TQ_ERROR_ALLOCATION, "sparse: output allocation failed");
    }
    hipLaunchKernelGGL(sp_scalar_out_kernel<T>, dim3(1), dim3(1), 0, h->stream, sum,
                       (int64_t)(total > 0), reinterpret_cast<i64 *>(oi),
                       reinterpret_cast<T *>(ov));
    TQ_HIP(h, hipGetLastError());
    TQ_HIP(h, hipStreamSynchronize(h->stream));
    *out_nnz = 1;
    *out_indices = reinterpret_cast<int64_t *>(oi);
    *out_values = ov;
    return TQ_SUCCESS;
  }

  int64_t n_seg = 0;
  DevBuf head, start;
  if (total > 0) {
    TQ_HIP(h, keys.alloc(sizeof(u64) * total));
    TQ_HIP(h, pay.alloc(sizeof(i64) * total));
    hipLaunchKernelGGL(sp_expand_kernel<T>, dim3(grid_for(total)), dim3(256), 0, h->stream,
                       total, na, cnt.as<i64>(), lo.as<i64>(), bpos.as<i64>(), fka.as<u64>(),
                       fkb.as<u64>(), va, vb, keys.as<u64>(), pay.as<i64>(), prod.as<double>());
    TQ_HIP(h, hipGetLastError());
    SortTmp st2;
    TQ_CHECK(st2.alloc(h, total));
    u64 *k = keys.as<u64>();
    i64 *v = pay.as<i64>();
    TQ_CHECK(radix_sort(h, &k, &v, total, bits_for((u64)(oacc - 1)), st2));
    TQ_HIP(h, head.alloc(sizeof(i64) * (total + 1)));
    DevBuf htmp;
    TQ_HIP(h, htmp.alloc(tq_scan_tmp_bytes(total + 1)));
    hipLaunchKernelGGL(sp_head_kernel, dim3(grid_for(total)), dim3(256), 0, h->stream, total, k,
                       head.as<i64>());
    TQ_CHECK(tq_scan_i64(h, head.as<i64>(), head.as<i64>(), total, head.as<i64>() + total,
                         htmp.p));
    TQ_HIP(h, hipMemcpyAsync(&n_seg, head.as<i64>() + total, sizeof(i64),
                             hipMemcpyDeviceToHost, h->stream));
    TQ_HIP(h, hipStreamSynchronize(h->stream));
    TQ_HIP(h, start.alloc(sizeof(i64) * std::max<int64_t>(n_seg, 1)));
    hipLaunchKernelGGL(sp_seg_start_kernel, dim3(grid_for(total)), dim3(256), 0, h->stream,
                       total, k, head.as<i64>(), start.as<i64>());
    TQ_HIP(h, hipMalloc(&oi, sizeof(i64) * std::max<int64_t>(n_seg, 1)));
    if (hipMalloc(&ov, sizeof(T) * std::max<int64_t>(n_seg, 1)) != hipSuccess) {
      (void)hipFree(oi);
      return tq_fail(h, TQ_ERROR_ALLOCATION, "sparse: output allocation failed");
    }
    hipLaunchKernelGGL(sp_segment_sum_kernel<T>, dim3(grid_for(n_seg)), dim3(256), 0, h->stream,
                       n_seg, total, start.as<i64>(), k, v, prod.as<double>(),
                       reinterpret_cast<i64 *>(oi), reinterpret_cast<T *>(ov));
    TQ_HIP(h, hipGetLastError());
  } else {
    TQ_HIP(h, hipMalloc(&oi, sizeof(i64)));
    if (hipMalloc(&ov, sizeof(T)) != hipSuccess) {
      (void)hipFree(oi);
      return tq_fail(h, TQ_ERROR_ALLOCATION, "sparse: output allocation failed");
    }
  }
  // the temporaries are freed on return: wait for the kernels using them
  TQ_HIP(h, hipStreamSynchronize(h->stream));
  *out_nnz = n_seg;
  *out_indices = reinterpret_cast<int64_t *>(oi);
  *out_values = ov;
  return TQ_SUCCESS;
}

}  // namespace

extern "C" {

tq_status tq_sparse_contract(tq_handle_t h, tq_dtype dtype, const tq_sp_desc *a,
                             const tq_sp_desc *b, int32_t n_legs, const int32_t *legs,
                             int64_t *out_dims, int64_t *out_nnz, int64_t **out_indices,
                             void **out_values) {
  if (!h) return TQ_ERROR_NOT_INITIALIZED;
  if (dtype != TQ_F32 && dtype != TQ_F64) return TQ_ERROR_UNSUPPORTED;
  if (!out_nnz || !out_indices || !out_values || n_legs < 0 || (n_legs > 0 && !legs))
    return TQ_ERROR_INVALID_VALUE;
  *out_nnz = 0;
  *out_indices = nullptr;
  *out_values = nullptr;
  SpIn A, B;
  TQ_CHECK(read_desc(h, a, A));
  TQ_CHECK(read_desc(h, b, B));
  if (n_legs > A.rank || n_legs > B.rank)
    return tq_fail(h, TQ_ERROR_INVALID_VALUE, "sparse: invalid contracted legs");
  if (dtype == TQ_F64)
    return sparse_contract_t<double>(h, a, b, A, B, n_legs, legs, out_dims, out_nnz,
                                     out_indices, out_values);
  return sparse_contract_t<float>(h, a, b, A, B, n_legs, legs, out_dims, out_nnz, out_indices,
                                  out_values);
}

tq_status tq_sparse_permute_indices(tq_handle_t h, int32_t rank, const int64_t *dims,
                                    const int64_t *table, const int32_t *perm, int64_t nnz,
                                    const int64_t *in_indices, int64_t *out_indices) {
  if (!h) return TQ_ERROR_NOT_INITIALIZED;
  if (rank < 1 || rank > TQ_MAX_RANK) return TQ_ERROR_UNSUPPORTED;
  if (!dims || !perm || nnz < 0 || (nnz > 0 && (!in_indices || !out_indices)))
    return TQ_ERROR_INVALID_VALUE;
  bool seen[TQ_MAX_RANK] = {false};
  for (int k = 0; k < rank; ++k) {
    if (perm[k] < 0 || perm[k] >= rank || seen[perm[k]])
      return tq_fail(h, TQ_ERROR_RANK_MISMATCH,
                     "The number of permutation does not match the rank of tensor");
    seen[perm[k]] = true;
  }
  tq_sp_desc d{rank, dims, table, nnz, in_indices, in_indices};
  SpIn s;
  TQ_CHECK(read_desc(h, &d, s));
  if (nnz == 0) return TQ_SUCCESS;
  // new table: compact column-major over dims[perm[k]]; leg perm[k] of the
  // input carries output stride new_table[k]
  std::vector<i64> cm(rank, 0), fm(rank, 0);
  i64 acc = 1;
  for (int k = 0; k < rank; ++k) {
    fm[perm[k]] = acc;
    acc *= s.dims[perm[k]];
  }
  const SpCode code = make_code(s, cm, fm);
  hipLaunchKernelGGL(sp_decode_kernel, dim3(grid_for(nnz)), dim3(256), 0, h->stream, nnz,
                     in_indices, code, nullptr, reinterpret_cast<u64 *>(out_indices));
  TQ_HIP(h, hipGetLastError());
  return TQ_SUCCESS;
}

}  // extern "C"
