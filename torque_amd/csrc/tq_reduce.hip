// tq_reduce.hip -- FP64 reductions on the GPU: deterministic sums, dot
// products, norms and scaling. They replace cublasDdot + add<<<1,1>>> of the
// reference's rank-0 branch (src/gpu/block_sparse.cuh:979-1010) and give the
// network-scalar and norm reductions of environment tensors a fixed
// summation order (two stages, fixed grid), so repeated runs agree bitwise.
#include <algorithm>
#include <cmath>
#include <mutex>
#include <vector>

#include "tq_common.h"

namespace {

constexpr int kThreads = 256;
constexpr int kMaxBlocks = 1024;

__device__ inline double block_sum(double v) {
  __shared__ double ws[kThreads / 64];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) ws[wave] = v;
  __syncthreads();
  double t = 0.0;
  if (threadIdx.x == 0)
    for (int w = 0; w < kThreads / 64; ++w) t += ws[w];
  __syncthreads();
  return t;
}

// Each block sums a contiguous slice [lo, hi) with a fixed thread order.
template <bool kDot>
__global__ __launch_bounds__(kThreads) void partial_kernel(
    const double *__restrict__ x, const double *__restrict__ y, int64_t n,
    int64_t per_block, double *__restrict__ out) {
  const int64_t lo = (int64_t)blockIdx.x * per_block;
  const int64_t hi = min(n, lo + per_block);
  double s = 0.0;
  for (int64_t i = lo + threadIdx.x; i < hi; i += kThreads)
    s += kDot ? x[i] * y[i] : x[i];
  s = block_sum(s);
  if (threadIdx.x == 0) out[blockIdx.x] = s;
}

__global__ __launch_bounds__(kThreads) void final_kernel(
    const double *__restrict__ part, int64_t n, double *__restrict__ out,
    int take_sqrt) {
  double s = 0.0;
  for (int64_t i = threadIdx.x; i < n; i += kThreads) s += part[i];
  s = block_sum(s);
  if (threadIdx.x == 0) out[0] = take_sqrt ? sqrt(s) : s;
}

__global__ void scale_kernel(double *__restrict__ x, int64_t n, double alpha) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    x[i] *= alpha;
}

// x[i] <- x[i] / norm[0], the norm read on the device; x is left unchanged
// when the norm is zero.
__global__ void inv_scale_kernel(double *__restrict__ x, int64_t n,
                                 const double *__restrict__ norm) {
  const double s = norm[0];
  if (s == 0.0) return;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    x[i] = x[i] / s;
}

int64_t n_blocks_for(int64_t n) {
  return std::max<int64_t>(1, std::min<int64_t>(kMaxBlocks, ceil_div(n, 8 * kThreads)));
}

tq_status reduce(tq_handle_t h, const double *x, const double *y, int64_t n,
                 double *out_dev, void *tmp, bool take_sqrt) {
  const int64_t nb = n_blocks_for(n);
  const int64_t per = ceil_div(std::max<int64_t>(n, 1), nb);
  double *part = reinterpret_cast<double *>(tmp);
  if (y)
    hipLaunchKernelGGL(partial_kernel<true>, dim3((unsigned)nb), dim3(kThreads), 0,
                       h->stream, x, y, n, per, part);
  else
    hipLaunchKernelGGL(partial_kernel<false>, dim3((unsigned)nb), dim3(kThreads), 0,
                       h->stream, x, nullptr, n, per, part);
  hipLaunchKernelGGL(final_kernel, dim3(1), dim3(kThreads), 0, h->stream, part, nb,
                     out_dev, take_sqrt ? 1 : 0);
  TQ_HIP(h, hipGetLastError());
  return TQ_SUCCESS;
}

// norm[0] = ||x|| (device), then x <- x / ||x||: two reduction launches and
// one scaling launch on the handle's stream, partial sums in tmp.
tq_status normalize(tq_handle_t h, int64_t n, double *x, double *norm, void *tmp) {
  TQ_CHECK(reduce(h, x, x, n, norm, tmp, true));
  if (n == 0) return TQ_SUCCESS;
  const int64_t g = std::min<int64_t>(ceil_div(n, 256), (int64_t)h->n_cu * 16);
  hipLaunchKernelGGL(inv_scale_kernel, dim3((unsigned)g), dim3(256), 0, h->stream, x, n,
                     norm);
  TQ_HIP(h, hipGetLastError());
  return TQ_SUCCESS;
}

}  // namespace

// Scratch of tq_normalize_f64_noalloc: the first stage's partial sums (at
// most kMaxBlocks of them, whatever n is). Each device the library runs on
// holds one copy of the array until the process ends; host code reaches a
// device's copy only through the address hipGetSymbolAddress returned for it
// with that device current (scratch_dev), never through the array's name.
static __device__ __align__(256) double scratch[kMaxBlocks];

namespace {
std::mutex scratch_mu;
std::vector<double *> scratch_dev;  // device index -> address of its copy
}  // namespace

tq_status tq_reduce_setup_device(tq_handle_t h) {
  std::lock_guard<std::mutex> lock(scratch_mu);
  if ((int)scratch_dev.size() <= h->device) scratch_dev.resize(h->device + 1, nullptr);
  if (scratch_dev[h->device]) return TQ_SUCCESS;
  void *p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p)
    return tq_fail(h, TQ_ERROR_HIP,
                   std::string("scratch address: ") + hipGetErrorString(e));
  scratch_dev[h->device] = reinterpret_cast<double *>(p);
  return TQ_SUCCESS;
}

size_t tq_sum_tmp_bytes(int64_t n) { return sizeof(double) * (n_blocks_for(n) + 1); }

tq_status tq_sum_f64_dev(tq_handle_t h, const double *x, int64_t n, double *out,
                         void *tmp) {
  return reduce(h, x, nullptr, n, out, tmp, false);
}

extern "C" {

tq_status tq_dot_f64(tq_handle_t h, int64_t n, const double *x, const double *y,
                     double *result) {
  if (!h) return TQ_ERROR_NOT_INITIALIZED;
  if (n < 0 || !result || (n > 0 && (!x || !y))) return TQ_ERROR_INVALID_VALUE;
  DevBuf tmp;
  TQ_HIP(h, tmp.alloc(tq_sum_tmp_bytes(n) + sizeof(double)));
  double *out = tmp.as<double>() + n_blocks_for(n) + 1;
  TQ_CHECK(reduce(h, x, y, n, out, tmp.p, false));
  TQ_HIP(h, hipMemcpyAsync(result, out, sizeof(double), hipMemcpyDeviceToHost, h->stream));
  TQ_HIP(h, hipStreamSynchronize(h->stream));
  return TQ_SUCCESS;
}

tq_status tq_nrm2_f64(tq_handle_t h, int64_t n, const double *x, double *result) {
  if (!h) return TQ_ERROR_NOT_INITIALIZED;
  if (n < 0 || !result || (n > 0 && !x)) return TQ_ERROR_INVALID_VALUE;
  DevBuf tmp;
  TQ_HIP(h, tmp.alloc(tq_sum_tmp_bytes(n) + sizeof(double)));
  double *out = tmp.as<double>() + n_blocks_for(n) + 1;
  TQ_CHECK(reduce(h, x, x, n, out, tmp.p, true));
  TQ_HIP(h, hipMemcpyAsync(result, out, sizeof(double), hipMemcpyDeviceToHost, h->stream));
  TQ_HIP(h, hipStreamSynchronize(h->stream));
  return TQ_SUCCESS;
}

tq_status tq_normalize_f64(tq_handle_t h, int64_t n, double *x, double *norm) {
  if (!h) return TQ_ERROR_NOT_INITIALIZED;
  if (n < 0 || !norm || (n > 0 && !x)) return TQ_ERROR_INVALID_VALUE;
  DevBuf tmp;
  TQ_HIP(h, tmp.alloc(tq_sum_tmp_bytes(n)));
  TQ_CHECK(normalize(h, n, x, norm, tmp.p));
  // the partial sums are released on return
  TQ_HIP(h, hipStreamSynchronize(h->stream));
  return TQ_SUCCESS;
}

tq_status tq_normalize_f64_noalloc(tq_handle_t h, int64_t n, double *x, double *norm) {
  if (!h) return TQ_ERROR_NOT_INITIALIZED;
  if (n < 0 || !norm || (n > 0 && !x)) return TQ_ERROR_INVALID_VALUE;
  double *tmp = nullptr;
  {
    std::lock_guard<std::mutex> lock(scratch_mu);
    if (h->device < (int)scratch_dev.size()) tmp = scratch_dev[h->device];
  }
  if (!tmp)
    return tq_fail(h, TQ_ERROR_NOT_INITIALIZED, "normalize: device scratch not set up");
  // every partial is written before it is read: nothing needs clearing
  return normalize(h, n, x, norm, tmp);
}

tq_status tq_scale_f64(tq_handle_t h, int64_t n, double *x, double alpha) {
  if (!h) return TQ_ERROR_NOT_INITIALIZED;
  if (n < 0 || (n > 0 && !x)) return TQ_ERROR_INVALID_VALUE;
  if (n == 0) return TQ_SUCCESS;
  const int64_t g = std::min<int64_t>(ceil_div(n, 256), (int64_t)h->n_cu * 16);
  hipLaunchKernelGGL(scale_kernel, dim3((unsigned)g), dim3(256), 0, h->stream, x, n, alpha);
  TQ_HIP(h, hipGetLastError());
  return TQ_SUCCESS;
}

}  // extern "C"
