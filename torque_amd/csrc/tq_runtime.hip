// tq_runtime.hip -- handle, stream, status strings and device-memory helpers
// of the C-ABI. Replaces the reference's cublasHandle_t plumbing and its
// gpuErrchk/cublasCheck exit-on-error macros (src/gpu/util/lib_helper.cuh:
// 39-77): every failure is returned as a tq_status with the message kept on
// the handle.
#include <cstring>

#include "tq_common.h"
#include "tq_plan.h"

tq_status tq_fail(tq_handle_t h, tq_status s, const std::string &msg) {
  if (h) h->last_error = msg;
  return s;
}

tq_status tq_stage(tq_handle_t h, size_t bytes, void **host_ptr) {
  if (h->pinned_bytes < bytes) {
    if (h->pinned) {
      TQ_HIP(h, hipStreamSynchronize(h->stream));
      (void)hipHostFree(h->pinned);
      h->pinned = nullptr;
      h->pinned_bytes = 0;
    }
    size_t want = bytes < (1u << 20) ? (1u << 20) : bytes;
    TQ_HIP(h, hipHostMalloc(&h->pinned, want, hipHostMallocDefault));
    h->pinned_bytes = want;
  }
  *host_ptr = h->pinned;
  return TQ_SUCCESS;
}

extern "C" {

tq_status tq_create(tq_handle_t *handle, int device) {
  if (!handle) return TQ_ERROR_INVALID_VALUE;
  *handle = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || device < 0 || device >= n)
    return TQ_ERROR_INVALID_VALUE;
  auto *h = new tq_handle_s();
  h->device = device;
  if (hipSetDevice(device) != hipSuccess) {
    delete h;
    return TQ_ERROR_HIP;
  }
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) == hipSuccess)
    h->n_cu = prop.multiProcessorCount;
  h->stream = nullptr;
  const tq_status st = tq_reduce_setup_device(h);
  if (st != TQ_SUCCESS) {
    delete h;
    return st;
  }
  *handle = h;
  return TQ_SUCCESS;
}

tq_status tq_destroy(tq_handle_t h) {
  if (!h) return TQ_ERROR_NOT_INITIALIZED;
  (void)hipStreamSynchronize(h->stream);
  for (PlanCacheEntry &e : h->plan_cache) {
    e.plan->cached = false;
    (void)tq_plan_destroy(e.plan);
  }
  if (h->capture_stream) (void)hipStreamDestroy(h->capture_stream);
  if (h->pinned) (void)hipHostFree(h->pinned);
  delete h;
  return TQ_SUCCESS;
}

tq_status tq_set_stream(tq_handle_t h, void *stream) {
  if (!h) return TQ_ERROR_NOT_INITIALIZED;
  if (h->capturing)
    return tq_fail(h, TQ_ERROR_INVALID_VALUE, "set stream: the handle is capturing");
  h->stream = reinterpret_cast<hipStream_t>(stream);
  return TQ_SUCCESS;
}

tq_status tq_synchronize(tq_handle_t h) {
  if (!h) return TQ_ERROR_NOT_INITIALIZED;
  TQ_HIP(h, hipStreamSynchronize(h->stream));
  return TQ_SUCCESS;
}

const char *tq_status_string(tq_status s) {
  switch (s) {
    case TQ_SUCCESS: return "success";
    case TQ_ERROR_INVALID_VALUE: return "invalid value";
    case TQ_ERROR_DIMENSION_MISMATCH:
      return "The dimensions from two tensors to be contracted do not match";
    case TQ_ERROR_RANK_MISMATCH:
      return "The number of permutation does not match the rank of tensor";
    case TQ_ERROR_OUT_OF_RANGE: return "Indices out of boundary";
    case TQ_ERROR_ALLOCATION: return "allocation failed";
    case TQ_ERROR_HIP: return "HIP runtime error";
    case TQ_ERROR_UNSUPPORTED: return "unsupported argument";
    case TQ_ERROR_NOT_INITIALIZED: return "Tensor not initialized";
  }
  return "unknown status";
}

const char *tq_last_error(tq_handle_t h) {
  if (!h) return "";
  return h->last_error.c_str();
}

tq_status tq_malloc(tq_handle_t h, void **ptr, size_t bytes) {
  if (!h) return TQ_ERROR_NOT_INITIALIZED;
  if (!ptr) return TQ_ERROR_INVALID_VALUE;
  *ptr = nullptr;
  if (bytes == 0) bytes = 8;
  hipError_t e = hipMalloc(ptr, bytes);
  if (e != hipSuccess)
    return tq_fail(h, TQ_ERROR_ALLOCATION,
                   std::string("hipMalloc: ") + hipGetErrorString(e));
  return TQ_SUCCESS;
}

tq_status tq_free(tq_handle_t h, void *ptr) {
  if (!h) return TQ_ERROR_NOT_INITIALIZED;
  if (!ptr) return TQ_SUCCESS;
  TQ_HIP(h, hipFree(ptr));
  return TQ_SUCCESS;
}

tq_status tq_memcpy(tq_handle_t h, void *dst, const void *src, size_t bytes,
                    int kind) {
  if (!h) return TQ_ERROR_NOT_INITIALIZED;
  if (bytes == 0) return TQ_SUCCESS;
  if (!dst || !src) return TQ_ERROR_INVALID_VALUE;
  hipMemcpyKind k;
  switch (kind) {
    case 0: k = hipMemcpyHostToDevice; break;
    case 1: k = hipMemcpyDeviceToHost; break;
    case 2: k = hipMemcpyDeviceToDevice; break;
    default: return TQ_ERROR_INVALID_VALUE;
  }
  TQ_HIP(h, hipMemcpyAsync(dst, src, bytes, k, h->stream));
  TQ_HIP(h, hipStreamSynchronize(h->stream));
  return TQ_SUCCESS;
}

tq_status tq_memset_zero(tq_handle_t h, void *ptr, size_t bytes) {
  if (!h) return TQ_ERROR_NOT_INITIALIZED;
  if (bytes == 0) return TQ_SUCCESS;
  if (!ptr) return TQ_ERROR_INVALID_VALUE;
  TQ_HIP(h, hipMemsetAsync(ptr, 0, bytes, h->stream));
  return TQ_SUCCESS;
}

}  // extern "C"
