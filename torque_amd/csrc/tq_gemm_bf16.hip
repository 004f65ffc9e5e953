// tq_gemm_bf16.hip -- split-BF16 grouped GEMM tier (three BF16 pieces per
// FP64 operand, six piece products on BF16 MFMA with FP32 accumulation).
#include "tq_common.h"

int tq_bf16x3_n_configs() { return 1; }
GemmConfig tq_bf16x3_config(int) { return GemmConfig{64, 64, 32}; }

tq_status tq_launch_bf16x3_gemm(tq_handle_t h, int, const GemmDesc *,
                                const int32_t *, const int32_t *, int64_t,
                                const void *, const void *, int64_t, int64_t,
                                void *) {
  return tq_fail(h, TQ_ERROR_UNSUPPORTED, "split-BF16 tier not available");
}
