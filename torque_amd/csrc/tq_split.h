// tq_split.h -- the three-way BF16 split of FP64 values shared by the
// split-BF16 pack (tq_permute.hip), the split-BF16 GEMM's in-place operand
// loads (tq_gemm_bf16.hip) and the small-K streaming kernel (tq_stream.hip).
//
// x = hi + mid + lo + r with
//   hi  = bf16(fp32(x))                 r1 = x - hi   (exact in FP64)
//   mid = bf16(g),  g = fp32(r1)        r2 = g - mid  (exact in FP32)
//   lo  = bf16(r2)
// every rounding to nearest even (v_cvt_pk_bf16_f32 converts two FP32 values
// at once). |r| <= |r1 - g| + |r2 - lo| <= 2^-24 |r1| + 2^-9 |r2|, below
// 2^-26 |x| for |x| in the normal FP32 range with its lo piece normal
// (|x| >= 2^-108); NaN and infinity propagate as NaN.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace tqs {

typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));

// Two FP32 values rounded to BF16 and packed (a in the low half).
__device__ inline uint32_t pack_bf16(float a, float b) {
  return __builtin_bit_cast(uint32_t, __builtin_convertvector((f32x2){a, b}, bf16x2));
}
__device__ inline float low_value(uint32_t u) { return __uint_as_float(u << 16); }
__device__ inline float high_value(uint32_t u) { return __uint_as_float(u & 0xffff0000u); }

// Splits two values at once: h, m, l hold the hi, mid and lo pieces of x0
// (low halves) and x1 (high halves).
__device__ inline void split3x2(double x0, double x1, uint32_t &h, uint32_t &m, uint32_t &l) {
  h = pack_bf16((float)x0, (float)x1);
  const float g0 = (float)(x0 - (double)low_value(h));
  const float g1 = (float)(x1 - (double)high_value(h));
  m = pack_bf16(g0, g1);
  l = pack_bf16(g0 - low_value(m), g1 - high_value(m));
}

// Whether x lies outside the domain in which the split represents it to
// 2^-26: |x| >= 2^127 (the hi piece leaves the BF16 range), 0 < |x| < 2^-100
// (the lo piece falls below the normal range and loses bits), NaN or
// infinity (they split into NaN).
__device__ inline bool split_out_of_range(double x) {
  const double ax = __builtin_fabs(x);
  return !(ax < 0x1p127) || (ax < 0x1p-100 && ax != 0.0);
}

// The value the split represents, hi + mid + lo summed in FP64 in that order
// (the operand rounding of the split-BF16 tier).
__device__ inline double round3(double x) {
  uint32_t h, m, l;
  split3x2(x, 0.0, h, m, l);
  return (double)low_value(h) + (double)low_value(m) + (double)low_value(l);
}

}  // namespace tqs
