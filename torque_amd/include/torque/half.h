// torque/half.h -- IEEE binary16 storage type for the drop-in classes (the
// reference's `half` instantiation, src/gpu/block_sparse.cu:94-148). Host
// code holds and converts the bits; the device reads them as FP16.
#pragma once

#include <cstdint>
#include <cstring>

namespace torque {

struct half {
  uint16_t bits = 0;

  half() = default;
  half(float f) : bits(from_float(f)) {}
  half(double d) : bits(from_float((float)d)) {}
  half(int i) : bits(from_float((float)i)) {}
  operator float() const { return to_float(bits); }

  half &operator+=(half o) {
    bits = from_float(to_float(bits) + to_float(o.bits));
    return *this;
  }

  // binary32 -> binary16, rounding to nearest even; overflow to infinity,
  // NaN kept quiet, values below the normal range as subnormals.
  static uint16_t from_float(float f) {
    uint32_t x;
    std::memcpy(&x, &f, 4);
    const uint32_t sign = (x >> 16) & 0x8000u;
    const uint32_t ax = x & 0x7fffffffu;
    if (ax >= 0x7f800000u)  // inf or nan
      return (uint16_t)(sign | 0x7c00u | (ax > 0x7f800000u ? 0x200u : 0u));
    if (ax >= 0x477ff000u) return (uint16_t)(sign | 0x7c00u);  // rounds past 65504
    if (ax < 0x38800000u) {
      // subnormal half (or zero): value = m * 2^-24
      if (ax < 0x33000000u) return (uint16_t)sign;  // below half the smallest subnormal
      const uint32_t e = ax >> 23;
      const uint32_t m = (ax & 0x7fffffu) | 0x800000u;
      const uint32_t shift = 126 - e;  // 14 .. 24 places
      uint32_t r = m >> shift;
      const uint32_t rem = m & ((1u << shift) - 1), halfway = 1u << (shift - 1);
      if (rem > halfway || (rem == halfway && (r & 1u))) ++r;
      return (uint16_t)(sign | r);
    }
    uint32_t r = ((ax - 0x38000000u) >> 13);  // rebias exponent 127 -> 15
    const uint32_t rem = ax & 0x1fffu;
    if (rem > 0x1000u || (rem == 0x1000u && (r & 1u))) ++r;
    return (uint16_t)(sign | r);
  }

  static float to_float(uint16_t h) {
    const uint32_t sign = (uint32_t)(h & 0x8000u) << 16;
    const uint32_t e = (h >> 10) & 0x1fu, m = h & 0x3ffu;
    uint32_t x;
    if (e == 0x1fu) {
      x = sign | 0x7f800000u | (m << 13);
    } else if (e == 0) {
      if (m == 0) {
        x = sign;
      } else {  // subnormal: normalise
        uint32_t mm = m, ee = 113;
        while (!(mm & 0x400u)) {
          mm <<= 1;
          --ee;
        }
        x = sign | (ee << 23) | ((mm & 0x3ffu) << 13);
      }
    } else {
      x = sign | ((e + 112) << 23) | (m << 13);
    }
    float f;
    std::memcpy(&f, &x, 4);
    return f;
  }
};

}  // namespace torque
