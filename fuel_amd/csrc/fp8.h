// FP8 E4M3 with power-of-two scales, as the gfx950 matrix cores read them
// (v_mfma_scale_f32_32x32x64_f8f6f4, operand format 0 = E4M3 with exponent
// bias 7, no infinities, 0x7F / 0xFF NaN, largest finite 448; scales E8M0,
// 2^(s - 127)).  The same code runs on the host (checks) and on the device
// (operand and query fragments), and works on the f32 bits, so its rounding
// does not depend on the floating-point mode.
//
// A vector v is stored as codes c_i = e4m3(v_i 2^-e) with one exponent e per
// vector (fp8_scale_exp of max |v_i|, so that every |v_i| 2^-e <= 448), and
// each element's error is at most 2^-4 |v_i| + 2^(e - 10) (half a unit in the
// last place of a normal code, or half the subnormal step 2^-9 scaled).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace fa {

constexpr int FP8_SCALE_MIN = -100;  // E8M0 code 27
constexpr int FP8_SCALE_MAX = 100;   // E8M0 code 227

__host__ __device__ inline uint32_t fp8_f32_bits(float x) {
  union {
    float f;
    uint32_t u;
  } c;
  c.f = x;
  return c.u;
}

// The smallest e with m 2^-e <= 448 (m = the largest |v_i|; m = 0 gives 0),
// clamped to [FP8_SCALE_MIN, FP8_SCALE_MAX].  With m = M 2^E, 1 <= M < 2:
// e = E - 8 when M <= 1.75, else E - 7, so m 2^-e lies in (224, 448].
__host__ __device__ inline int fp8_scale_exp(float m) {
  const uint32_t u = fp8_f32_bits(m) & 0x7fffffffu;
  if (u == 0) return 0;
  if (u < 0x00800000u) return FP8_SCALE_MIN;  // f32-subnormal maximum
  if (u >= 0x7f800000u) return FP8_SCALE_MAX;
  const int E = int(u >> 23) - 127;
  const int e = (u & 0x7fffffu) <= 0x600000u ? E - 8 : E - 7;
  return e < FP8_SCALE_MIN ? FP8_SCALE_MIN : (e > FP8_SCALE_MAX ? FP8_SCALE_MAX : e);
}

// Code of x 2^-e rounded to the nearest E4M3 value, ties to even, for
// |x| 2^-e <= 448.  f32-subnormal x (|x| < 2^-126) and scaled values below
// 2^-10 round to zero.
__host__ __device__ inline uint32_t fp8_e4m3_encode(float x, int e) {
  uint32_t u = fp8_f32_bits(x);
  const uint32_t s = (u >> 24) & 0x80u;
  u &= 0x7fffffffu;
  if (u < 0x00800000u) return s;
  const int E = int(u >> 23) - 127 - e;  // exponent of x 2^-e
  if (E < -10) return s;
  const uint32_t mant = (u & 0x7fffffu) | 0x800000u;  // x 2^-e = mant 2^(E - 23)
  // units of the result: 2^(E - 3) for a normal code, 2^-9 below 2^-6
  const int shift = 20 + (E < -6 ? -6 - E : 0);
  uint32_t r = mant >> shift;
  const uint32_t rem = mant & ((1u << shift) - 1u);
  const uint32_t half = 1u << (shift - 1);
  if (rem > half || (rem == half && (r & 1u))) ++r;
  uint32_t code;
  if (E < -6) {
    code = r;  // up to 8: the step to the smallest normal 2^-6 carries into the exponent field
  } else {
    uint32_t eb = uint32_t(E + 7);
    if (r == 16) {
      r = 8;
      ++eb;
    }
    code = (eb << 3) | (r - 8);
  }
  if (code > 0x7Eu) code = 0x7Eu;  // unreachable with a scale from fp8_scale_exp
  return s | code;
}

// Value of an E4M3 code (without its scale).
__host__ __device__ inline float fp8_e4m3_value(uint32_t c) {
  const int E = int(c >> 3) & 15;
  const int m = int(c & 7u);
  const float v = E == 0 ? float(m) * 0x1p-9f : float(8 + m) * float(1u << E) * 0x1p-10f;
  return (c & 0x80u) ? -v : v;
}

// E8M0 code of a scale exponent.
__host__ __device__ inline uint32_t fp8_scale_code(int e) { return uint32_t(e + 127); }

}  // namespace fa
