// Lossless two-word storage of an FP64 value as its FP32 rounding plus a
// 32-bit integer residual, used for the FP64 velocity of the fused FP32-bulk
// step: the FP32 word is the velocity copy that the FP32 sweeps read anyway,
// so the FP64 state costs one extra 4-byte word per point instead of eight.
//
// For a finite double u with hi = (float)u (round to nearest) and e the
// exponent of hi (e = -126 for hi zero or subnormal), |u - hi| <= 2^(e-24),
// and u - hi is a multiple of 2^(e-53) whenever |u| >= 2^-127 (the ulp of u
// is at least 2^(e-53)).  So  r = (u - hi) * 2^(53-e)  is an integer with
// |r| <= 2^29, and  u = hi + r * 2^(e-53)  exactly: both steps are exact in
// FP64 (a difference of two nearby doubles, then a power-of-two scaling).
// Below 2^-127 in magnitude the residual is rounded to a multiple of 2^-179
// (an absolute error under 1e-54); values beyond the FP32 range are not
// representable (hi is infinite).
//
// The functions compile for the host as well, so the round trip is testable
// on the CPU with the same code.
#pragma once

#include <cmath>
#include <cstdint>

#ifdef __HIPCC__
#define X3D_HD __host__ __device__
#else
#define X3D_HD
#endif

namespace x3d {

// exponent e of hi as above (-126 for zero and subnormals)
X3D_HD inline int split_exponent(float hi) {
  const uint32_t bits = __builtin_bit_cast(uint32_t, hi);
  const int be = (int)((bits >> 23) & 0xffu);
  return be == 0 ? -126 : be - 127;
}

// residual word of u given hi = (float)u
X3D_HD inline int32_t split_residual(double u, float hi) {
  const double d = u - (double)hi;
  return (int32_t)std::rint(std::ldexp(d, 53 - split_exponent(hi)));
}

// (r == 0 returns hi itself, which keeps the sign of a zero)
X3D_HD inline double join_split(float hi, int32_t r) {
  return r == 0 ? (double)hi : (double)hi + std::ldexp((double)r, split_exponent(hi) - 53);
}

}  // namespace x3d
