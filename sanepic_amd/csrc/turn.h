// turn.h -- a polarisation angle 2psi stored as a 32-bit fraction of a turn:
// the FP32 operator's per-sample weights (cos 2psi, sin 2psi) in 4 bytes
// instead of the 8 of an FP32 pair, decoded in the kernels that read them.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

// Angles are even multiples of 2^-32 turn (resolution 2.9e-9 rad); the odd
// value kNanTurn marks a NaN angle (a segment whose psi starts with NaN).
constexpr uint32_t kNanTurn = 0xFFFFFFFFu;

// atan2(s, c) as a fraction of a turn, rounded to the nearest 2^-31 turn;
// negative angles wrap into [1/2, 1) through two's complement.
__device__ __forceinline__ uint32_t turn_of(double c, double s) {
  const double t = atan2(s, c) * 0.15915494309189533577;  // 1 / (2 pi): t in [-1/2, 1/2]
  if (!(t == t)) return kNanTurn;
  const int64_t k = (int64_t)llrint(t * 2147483648.0);   // |k| <= 2^30
  return (uint32_t)k << 1;
}

// (cos, sin) of a stored angle in FP32.  The nearest quarter turn q is taken
// off exactly in integer arithmetic, leaving x = f 2 pi / 2^32 with
// |x| <= pi/4; sin x and cos x are their Taylor series to x^9 and x^10
// (truncation below 2e-9 there); the quarter turn is a rotation,
// (cos, sin)(q pi/2 + x) = (c, s), (-s, c), (-c, -s), (s, -c) for q = 0..3.
__device__ __forceinline__ float2 cs_of_turn(uint32_t u) {
  const uint32_t q = (u + 0x20000000u) >> 30;
  const int32_t f = (int32_t)(u - (q << 30));             // |f| <= 2^29
  const float x = (float)f * 1.46291807926715968e-9f;     // 2 pi / 2^32
  const float x2 = x * x;
  const float s = x + x * x2 * (-1.66666667e-1f + x2 * (8.33333333e-3f + x2 * (-1.98412698e-4f + x2 * 2.75573192e-6f)));
  const float c = 1.0f + x2 * (-0.5f + x2 * (4.16666667e-2f + x2 * (-1.38888889e-3f +
                                                                  x2 * (2.48015873e-5f - x2 * 2.75573192e-7f))));
  float a = (q & 1) ? -s : c;
  float b = (q & 1) ? c : s;
  if (q & 2) {
    a = -a;
    b = -b;
  }
  if (u == kNanTurn) a = b = __int_as_float(0x7fc00000);
  return make_float2(a, b);
}
