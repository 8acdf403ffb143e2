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

// The FP64 operator's angle as a 64-bit fraction of a turn (even multiples of
// 2^-64 turn; the odd value kNanTurn64 marks a NaN angle): 8 bytes instead of
// the 16 of an FP64 (cos, sin) pair.  atan2 / (2 pi) carries ~1e-16 of
// relative error, so the decoded pair is within a few 1e-16 of the stored one.
constexpr uint64_t kNanTurn64 = ~(uint64_t)0;

__device__ __forceinline__ uint64_t turn64_of(double c, double s) {
  const double t = atan2(s, c) * 0.15915494309189533577;  // t in [-1/2, 1/2]
  if (!(t == t)) return kNanTurn64;
  const int64_t k = (int64_t)llrint(t * 9223372036854775808.0 * 0.5);  // t 2^62, |k| <= 2^61
  return (uint64_t)k << 2;
}

// (cos, sin) in FP64 of a 64-bit turn: the nearest quarter turn taken off in
// integer arithmetic, x = f 2 pi / 2^64 with |x| <= pi/4, then the Taylor
// series of sin to x^15 and of cos to x^16 (the next terms are below 5e-17
// there) and the quarter-turn rotation.
__device__ __forceinline__ double2 cs_of_turn64(uint64_t u) {
  const uint64_t q = (u + ((uint64_t)1 << 61)) >> 62;
  const int64_t f = (int64_t)(u - (q << 62));  // |f| <= 2^61
  const double x = (double)f * 3.4061215800865545e-19;  // 2 pi / 2^64
  // explicit fused multiply-adds: every kernel that decodes gets the same bits
  const double x2 = x * x;
  double ps = -1.0 / 1307674368000.0;
  ps = fma(ps, x2, 1.0 / 6227020800.0);
  ps = fma(ps, x2, -1.0 / 39916800);
  ps = fma(ps, x2, 1.0 / 362880);
  ps = fma(ps, x2, -1.0 / 5040);
  ps = fma(ps, x2, 1.0 / 120);
  ps = fma(ps, x2, -1.0 / 6);
  const double s = fma(x * x2, ps, x);
  double pc = 1.0 / 20922789888000.0;
  pc = fma(pc, x2, -1.0 / 87178291200.0);
  pc = fma(pc, x2, 1.0 / 479001600);
  pc = fma(pc, x2, -1.0 / 3628800);
  pc = fma(pc, x2, 1.0 / 40320);
  pc = fma(pc, x2, -1.0 / 720);
  pc = fma(pc, x2, 1.0 / 24);
  pc = fma(pc, x2, -0.5);
  const double c = fma(x2, pc, 1.0);
  double a = (q & 1) ? -s : c;
  double b = (q & 1) ? c : s;
  if (q & 2) {
    a = -a;
    b = -b;
  }
  if (u == kNanTurn64) a = b = __longlong_as_double(0x7ff8000000000000LL);
  return make_double2(a, b);
}

// The same through a table of (cos, sin) of j/256 turn, j = 0..255 (made by
// cs_of_turn64, e.g. in LDS): the nearest 1/256 turn A = j/256 is taken off
// in integer arithmetic and read from the table; the remainder
// x = f 2 pi / 2^64, |x| <= pi/256, needs only sin x to x^5 and cos x - 1 to
// x^6 (next terms below 1e-17), then
// cos(A + x) = cos A + (cos A (cos x - 1) - sin A sin x) and
// sin(A + x) = sin A + (sin A (cos x - 1) + cos A sin x):
// about 20 fewer vector instructions a sample, within ~2e-16 of
// cs_of_turn64.
__device__ __forceinline__ double2 cs_of_turn64_tab(uint64_t u, const double2 *tab) {
  const uint32_t j = (uint32_t)((u + ((uint64_t)1 << 55)) >> 56);
  const int64_t f = (int64_t)(u - ((uint64_t)j << 56));  // |f| <= 2^55
  const double x = (double)f * 3.4061215800865545e-19;  // 2 pi / 2^64
  const double x2 = x * x;
  const double s = fma(x * x2, fma(x2, 1.0 / 120, -1.0 / 6), x);
  const double cm1 = x2 * fma(x2, fma(x2, -1.0 / 720, 1.0 / 24), -0.5);
  const double2 t = tab[j];
  double a = t.x + fma(t.x, cm1, -(t.y * s));
  double b = t.y + fma(t.y, cm1, t.x * s);
  if (u == kNanTurn64) a = b = __longlong_as_double(0x7ff8000000000000LL);
  return make_double2(a, b);
}
