// Scalar FP64 math of the radiation kernels: large-argument sin/cos, correctly
// rounded quotients from a reciprocal, 2*pi reduction.  Plain C++ so that the same
// source is compiled into the gfx950 kernels (SRS_FN = __device__ __forceinline__)
// and into the CPU unit test of tests/test_device_math.py (SRS_FN = inline).
#pragma once
#include <stdint.h>

#ifndef SRS_FN
#error "define SRS_FN before including srs_math.h"
#endif

namespace srs {

// ---------------------------------------------------------------------------
// FP64 sine and cosine of large arguments (|x| up to ~1e15).
//
// Reduction: k = rint(x * 2/pi), r = x - k*(PIO2_HI + PIO2_LO) with two fused
// multiply-adds.  PIO2_HI = RN(pi/2) is a multiple of 2^-52 and so is every
// x >= 1, hence x - k*PIO2_HI is an exact multiple of 2^-52 below 2 in
// magnitude and the first FMA is exact; the second FMA rounds once.  The
// reduced argument is within one ulp for |x| < 2^51 (Cody & Waite's additive
// reduction, with the FMA removing the need to split k).  The quadrant comes
// from the low bits of k + 1.5*2^52 (an integer in [2^52, 2^53) for |k| < 2^51, so its
// mantissa holds k + 2^51, and 2^51 = 0 mod 4).
//
// Kernels: Taylor series of sin and cos on |r| <= pi/4 (+ rounding slack),
// truncated after r^17/17! and r^16/16!: truncation error < 1e-19, so the
// result is within about one ulp.
// ---------------------------------------------------------------------------
constexpr double kTwoOverPi = 0x1.45f306dc9c883p-1;
constexpr double kPio2Hi = 0x1.921fb54442d18p+0;
constexpr double kPio2Lo = 0x1.1a62633145c07p-54;

SRS_FN void sincos_fp64(double x, double* s_out, double* c_out) {
  double const k = __builtin_rint(x * kTwoOverPi);
  double r = __builtin_fma(-k, kPio2Hi, x);
  r = __builtin_fma(-k, kPio2Lo, r);
  double const kb = k + 0x1.8p52;
  uint64_t bits;
  __builtin_memcpy(&bits, &kb, sizeof bits);
  int const q = (int)(bits & 3u);
  double const r2 = r * r;
  // sin(r) = r + r^3 * P(r^2),  P = -1/3! + r^2/5! - ... - r^14/17!
  double ps = -2.8114572543455206e-15;
  ps = __builtin_fma(ps, r2, 7.647163731819816e-13);
  ps = __builtin_fma(ps, r2, -1.6059043836821613e-10);
  ps = __builtin_fma(ps, r2, 2.505210838544172e-08);
  ps = __builtin_fma(ps, r2, -2.7557319223985893e-06);
  ps = __builtin_fma(ps, r2, 0.0001984126984126984);
  ps = __builtin_fma(ps, r2, -0.008333333333333333);
  ps = __builtin_fma(ps, r2, 0.16666666666666666);
  double const sn = __builtin_fma(-(r * r2), ps, r);
  // cos(r) = 1 + r^2 * Q(r^2),  Q = -1/2! + r^2/4! - ... + r^14/16!
  double pc = 4.779477332387385e-14;
  pc = __builtin_fma(pc, r2, -1.1470745597729725e-11);
  pc = __builtin_fma(pc, r2, 2.08767569878681e-09);
  pc = __builtin_fma(pc, r2, -2.755731922398589e-07);
  pc = __builtin_fma(pc, r2, 2.48015873015873e-05);
  pc = __builtin_fma(pc, r2, -0.001388888888888889);
  pc = __builtin_fma(pc, r2, 0.041666666666666664);
  pc = __builtin_fma(pc, r2, -0.5);
  double const cs = __builtin_fma(r2, pc, 1.0);
  // quadrant q: (sin, cos) of x = (s, c), (c, -s), (-s, -c), (-c, s)
  double const s1 = (q & 1) ? cs : sn;
  double const c1 = (q & 1) ? sn : cs;
  *s_out = (q & 2) ? -s1 : s1;
  *c_out = ((q + 1) & 2) ? -c1 : c1;
}

// The same sine and cosine with the polynomials in Estrin's form (powers r^2, r^4, r^8
// shared by both): 4 dependent levels after r^2 instead of Horner's 8, for callers whose
// time is one long dependency chain (the RK4 field evaluation); rounding differs from
// sincos_fp64 by about an ulp.
SRS_FN void sincos_fp64_short(double x, double* s_out, double* c_out) {
  double const k = __builtin_rint(x * kTwoOverPi);
  double r = __builtin_fma(-k, kPio2Hi, x);
  r = __builtin_fma(-k, kPio2Lo, r);
  double const kb = k + 0x1.8p52;
  uint64_t bits;
  __builtin_memcpy(&bits, &kb, sizeof bits);
  int const q = (int)(bits & 3u);
  double const r2 = r * r, r4 = r2 * r2, r8 = r4 * r4;
  // P(r^2) = p0 + p1 r^2 + ... + p7 r^14 (the Taylor terms of sincos_fp64)
  double const s01 = __builtin_fma(-0.008333333333333333, r2, 0.16666666666666666);
  double const s23 = __builtin_fma(-2.7557319223985893e-06, r2, 0.0001984126984126984);
  double const s45 = __builtin_fma(-1.6059043836821613e-10, r2, 2.505210838544172e-08);
  double const s67 = __builtin_fma(-2.8114572543455206e-15, r2, 7.647163731819816e-13);
  double const ps = __builtin_fma(r8, __builtin_fma(r4, s67, s45), __builtin_fma(r4, s23, s01));
  double const sn = __builtin_fma(-(r * r2), ps, r);
  double const c01 = __builtin_fma(0.041666666666666664, r2, -0.5);
  double const c23 = __builtin_fma(2.48015873015873e-05, r2, -0.001388888888888889);
  double const c45 = __builtin_fma(2.08767569878681e-09, r2, -2.755731922398589e-07);
  double const c67 = __builtin_fma(4.779477332387385e-14, r2, -1.1470745597729725e-11);
  double const pc = __builtin_fma(r8, __builtin_fma(r4, c67, c45), __builtin_fma(r4, c23, c01));
  double const cs = __builtin_fma(r2, pc, 1.0);
  double const s1 = (q & 1) ? cs : sn;
  double const c1 = (q & 1) ? sn : cs;
  *s_out = (q & 2) ? -s1 : s1;
  *c_out = ((q + 1) & 2) ? -c1 : c1;
}

// Correctly rounded quotient a/b from the correctly rounded reciprocal y = RN(1/b):
// q = RN(a*y), r = a - b*q (exact with an FMA), RN(q + r*y) = RN(a/b) (Markstein's
// correction step).  Three operations instead of a full IEEE division; the
// reference rounds every a/b it computes, so the quotients must match it bit for bit.
SRS_FN double div_rn(double a, double b, double y) {
  double const q = a * y;
  double const r = __builtin_fma(-q, b, a);
  return __builtin_fma(r, y, q);
}
constexpr double kInvC = 0x1.ca726eb25f9dbp-29;  // RN(1/c)

// RN(sqrt(x)) for 2^-900 <= x <= 2^1000 from an estimate y0 of 1/sqrt(x) with relative
// error below 2^-16 (the hardware reciprocal square root is far closer).  With
// g = x y0 ~ sqrt(x) and h = y0/2 ~ 1/(2 sqrt(x)), e = 1/2 - h g is, to first order,
// the relative error of both; g += g e, h += h e squares it (Goldschmidt), and each
// g += (x - g^2) h (the residual exact with an FMA) is a Newton step: the error falls
// far below half an ulp before the last one rounds.  No scaling or special cases:
// in this range no intermediate is subnormal or overflows (the squared distances of
// the flux terms, from ~1e-30 to ~1e30 m^2).
SRS_FN double sqrt_rn(double x, double y0) {
  double g = x * y0;
  double h = 0.5 * y0;
  double const e = __builtin_fma(-h, g, 0.5);
  g = __builtin_fma(g, e, g);
  h = __builtin_fma(h, e, h);
  double d = __builtin_fma(-g, g, x);
  g = __builtin_fma(d, h, g);
  d = __builtin_fma(-g, g, x);
  return __builtin_fma(d, h, g);
}

// Reduce an FP64 phase to r in [-pi, pi] (two FMAs, same argument as above with
// 2*pi) for the single-precision trig of the mixed form.
constexpr double kInvTwoPi = 0x1.45f306dc9c883p-3;
constexpr double kTwoPiHi = 0x1.921fb54442d18p+2;
constexpr double kTwoPiLo = 0x1.1a62633145c07p-52;

SRS_FN double reduce_two_pi(double x) {
  double const k = __builtin_rint(x * kInvTwoPi);
  double r = __builtin_fma(-k, kTwoPiHi, x);
  return __builtin_fma(-k, kTwoPiLo, r);
}

}  // namespace srs
