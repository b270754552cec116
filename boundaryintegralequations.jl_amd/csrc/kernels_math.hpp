// FP64 device math for the assembly kernels (gfx950).
#pragma once

#include <hip/hip_runtime.h>

#include "bie_common.hpp"

namespace bie {

// sin and cos of x for |x| < 2^26 by Cody-Waite reduction modulo pi/2 with two fused
// multiply-add terms (the product n * pio2_hi is exact inside the FMA, the dropped tail of
// pi/2 is ~1e-33, so the reduced argument carries only two roundings of ~5.5e-17), then
// degree-11 (sin) and degree-12 (cos) polynomials on |r| <= pi/4 whose coefficients were
// fitted by a Remez exchange in 60-digit arithmetic (csrc/fit_sincos.py; max error of the
// polynomials 1.4e-17 and 8.7e-19).  Absolute error of the results ~2e-16, which is what
// e^{ikr} needs.  Larger arguments take the library sincos.
__device__ __forceinline__ void sincos_f64(double x, double* s, double* c) {
  if (fabs(x) >= 67108864.0) {  // 2^26
    sincos(x, s, c);
    return;
  }
  constexpr double kTwoOverPi = 0.63661977236758138243;
  constexpr double kPio2Hi = 1.5707963267948965580;   // pi/2 rounded to double
  constexpr double kPio2Lo = 6.1232339957367658e-17;  // pi/2 - kPio2Hi rounded to double
  const double n = rint(x * kTwoOverPi);
  double r = fma(-n, kPio2Hi, x);
  r = fma(-n, kPio2Lo, r);
  const int q = (int)n;
  const double t = r * r;
  double S = fma(t, 1.5918142569704652e-10, -2.505113204972727e-08);
  S = fma(t, S, 2.7557316103657487e-06);
  S = fma(t, S, -0.00019841269836761008);
  S = fma(t, S, 0.00833333333333095);
  S = fma(t, S, -0.16666666666666666);
  const double sr = fma(r * t, S, r);
  double C = fma(t, -1.1382639805756885e-11, 2.0876146382220145e-09);
  C = fma(t, C, -2.7557317272344146e-07);
  C = fma(t, C, 2.480158729876704e-05);
  C = fma(t, C, -0.0013888888888887398);
  C = fma(t, C, 0.041666666666666664);
  const double cr = fma(t * t, C, fma(-0.5, t, 1.0));
  double ss = (q & 1) ? cr : sr;
  double cc = (q & 1) ? sr : cr;
  if (q & 2) ss = -ss;
  if ((q + 1) & 2) cc = -cc;
  *s = ss;
  *c = cc;
}

// sin and cos of x for |x| < 2^26 from the table tab[j] = (cos, sin)(j pi / 64), j < 128,
// in LDS: x = n pi / 64 + d with n = rint(64 x / pi) and the two-term reduction above
// (|d| <= pi / 128), j = n mod 128, then sin x = s_j cos d + c_j sin d and
// cos x = c_j cos d - s_j sin d with the Taylor polynomials sin d = d - d^3/3! + d^5/5! -
// d^7/7! and cos d = 1 - d^2/2! + d^4/4! - d^6/6! (truncation below 9e-21 and 3.3e-18 on
// |d| <= pi / 128).  Absolute error ~2e-16 with the table entries correctly rounded; no
// quadrant selection and three polynomial constants per function.  A NaN x reads entry 0
// and returns NaN; larger arguments take the library sincos.
__device__ __forceinline__ void sincos_tab(double x, const double2* __restrict__ tab,
                                           double* s, double* c) {
  if (fabs(x) >= 67108864.0) {  // 2^26
    sincos(x, s, c);
    return;
  }
  constexpr double kInvStep = 20.371832715762602978;   // 64 / pi
  constexpr double kStepHi = 1.5707963267948965580 / 32;   // pi/2 rounded, / 32 (exact)
  constexpr double kStepLo = 6.1232339957367658e-17 / 32;  // its remainder, / 32 (exact)
  const double n = rint(x * kInvStep);
  double d = fma(-n, kStepHi, x);
  d = fma(-n, kStepLo, d);
  const double2 t = tab[(int)n & (kTrigN - 1)];
  const double d2 = d * d;
  const double sp = fma(d2, fma(d2, -1.0 / 5040.0, 1.0 / 120.0), -1.0 / 6.0);
  const double sd = fma(d * d2, sp, d);
  const double cp = fma(d2, fma(d2, -1.0 / 720.0, 1.0 / 24.0), -0.5);
  const double cd = fma(d2, cp, 1.0);
  *s = fma(t.y, cd, t.x * sd);
  *c = fma(t.x, cd, -(t.y * sd));
}

// sincos_tab without the large-argument branch, for |x| < kTabFastMaxArg (bie_common.hpp):
// n = rint(64 x / pi) is formed by adding 1.5 * 2^52 to x * 64 / pi inside one fused multiply-add (the sum's
// unit in the last place is 1, so it is rounded to an integer there), the reduction uses
// n = sum - 1.5 * 2^52 (exact), and the table index n mod 128 is the low bits of the sum's
// mantissa (two's complement for negative n), so no float-to-integer conversion is needed.
// Same reduction, table and polynomials as sincos_tab otherwise; NaN and infinite x give
// NaN (the index is masked into the table whatever the bits are).
__device__ __forceinline__ void sincos_tab_fast(double x, const double2* __restrict__ tab,
                                                double* s, double* c) {
  constexpr double kInvStep = 20.371832715762602978;   // 64 / pi
  constexpr double kStepHi = 1.5707963267948965580 / 32;   // pi/2 rounded, / 32 (exact)
  constexpr double kStepLo = 6.1232339957367658e-17 / 32;  // its remainder, / 32 (exact)
  constexpr double kShift = 6755399441055744.0;            // 1.5 * 2^52
  const double sh = fma(x, kInvStep, kShift);
  const double n = sh - kShift;
  double d = fma(-n, kStepHi, x);
  d = fma(-n, kStepLo, d);
  const double2 t = tab[__double2loint(sh) & (kTrigN - 1)];
  const double d2 = d * d;
  const double sp = fma(d2, fma(d2, -1.0 / 5040.0, 1.0 / 120.0), -1.0 / 6.0);
  const double sd = fma(d * d2, sp, d);
  const double cp = fma(d2, fma(d2, -1.0 / 720.0, 1.0 / 24.0), -0.5);
  const double cd = fma(d2, cp, 1.0);
  *s = fma(t.y, cd, t.x * sd);
  *c = fma(t.x, cd, -(t.y * sd));
}

// 1/sqrt(x) from the hardware estimate y0 (v_rsq_f64; relative error below 6e-8 on
// gfx950, measured over 4M arguments spanning e^-20..e^20) refined by one step of the
// third-order iteration y = y0 (1 + e/2 + 3 e^2/8), e = 1 - x y0^2, whose truncation
// (5/16) e^3 < 1e-22 leaves only the roundings of e and of the last fused multiply-add
// (~2e-16): five instructions where two Newton steps take seven.
__device__ __forceinline__ double rsqrt_nr(double x) {
  const double y = __builtin_amdgcn_rsq(x);
  const double e = fma(-x * y, y, 1.0);
  return fma(y * e, fma(e, 0.375, 0.5), y);
}

}  // namespace bie
