// sin/cos of phase arguments, written for this library from textbook methods:
// Cody-Waite reduction by pi/2 with a three-part constant (W. J. Cody and
// W. Waite, "Software Manual for the Elementary Functions", 1980), then the
// Taylor series of sin and cos on |r| <= pi/4 truncated where the next term is
// below half an ulp (FP64: through r^17 / r^18, FP32: r^9 / r^10).  Each part
// k*C_i is formed exactly inside an FMA and C1 + C2 + C3 equals pi/2 to ~2^-120,
// so the reduced argument's absolute error is ~1e-16 for |phi| < 1e6 and stays
// below ulp(phi) (the precision of the phase itself) for every |phi| < 2^40.
// Larger phases (a time step that
// turns a phase by more than 1e12 rad is not resolved by any grid) give NaN, so the
// per-element code stays a short branch-free sequence, which matters when a lane
// evaluates 16 phases back to back.
#pragma once
#include <hip/hip_runtime.h>

namespace wh {

constexpr double kTwoOverPi = 0.6366197723675814;
constexpr double kPio2C1 = 1.5707963267341256;
constexpr double kPio2C2 = 6.077100506303966e-11;
constexpr double kPio2C3 = 2.0222662487959506e-21;

// phi = k*pi/2 + r, |r| <= pi/4 (+ rounding)
__host__ __device__ __forceinline__ double reduce_pio2(double phi, int& q) {
  const double k = rint(phi * kTwoOverPi);
  double r = fma(-k, kPio2C1, phi);
  r = fma(-k, kPio2C2, r);
  r = fma(-k, kPio2C3, r);
  q = int(k - 4.0 * floor(k * 0.25));  // k mod 4, exact for |k| < 2^53
  return r;
}

constexpr double kMaxPhase = 1099511627776.0;  // 2^40

__host__ __device__ __forceinline__ void sincos_fp64(double phi, double* s, double* c) {
  if (!(fabs(phi) < kMaxPhase)) phi = __builtin_nan("");
  int q;
  const double r = reduce_pio2(phi, q);
  const double r2 = r * r;
  // sin r = r + r^3 * sum_n (-1)^(n+1) r^(2n) / (2n+3)!
  double ps = 2.8114572543455206e-15;   // 1/17!
  ps = fma(ps, r2, -7.647163731819816e-13);  // -1/15!
  ps = fma(ps, r2, 1.6059043836821613e-10);  // 1/13!
  ps = fma(ps, r2, -2.505210838544172e-08);  // -1/11!
  ps = fma(ps, r2, 2.7557319223985893e-06);  // 1/9!
  ps = fma(ps, r2, -1.984126984126984e-04);  // -1/7!
  ps = fma(ps, r2, 8.333333333333333e-03);   // 1/5!
  ps = fma(ps, r2, -1.6666666666666666e-01); // -1/3!
  const double sn = fma(ps * r2, r, r);
  // cos r = 1 + r^2 * sum_n (-1)^(n+1) r^(2n) / (2n+2)!
  double pc = -1.5619206968586225e-16;  // -1/18!
  pc = fma(pc, r2, 4.779477332387385e-14);   // 1/16!
  pc = fma(pc, r2, -1.1470745597729725e-11); // -1/14!
  pc = fma(pc, r2, 2.08767569878681e-09);    // 1/12!
  pc = fma(pc, r2, -2.755731922398589e-07);  // -1/10!
  pc = fma(pc, r2, 2.48015873015873e-05);    // 1/8!
  pc = fma(pc, r2, -1.388888888888889e-03);  // -1/6!
  pc = fma(pc, r2, 4.1666666666666664e-02);  // 1/4!
  pc = fma(pc, r2, -0.5);
  const double cs = fma(pc, r2, 1.0);
  // quadrant q: (sin, cos) = (sn, cs), (cs, -sn), (-sn, -cs), (-cs, sn), branch-free
  const double a = (q & 1) ? cs : sn, b = (q & 1) ? sn : cs;
  *s = (q & 2) ? -a : a;
  *c = ((q + 1) & 2) ? -b : b;
}

// FP64 argument, FP64 reduction, FP32 polynomials.
__host__ __device__ __forceinline__ void sincos_fp32(double phi, float* s, float* c) {
  if (!(fabs(phi) < kMaxPhase)) phi = __builtin_nan("");
  int q;
  const float r = float(reduce_pio2(phi, q));
  const float r2 = r * r;
  float ps = 2.7557319e-06f;         // 1/9!
  ps = fmaf(ps, r2, -1.9841270e-04f);  // -1/7!
  ps = fmaf(ps, r2, 8.3333333e-03f);   // 1/5!
  ps = fmaf(ps, r2, -1.6666667e-01f);  // -1/3!
  const float sn = fmaf(ps * r2, r, r);
  float pc = -2.7557319e-07f;        // -1/10!
  pc = fmaf(pc, r2, 2.4801587e-05f);   // 1/8!
  pc = fmaf(pc, r2, -1.3888889e-03f);  // -1/6!
  pc = fmaf(pc, r2, 4.1666668e-02f);   // 1/4!
  pc = fmaf(pc, r2, -0.5f);
  const float cs = fmaf(pc, r2, 1.0f);
  const float a = (q & 1) ? cs : sn, b = (q & 1) ? sn : cs;
  *s = (q & 2) ? -a : a;
  *c = ((q + 1) & 2) ? -b : b;
}

}  // namespace wh
