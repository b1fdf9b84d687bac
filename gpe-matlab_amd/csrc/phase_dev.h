// Pointwise exponentials of the split step: e^{x} and (cos y, sin y) in FP64.
//
// Method: Cody-Waite additive argument reduction (W. J. Cody and W. Waite,
// "Software Manual for the Elementary Functions", 1980) followed by the Taylor
// series of exp / sin / cos truncated where the remainder is below 1e-17
// relative on the reduced interval, evaluated by Horner's rule with FMA.
//   sincos: y = k pi/2 + r, |r| <= pi/4, pi/2 split in three parts (33 + 33 + 53
//           significant bits) so k*P1 and k*P2 are exact for |k| < 2^20; sin to
//           r^17, cos to r^16.
//   exp:    x = k ln2 + r, |r| <= ln2/2, ln2 split in 32 + 53 bits; series to
//           r^13; result scaled by 2^k.
// Both are branch-free: NaN and Inf propagate (sin/cos of +-Inf is NaN, exp
// saturates to 0 / Inf beyond +-800).  Results are within about 1 ulp of the
// correctly rounded value for |y| < 2^20 pi/2 (about 1.6e6 rad, far beyond any
// physical split-step phase) and any x; past that range the pi/2 reduction
// loses about log2(|k|) - 20 bits of the reduced argument.  Checked against
// numpy on the GPU in the tests.
#pragma once
#include <hip/hip_runtime.h>

namespace gpe {

__device__ __forceinline__ void sincos_cw(double y, double* s, double* c) {
  const double k = rint(y * 0.63661977236758138);   // 2/pi
  double r = fma(-k, 1.5707963267341256, y);         // P1
  r = fma(-k, 6.077100506303966e-11, r);             // P2
  r = fma(-k, 2.0222662487959506e-21, r);            // P3
  const double z = r * r;
  // sin r = r - r^3/3! + r^5/5! - ... up to r^17
  double a = 1.0 / 355687428096000.0;                //  1/17!
  a = fma(a, z, -1.0 / 1307674368000.0);             // -1/15!
  a = fma(a, z, 1.0 / 6227020800.0);                 //  1/13!
  a = fma(a, z, -1.0 / 39916800.0);                  // -1/11!
  a = fma(a, z, 1.0 / 362880.0);                     //  1/9!
  a = fma(a, z, -1.0 / 5040.0);                      // -1/7!
  a = fma(a, z, 1.0 / 120.0);                        //  1/5!
  a = fma(a, z, -1.0 / 6.0);                         // -1/3!
  const double sn = fma(r * z, a, r);
  // cos r = 1 - r^2/2! + r^4/4! - ... up to r^16
  double b = 1.0 / 20922789888000.0;                 //  1/16!
  b = fma(b, z, -1.0 / 87178291200.0);               // -1/14!
  b = fma(b, z, 1.0 / 479001600.0);                  //  1/12!
  b = fma(b, z, -1.0 / 3628800.0);                   // -1/10!
  b = fma(b, z, 1.0 / 40320.0);                      //  1/8!
  b = fma(b, z, -1.0 / 720.0);                       // -1/6!
  b = fma(b, z, 1.0 / 24.0);                         //  1/4!
  b = fma(b, z, -0.5);                               // -1/2!
  const double cs = fma(z, b, 1.0);
  const int qd = ((int)k) & 3;
  const double s0 = (qd & 1) ? cs : sn;
  const double c0 = (qd & 1) ? sn : cs;
  *s = (qd & 2) ? -s0 : s0;
  *c = ((qd + 1) & 2) ? -c0 : c0;
}

__device__ __forceinline__ double exp_cw(double x) {
  x = x > 800.0 ? 800.0 : (x < -800.0 ? -800.0 : x);  // NaN passes through
  const double k = rint(x * 1.4426950408889634);    // log2(e)
  double r = fma(-k, 0.69314718060195446, x);        // ln2 high part (32 bits)
  r = fma(-k, -4.2009150726810846e-11, r);           // ln2 low part
  double p = 1.0 / 6227020800.0;                     // 1/13!
  p = fma(p, r, 1.0 / 479001600.0);                  // 1/12!
  p = fma(p, r, 1.0 / 39916800.0);                   // 1/11!
  p = fma(p, r, 1.0 / 3628800.0);                    // 1/10!
  p = fma(p, r, 1.0 / 362880.0);                     // 1/9!
  p = fma(p, r, 1.0 / 40320.0);                      // 1/8!
  p = fma(p, r, 1.0 / 5040.0);                       // 1/7!
  p = fma(p, r, 1.0 / 720.0);                        // 1/6!
  p = fma(p, r, 1.0 / 120.0);                        // 1/5!
  p = fma(p, r, 1.0 / 24.0);                         // 1/4!
  p = fma(p, r, 1.0 / 6.0);                          // 1/3!
  p = fma(p, r, 0.5);                                // 1/2!
  p = fma(p, r, 1.0);
  p = fma(p, r, 1.0);
  return ldexp(p, (int)k);
}

// v * exp(er + i ei), with w = exp(E) formed first and then w .* v, the order of
// exp(...) .* phi in the reference.
__device__ __forceinline__ double2 mul_cexp(double2 v, double er, double ei, bool has_re) {
  double s, c;
  sincos_cw(ei, &s, &c);
  const double m = has_re ? exp_cw(er) : 1.0;
  const double wr = m * c, wi = m * s;
  return make_double2(wr * v.x - wi * v.y, wr * v.y + wi * v.x);
}

}  // namespace gpe
