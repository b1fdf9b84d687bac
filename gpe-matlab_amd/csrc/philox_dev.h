// Counter-based normal variates for the stochastic GPE.
//
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy as
// 1, 2, 3", SC'11) with counter (point, step_lo, step_hi, 0) and key (seed_lo,
// seed_hi); the four output words make two uniforms u = ((w_a << 32 | w_b) >> 11
// + 1/2) 2^-53 in (0, 1), and Box-Muller gives the real and imaginary parts of one
// complex normal variate: sqrt(-2 ln u1) (cos, sin)(2 pi u2).
#pragma once
#include <hip/hip_runtime.h>
#include "phase_dev.h"

namespace gpe {

__device__ __forceinline__ void philox4x32_10(unsigned (&c)[4], unsigned k0, unsigned k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    const unsigned long long p0 = 0xD2511F53ull * c[0];
    const unsigned long long p1 = 0xCD9E8D57ull * c[2];
    const unsigned hi0 = (unsigned)(p0 >> 32), lo0 = (unsigned)p0;
    const unsigned hi1 = (unsigned)(p1 >> 32), lo1 = (unsigned)p1;
    const unsigned n0 = hi1 ^ c[1] ^ k0, n2 = hi0 ^ c[3] ^ k1;
    c[0] = n0;
    c[1] = lo1;
    c[2] = n2;
    c[3] = lo0;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
}

__device__ __forceinline__ double u53(unsigned a, unsigned b) {
  const unsigned long long w = ((unsigned long long)a << 32) | b;
  return ((double)(w >> 11) + 0.5) * 0x1.0p-53;
}

// One complex standard normal variate (re, im independent N(0, 1)) of point idx at
// global step `step` for key `seed`.
__device__ __forceinline__ double2 complex_normal(unsigned idx, unsigned long long step, unsigned long long seed) {
  unsigned c[4] = {idx, (unsigned)step, (unsigned)(step >> 32), 0u};
  philox4x32_10(c, (unsigned)seed, (unsigned)(seed >> 32));
  const double u1 = u53(c[0], c[1]), u2 = u53(c[2], c[3]);
  const double r = sqrt(-2.0 * log(u1));
  double s, co;
  sincos_cw((2.0 * 3.141592653589793) * u2, &s, &co);
  return make_double2(r * co, r * s);
}

}  // namespace gpe
