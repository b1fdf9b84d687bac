// Two-component x-sweeps of GPEtaskSpinor for one FFT length (-DGPE_N=<n>).
//
// SP_PR : inverse x-FFT of both components, the fused potential/interaction
//         phase and 2x2 Rabi rotation (lib/@GPEtaskSpinor/solve_split.m:77-91),
//         forward x-FFT of both components.
// SP_KEX: reductions of the block-end chemical potential on the real-space
//         state (solve_split.m:100-115, GPEtaskSpinor.m:55-92) plus the x part of
//         the kinetic energy of both components.
#include "fft_dev.h"
#include "phase_dev.h"
#include "reduce_dev.h"
#include "sweep.h"

#ifndef GPE_N
#error "compile with -DGPE_N=<pencil length>"
#endif

namespace gpe {

// cosh and sinh of the complex number z = (x, y).
__device__ __forceinline__ void ccoshsinh(double x, double y, double2* ch, double2* sh) {
  double s, c;
  sincos_cw(y, &s, &c);
  const double e = exp_cw(x);
  const double ei = 1.0 / e;
  const double chx = 0.5 * (e + ei), shx = 0.5 * (e - ei);
  *ch = make_double2(chx * c, shx * s);
  *sh = make_double2(shx * c, chx * s);
}

template <int N, int C, int SEQ>
__global__ void __launch_bounds__(256) spinor_kernel(const SpinorParams p) {
  using PL = Plan<N>;
  constexpr int T = PL::T, E = PL::E;
  constexpr bool COLFAST = true;
  __shared__ double lds[Lds<N, C, COLFAST>::SIZE];
  __shared__ double2 twl[N];
  __shared__ double red[4];

  const int tid = threadIdx.x;
  const int c = tid % C;
  const int t = tid / C;
  for (int i = tid; i < N; i += 256) twl[i] = p.tw[i];

  const unsigned cgu = blockIdx.x * (unsigned)C + (unsigned)c;
  const bool active = cgu < p.ncols;
  const unsigned cg = active ? cgu : p.ncols - 1u;
  const unsigned base = (cg % p.A) * p.sa + (cg / p.A) * p.sb + (unsigned)t * p.stride;
  const unsigned step = (unsigned)T * p.stride;

  double2 u[E], w[E];
#pragma unroll
  for (int m = 0; m < E; ++m) {
    u[m] = p.psi1[base + (unsigned)m * step];
    w[m] = p.psi2[base + (unsigned)m * step];
  }

  if constexpr (SEQ == SP_PR) {
    fft_pencil<PL, +1, C, COLFAST>(u, t, c, lds, twl);
    fft_pencil<PL, +1, C, COLFAST>(w, t, c, lds, twl);
    const double mu = p.st->mu;
    const bool has_re = p.dr != 0.0;
#pragma unroll
    for (int m = 0; m < E; ++m) {
      const unsigned idx = base + (unsigned)m * step;
      const double n1 = cabs2(u[m]), n2 = cabs2(w[m]);
      const double a1 = ((p.V1[idx] - mu) + p.g11 * n1) + p.g12 * n2;   // solve_split.m:78-83
      const double a2 = ((p.V2[idx] - mu) + p.g21 * n1) + p.g22 * n2;
      double2 cosom, sinomm, sinomp;
      if (p.nlincpl != 0.0) {                                           // solve_split.m:84-89
        const double2 om = p.cpl ? p.cpl[idx] : p.cpl_s;
        const double2 t3 = make_double2(om.x + p.nlincpl * (n1 + n2), om.y);
        const double2 z = cmul(make_double2(p.dr, p.di), t3);
        ccoshsinh(z.x, z.y, &cosom, &sinomm);
        sinomp = sinomm;
      } else if (p.cpl) {                                               // solve_split.m:61-65, array Omega
        const double2 om = p.cpl[idx];
        const double r = hypot(om.x, om.y);
        double2 sh;
        ccoshsinh(p.dr * r, p.di * r, &cosom, &sh);
        const double2 ph = r > 0.0 ? make_double2(om.x / r, om.y / r) : make_double2(1.0, 0.0);   // e^{i angle}
        sinomm = cmulc(sh, ph);
        sinomp = cmul(sh, ph);
      } else {
        cosom = p.cosom;
        sinomm = p.sinomm;
        sinomp = p.sinomp;
      }
      const double2 r1 = csub(cmul(cosom, u[m]), cmul(sinomp, w[m]));   // solve_split.m:90
      const double2 r2 = csub(cmul(cosom, w[m]), cmul(sinomm, u[m]));   // solve_split.m:91
      u[m] = mul_cexp(r1, -a1 * p.dr, -a1 * p.di, has_re);
      w[m] = mul_cexp(r2, -a2 * p.dr, -a2 * p.di, has_re);
    }
    fft_pencil<PL, -1, C, COLFAST>(u, t, c, lds, twl);
    fft_pencil<PL, -1, C, COLFAST>(w, t, c, lds, twl);
    if (active) {
#pragma unroll
      for (int m = 0; m < E; ++m) {
        p.psi1[base + (unsigned)m * step] = u[m];
        p.psi2[base + (unsigned)m * step] = w[m];
      }
    }
  } else {
    double accA = 0.0, accB = 0.0, accC1 = 0.0, accC12 = 0.0, accC2 = 0.0, accX = 0.0, k1 = 0.0, k2 = 0.0;
    if (active) {
#pragma unroll
      for (int m = 0; m < E; ++m) {
        const unsigned idx = base + (unsigned)m * step;
        const double n1 = cabs2(u[m]), n2 = cabs2(w[m]);
        const double2 om = p.cpl ? p.cpl[idx] : p.cpl_s;
        accA += n1 + n2;
        accB += p.V1[idx] * n1 + p.V2[idx] * n2;
        accC1 += n1 * n1;
        accC12 += n1 * n2;
        accC2 += n2 * n2;
        const double2 ow = cmul(om, w[m]);
        accX += u[m].x * ow.x + u[m].y * ow.y;   // Re conj(psi1) Omega psi2
      }
    }
    fft_pencil<PL, -1, C, COLFAST>(u, t, c, lds, twl);
    fft_pencil<PL, -1, C, COLFAST>(w, t, c, lds, twl);
    if (active) {
#pragma unroll
      for (int m = 0; m < E; ++m) {
        k1 += p.keax[t + T * m] * cabs2(u[m]);
        k2 += p.keax[t + T * m] * cabs2(w[m]);
      }
    }
    const double vals[8] = {accA, accB, accC1, accC12, accC2, accX, k1, k2};
    const int slots[8] = {SL_A, SL_B, SL_C, SL_C12, SL_C2, SL_X, SL_KX, SL_KX2};
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const double r = block_sum(vals[q], red);
      if (tid == 0) p.partials[slots[q] * p.maxwg + blockIdx.x] = r;
    }
  }
}

#define GPE_CAT2(a, b) a##b
#define GPE_CAT(a, b) GPE_CAT2(a, b)

hipError_t GPE_CAT(launch_spinor_sweep_, GPE_N)(int seq, const SpinorParams& p, hipStream_t s, int* nblocks) {
  constexpr int N = GPE_N;
  constexpr int C = 256 / Plan<N>::T;
  const long long nb = ((long long)p.ncols + C - 1) / C;
  if (nblocks) *nblocks = (int)nb;
  if (nb <= 0) return hipSuccess;
  if (seq == SP_PR)
    hipLaunchKernelGGL((spinor_kernel<N, C, SP_PR>), dim3((unsigned)nb), dim3(256), 0, s, p);
  else if (seq == SP_KEX)
    hipLaunchKernelGGL((spinor_kernel<N, C, SP_KEX>), dim3((unsigned)nb), dim3(256), 0, s, p);
  else
    return hipErrorInvalidValue;
  return hipGetLastError();
}

}  // namespace gpe
