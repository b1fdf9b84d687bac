// Two-component x-sweeps of GPEtaskSpinor for one FFT length (-DGPE_N=<n>).
//
// SP_PR : inverse x-FFT of both components, the fused potential/interaction
//         phase and 2x2 Rabi rotation (lib/@GPEtaskSpinor/solve_split.m:77-91),
//         forward x-FFT of both components.
// SP_KEX: reductions of the block-end chemical potential on the real-space
//         state (solve_split.m:100-115, GPEtaskSpinor.m:55-92) plus the x part of
//         the kinetic energy of both components.
// SP_ADJ: adjoint of SP_PR for gamma = 0 and Omega(x) = om0 exp(i kw x): inverse
//         x-FFT of both adjoint components, the transposed step linearised at the
//         stored pre-step state, forward x-FFT, and the per-work-group sums of
//         dL/d om0 and dL/d kw.
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
    if (p.ust1 && active) {   // keep the pre-step state for the adjoint's segment recomputation
#pragma unroll
      for (int m = 0; m < E; ++m) {
        p.ust1[base + (unsigned)m * step] = u[m];
        p.ust2[base + (unsigned)m * step] = w[m];
      }
    }
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
  } else if constexpr (SEQ == SP_ADJ) {
    // u, w hold the adjoint components (x in k space); ust1/ust2 hold the pre-step state
    fft_pencil<PL, +1, C, COLFAST>(u, t, c, lds, twl);
    fft_pencil<PL, +1, C, COLFAST>(w, t, c, lds, twl);
    const double mu = p.st->mu;
    const double h = p.di;   // gamma = 0: dt = i h
    double sn, cs;
    sincos_cw(h * p.om0, &sn, &cs);
    double g0 = 0.0, gk = 0.0;
#pragma unroll
    for (int m = 0; m < E; ++m) {
      const unsigned idx = base + (unsigned)m * step;
      const double2 s1 = p.ust1[idx], s2 = p.ust2[idx];
      const double n1 = cabs2(s1), n2 = cabs2(s2);
      const double a1 = ((p.V1[idx] - mu) + p.g11 * n1) + p.g12 * n2;
      const double a2 = ((p.V2[idx] - mu) + p.g21 * n1) + p.g22 * n2;
      // R = [[c, -i s e], [-i s conj(e), c]] with c = cos(h om0), s = sin(h om0),
      // e = exp(i kw X) (solve_split.m:61-65, 90-91 for dt = i h)
      const double X = p.xc[t + T * m];
      double ex, ey;
      sincos_cw(p.kw * X, &ey, &ex);
      const double2 sinomp = make_double2(-sn * ey, sn * ex);    // i s e
      const double2 sinomm = make_double2(sn * ey, sn * ex);     // i s conj(e)
      const double2 cosom = make_double2(cs, 0.0);
      // forward map of this point: y = R s, w = e^{-i theta} y
      const double2 y1 = csub(cmul(cosom, s1), cmul(sinomp, s2));
      const double2 y2 = csub(cmul(cosom, s2), cmul(sinomm, s1));
      const double2 w1 = mul_cexp(y1, 0.0, -a1 * h, false);
      const double2 w2 = mul_cexp(y2, 0.0, -a2 * h, false);
      const double2 l1 = u[m], l2 = w[m];
      const double2 m1 = mul_cexp(l1, 0.0, a1 * h, false);   // e^{i theta_1} lambda_1
      const double2 m2 = mul_cexp(l2, 0.0, a2 * h, false);
      const double q1 = l1.x * w1.y - l1.y * w1.x;           // Im(conj(lambda) w)
      const double q2 = l2.x * w2.y - l2.y * w2.x;
      const double A1 = h * (p.g11 * q1 + p.g21 * q2);
      const double A2 = h * (p.g12 * q1 + p.g22 * q2);
      // lambda_u = R^H m + 2 A s
      const double2 nm = make_double2(-sinomm.x, -sinomm.y), np_ = make_double2(-sinomp.x, -sinomp.y);
      const double2 lu1 = cadd(cadd(cmulc(m1, cosom), cmulc(m2, nm)), cscale(s1, 2.0 * A1));
      const double2 lu2 = cadd(cadd(cmulc(m1, np_), cmulc(m2, cosom)), cscale(s2, 2.0 * A2));
      // dL/dp += Re m^H (dR/dp) s:  dR/dom0 = [[-h s, -i h c e], [-i h c conj(e), -h s]],
      // dR/dkw = [[0, s X e], [-s X conj(e), 0]]
      const double2 d12 = make_double2(h * cs * ey, -h * cs * ex);
      const double2 d21 = make_double2(-h * cs * ey, -h * cs * ex);
      const double2 t01 = cadd(cscale(s1, -h * sn), cmul(d12, s2));
      const double2 t02 = cadd(cmul(d21, s1), cscale(s2, -h * sn));
      g0 += (m1.x * t01.x + m1.y * t01.y) + (m2.x * t02.x + m2.y * t02.y);
      const double2 k12 = make_double2(sn * X * ex, sn * X * ey);
      const double2 k21 = make_double2(-sn * X * ex, sn * X * ey);
      const double2 tk1 = cmul(k12, s2), tk2 = cmul(k21, s1);
      gk += (m1.x * tk1.x + m1.y * tk1.y) + (m2.x * tk2.x + m2.y * tk2.y);
      u[m] = lu1;
      w[m] = lu2;
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
    if (!active) {
      g0 = 0.0;
      gk = 0.0;
    }
    const double r0 = block_sum(g0, red);
    const double r1 = block_sum(gk, red);
    if (tid == 0) {   // accumulated over the steps in a fixed order
      p.gacc[blockIdx.x] += r0;
      p.gacc[p.maxwg + blockIdx.x] += r1;
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
  else if (seq == SP_ADJ)
    hipLaunchKernelGGL((spinor_kernel<N, C, SP_ADJ>), dim3((unsigned)nb), dim3(256), 0, s, p);
  else
    return hipErrorInvalidValue;
  return hipGetLastError();
}

}  // namespace gpe
