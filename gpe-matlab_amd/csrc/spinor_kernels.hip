// Two-component x-sweeps of GPEtaskSpinor for one FFT length (-DGPE_N=<n>).
//
// Lane pairs: lane 2i holds component 1 and lane 2i+1 component 2 of the same
// pencil positions, so each lane transforms one component (E values, the
// register footprint of a one-component sweep) and the pointwise 2x2 step reads
// the other component from its partner lane with a DPP swap instead of holding
// both.  The two components' FFTs run side by side in two halves of LDS.
//
// SP_PR : inverse x-FFT of both components, the fused potential/interaction
//         phase and 2x2 Rabi rotation (lib/@GPEtaskSpinor/solve_split.m:77-91),
//         forward x-FFT of both components.
// SP_KEX: reductions of the block-end chemical potential on the real-space
//         state (solve_split.m:100-115, GPEtaskSpinor.m:55-92) plus the x part of
//         the kinetic energy of both components.
// SP_KEXS: the same sums from the stored x spectrum S = F_x psi / nx that the steps
//         leave between the sweeps (block ends inside an outer iteration): the x
//         kinetic part straight from S, then the inverse x-FFT for the real-space
//         sums; nothing is stored.
// SP_ITPA/B/C: the x-sweeps of one two-component imaginary-time iteration
//         (GPEtaskSpinor/groundstate_itp.m:49-88, GPEtask2comp/groundstate_itp.m:58-90):
//         A renormalises the real-space state, stores exp(-(V_k + sum_j g_kj n_j) dt)
//         from that density and transforms forward; B (after a kinetic half step)
//         transforms back, applies the stored factor (and the Rabi rotation) and
//         transforms forward; C (after the second half step) transforms back and
//         reduces each component's norm.
// SP_ADJ: adjoint of SP_PR for gamma = 0 and Omega(x) = om0 exp(i kw x): inverse
//         x-FFT of both adjoint components, the transposed step linearised at the
//         stored pre-step state, forward x-FFT, and the per-work-group sums of
//         dL/d om0 and dL/d kw.
#include <type_traits>

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

// cos(h r) and sin(h r) / r from z = (h r)^2, r = |Omega| (the Rabi factors for
// dt = i h): Taylor series in z up to (h r)^16, truncation below 1e-17 for z <= 0.6.
__device__ __forceinline__ void rabi_taylor(double h, double z, double* cs, double* f) {
  double c = 1.0 / 20922789888000.0;                  //  1/16!
  c = fma(c, z, -1.0 / 87178291200.0);                // -1/14!
  c = fma(c, z, 1.0 / 479001600.0);                   //  1/12!
  c = fma(c, z, -1.0 / 3628800.0);                    // -1/10!
  c = fma(c, z, 1.0 / 40320.0);                       //  1/8!
  c = fma(c, z, -1.0 / 720.0);                        // -1/6!
  c = fma(c, z, 1.0 / 24.0);                          //  1/4!
  c = fma(c, z, -0.5);                                // -1/2!
  *cs = fma(c, z, 1.0);
  double s = 1.0 / 355687428096000.0;                 //  1/17!
  s = fma(s, z, -1.0 / 1307674368000.0);              // -1/15!
  s = fma(s, z, 1.0 / 6227020800.0);                  //  1/13!
  s = fma(s, z, -1.0 / 39916800.0);                   // -1/11!
  s = fma(s, z, 1.0 / 362880.0);                      //  1/9!
  s = fma(s, z, -1.0 / 5040.0);                       // -1/7!
  s = fma(s, z, 1.0 / 120.0);                         //  1/5!
  s = fma(s, z, -1.0 / 6.0);                          // -1/3!
  *f = h * fma(s, z, 1.0);
}


// Coupling kinds of SP_PR, fixed per launch.
// CK_ARRAY_REAL: array Omega with dt = i di; CK_ARRAY_TAYLOR: the same with
// di^2 max|Omega|^2 <= 0.6 (rabi_maxz), where the Rabi factors are power series.
// CK_SCALAR_REAL: scalar Omega with dt = i di.  CK_NONE: no coupling (mixture).
enum CplKind : int {
  CK_SCALAR = 0, CK_ARRAY = 1, CK_NONLIN = 2, CK_ARRAY_REAL = 3, CK_ARRAY_TAYLOR = 4, CK_SCALAR_REAL = 5,
  CK_NONE = 6
};

#ifndef GPE_ADJ_WAVES
#define GPE_ADJ_WAVES 2
#endif

template <int N, int C2, int SEQ, int CK>
__device__ __forceinline__ void spinor_body(const SpinorParams& p) {
  using PL = Plan<N>;
  using L = Lds<N, C2, true>;
  constexpr int T = PL::T, E = PL::E;
  constexpr bool COLFAST = true;
  // The transposed step of the adjoint moves real and imaginary parts together through
  // LDS (half the LDS instructions and barriers; 2 x 24 KB at 192): measured at 192^3 it
  // spills 48 VGPRs instead of 70 and runs 9 % faster (352 -> 320 us per launch).  The forward step (same speed
  // either way, c5: 2437-2445 against 2448-2453 steps/s) and the block-end sums keep the
  // split exchanges, and so do the imaginary-time sweeps, which would lose a wave per
  // SIMD (registers and LDS).
  constexpr bool CPLX_LDS = SEQ == SP_ADJ;
  __shared__ typename std::conditional<CPLX_LDS, double2, double>::type lds[2 * L::SIZE];
  __shared__ double2 twl[N];
  __shared__ double2 exl[SEQ == SP_ADJ ? N : 1];   // SP_ADJ: exp(i kw x) per x index
  __shared__ double xl[SEQ == SP_ADJ ? N : 1];     // SP_ADJ: x per x index
  __shared__ double red[4];

  const int tid = threadIdx.x;
  const int k = tid & 1;           // component of this lane
  const int q = tid >> 1;
  const int c = q % C2;
  const int t = q / C2;
  auto* ldsk = lds + k * L::SIZE;
  // Lanes 2i and 2i+1 address the same slot of their component's half of LDS, and the
  // halves are whole 64-bank lines apart, so every exchange access would be a 2-way bank
  // conflict.  Component 2 XORs its slot index with 28 instead: within a 32-slot line
  // that moves it to another column group (bit 2) and row slot (bits 3-4), which takes
  // the partner lane to a bank of its own for every plan's exchange pattern, for groups
  // of 16 and of 32 lanes alike.  Measured at 192^3: 608 -> 0 conflict cycles per wave,
  // the same VALU count.
  static_assert(L::SIZE % 32 == 0, "the slot XOR stays inside whole 32-slot lines");
  // The adjoint's 16-byte exchanges run without it (320 us per launch at 192^3, 326 with).
  const unsigned swz = (k && SEQ != SP_ADJ) ? 28u : 0u;
  for (int i = tid; i < N; i += 256) twl[i] = p.tw[i];
  if constexpr (SEQ == SP_ADJ) {
    for (int i = tid; i < N; i += 256) {
      double ey, ex;
      sincos_cw(p.kw * p.xc[i], &ey, &ex);
      exl[i] = make_double2(ex, ey);
      xl[i] = p.xc[i];
    }
    __syncthreads();
  }

  const unsigned cgu = xcd_pair_tile(blockIdx.x, gridDim.x) * (unsigned)C2 + (unsigned)c;
  const bool active = cgu < p.ncols;   // equal on both lanes of a pair
  const unsigned cg = active ? cgu : p.ncols - 1u;
  const unsigned base = (cg % p.A) * p.sa + (cg / p.A) * p.sb + (unsigned)t * p.stride;
  const unsigned step = (unsigned)T * p.stride;
  double2* __restrict__ psik = k ? p.psi2 : p.psi1;
  const double* __restrict__ Vk = k ? p.V2 : p.V1;

  double2 v[E];
#pragma unroll
  for (int m = 0; m < E; ++m) v[m] = psik[base + (unsigned)m * step];

  if constexpr (SEQ == SP_PR) {
    // with a real time step and an array Omega, V_k and Omega are loaded with psi, before
    // the transform, so their latency overlaps it (measured: +2 % at 192^3 although the
    // registers allow 2 waves per SIMD instead of 3; the other forms keep 1 or 3 waves
    // and load them in the phase loop)
    constexpr bool pre_v = CK == CK_ARRAY_REAL || CK == CK_ARRAY_TAYLOR;
    constexpr bool pre_o = pre_v;
    double vpre[pre_v ? E : 1];
    double2 opre[pre_o ? E : 1];
    if constexpr (pre_v) {
#pragma unroll
      for (int m = 0; m < E; ++m) vpre[m] = Vk[base + (unsigned)m * step];
    }
    if constexpr (pre_o) {
#pragma unroll
      for (int m = 0; m < E; ++m) opre[m] = p.cpl[base + (unsigned)m * step];
    }
    fft_pencil<PL, +1, C2, COLFAST>(v, t, c, ldsk, twl, swz);
    if (p.ust1 && active) {   // keep the pre-step state for the adjoint's segment recomputation
      double2* __restrict__ us = k ? p.ust2 : p.ust1;
#pragma unroll
      for (int m = 0; m < E; ++m) us[base + (unsigned)m * step] = v[m];
    }
    const double mu = p.st->mu;
    constexpr bool real_dt = CK == CK_ARRAY_REAL || CK == CK_ARRAY_TAYLOR || CK == CK_SCALAR_REAL;
    const bool has_re = !real_dt && p.dr != 0.0;
    const double sgn = k ? 1.0 : -1.0;   // row 1: i f Omega, row 2: i f conj(Omega)
    // this lane's row of the interaction matrix: a_k = V_k - mu + g_k1 n1 + g_k2 n2
    const double gk1 = k ? p.g21 : p.g11, gk2 = k ? p.g22 : p.g12;
#pragma unroll
    for (int m = 0; m < E; ++m) {
      const unsigned idx = base + (unsigned)m * step;
      const double2 o = xchg(v[m]);
      const double nown = cabs2(v[m]), noth = cabs2(o);
      const double n1 = k ? noth : nown, n2 = k ? nown : noth;
      const double vkm = pre_v ? vpre[pre_v ? m : 0] : Vk[idx];
      const double a = ((vkm - mu) + gk1 * n1) + gk2 * n2;        // solve_split.m:78-83
      double2 cosom, soff;   // diagonal and this row's off-diagonal Rabi factor
      if constexpr (CK == CK_NONLIN) {                            // solve_split.m:84-89
        const double2 om = p.cpl ? p.cpl[idx] : p.cpl_s;
        const double2 t3 = make_double2(om.x + p.nlincpl * (n1 + n2), om.y);
        const double2 z = cmul(make_double2(p.dr, p.di), t3);
        ccoshsinh(z.x, z.y, &cosom, &soff);
      } else if constexpr (CK == CK_ARRAY) {                      // solve_split.m:61-65, array Omega
        const double2 om = pre_o ? opre[pre_o ? m : 0] : p.cpl[idx];
        const double r = hypot(om.x, om.y);
        double2 sh;
        ccoshsinh(p.dr * r, p.di * r, &cosom, &sh);
        const double2 ph = r > 0.0 ? make_double2(om.x / r, om.y / r) : make_double2(1.0, 0.0);   // e^{i angle}
        soff = k ? cmulc(sh, ph) : cmul(sh, ph);
      } else if constexpr (CK == CK_ARRAY_REAL || CK == CK_ARRAY_TAYLOR) {   // the same for dt = i di:
        const double2 om = pre_o ? opre[pre_o ? m : 0] : p.cpl[idx];   // cosh(dt r) = cos(di r), sinh = i sin
        double cs, f;                           // sin(di r) e^{+-i angle} = f Omega or f conj(Omega)
        if constexpr (CK == CK_ARRAY_TAYLOR) {
          rabi_taylor(p.di, p.di * p.di * (om.x * om.x + om.y * om.y), &cs, &f);
        } else {
          const double r = sqrt(om.x * om.x + om.y * om.y);
          double sn;
          sincos_cw(p.di * r, &sn, &cs);
          f = r > 0.0 ? sn / r : p.di;
        }
        cosom = make_double2(cs, 0.0);
        soff = make_double2(sgn * f * om.y, f * om.x);
      } else {
        cosom = p.cosom;
        soff = k ? p.sinomm : p.sinomp;
      }
      const double2 r = csub(cmul(cosom, v[m]), cmul(soff, o));   // solve_split.m:90-91
      v[m] = mul_cexp(r, -a * p.dr, -a * p.di, has_re);
    }
    fft_pencil<PL, -1, C2, COLFAST>(v, t, c, ldsk, twl, swz);
    const unsigned bs = opaque(base);   // the store offset is formed here, not kept live
    if (active) {
#pragma unroll
      for (int m = 0; m < E; ++m) psik[bs + (unsigned)m * step] = v[m];
    }
  } else if constexpr (SEQ == SP_ADJ) {
    // v holds this lane's adjoint component (x in k space); ust1/ust2 hold the
    // pre-step state.  Per point, with e_k = e (k = 0) or conj(e) (k = 1),
    // e = exp(i kw X), c = cos(h om0), s = sin(h om0) and z = i e_k s_other
    // (solve_split.m:61-65, 90-91 for dt = i h):
    //   forward row k    y = c s_k - s z,  w_k = e^{-i theta_k} y
    //   adjoint          m_k = e^{i theta_k} lambda_k,  q_k = Im(conj(m_k) y)
    //                    lambda_k <- c m_k + s (i e_k m_other) + 2 A_k s_k,
    //                    A_k = h (g_1k q_1 + g_2k q_2)
    //   gradient         dL/dom0 += -h (s Re(conj(m_k) s_k) + c Re(conj(m_k) z))
    //                    dL/dkw  += sigma_k s X Im(conj(m_k) z),  sigma = +1, -1
    fft_pencil<PL, +1, C2, COLFAST>(v, t, c, ldsk, twl, swz);
    const double2* __restrict__ usk = k ? p.ust2 : p.ust1;
    const unsigned base2 = opaque(base);
    const double mu = p.st->mu;
    const double h = p.di;   // gamma = 0: dt = i h
    double sn, cs;
    sincos_cw(h * p.om0, &sn, &cs);
    const double sig = k ? -1.0 : 1.0;
    const double gown = k ? p.g22 : p.g11, goth = k ? p.g21 : p.g12;            // theta_k / h
    const double hown = h * (k ? p.g22 : p.g11), hoth = h * (k ? p.g12 : p.g21);  // A_k
    double g0 = 0.0, gk = 0.0;
#pragma unroll
    for (int m = 0; m < E; ++m) {
      const unsigned idx = base2 + (unsigned)m * step;
      const double2 sk = usk[idx];
      const double2 so = xchg(sk);
      const double a = ((Vk[idx] - mu) + gown * cabs2(sk)) + goth * cabs2(so);
      const double X = xl[t + T * m];
      const double2 e = exl[t + T * m];
      const double2 ik = make_double2(-sig * e.y, e.x);   // i e_k
      const double2 z = cmul(ik, so);
      const double2 y = make_double2(cs * sk.x - sn * z.x, cs * sk.y - sn * z.y);
      double sth, cth;
      sincos_cw(a * h, &sth, &cth);
      const double2 mk = cmul(v[m], make_double2(cth, sth));
      const double qk = mk.x * y.y - mk.y * y.x;
      const double qo = xchg(qk);
      const double2 mo = xchg(mk);
      const double A2 = 2.0 * (hown * qk + hoth * qo);
      const double2 im = cmul(ik, mo);
      v[m] = make_double2(cs * mk.x + sn * im.x + A2 * sk.x, cs * mk.y + sn * im.y + A2 * sk.y);
      const double re1 = mk.x * sk.x + mk.y * sk.y;
      const double re2 = mk.x * z.x + mk.y * z.y;
      const double im2 = mk.x * z.y - mk.y * z.x;
      g0 -= h * (sn * re1 + cs * re2);
      gk += sig * sn * X * im2;
    }
    fft_pencil<PL, -1, C2, COLFAST>(v, t, c, ldsk, twl, swz);
    const unsigned base3 = opaque(base);
    if (active) {
#pragma unroll
      for (int m = 0; m < E; ++m) psik[base3 + (unsigned)m * step] = v[m];
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
  } else if constexpr (SEQ == SP_ITPA) {
    const double sk = p.st->sc[k];
    const double gk1 = k ? p.g21 : p.g11, gk2 = k ? p.g22 : p.g12;
    double* __restrict__ pk = p.pbuf + (k ? p.npts : 0u);
#pragma unroll
    for (int m = 0; m < E; ++m) {
      const unsigned idx = base + (unsigned)m * step;
      v[m] = cscale(v[m], sk);                                   // phi{j} .* mu (88)
      const double nown = cabs2(v[m]);
      const double noth = xchg(nown);
      const double n1 = k ? noth : nown, n2 = k ? nown : noth;
      const double a = (Vk[idx] + gk1 * n1) + gk2 * n2;         // 49-54
      if (active) pk[idx] = exp_cw(-a * p.dr);                 // exp(-tmp2 dt) (70-71)
    }
    fft_pencil<PL, -1, C2, COLFAST>(v, t, c, ldsk, twl, swz);
    const unsigned bs = opaque(base);   // the store offset is formed here, not kept live
    if (active) {
#pragma unroll
      for (int m = 0; m < E; ++m) psik[bs + (unsigned)m * step] = v[m];
    }
  } else if constexpr (SEQ == SP_ITPB) {
    fft_pencil<PL, +1, C2, COLFAST>(v, t, c, ldsk, twl, swz);
    const double* __restrict__ pk = p.pbuf + (k ? p.npts : 0u);
#pragma unroll
    for (int m = 0; m < E; ++m) {
      const unsigned idx = base + (unsigned)m * step;
      const double f = pk[idx];
      if constexpr (CK == CK_NONE) {
        v[m] = cscale(v[m], f);                                  // GPEtask2comp 70-71
      } else {
        const double2 o = xchg(v[m]);
        double2 cosom, soff;
        if constexpr (CK == CK_ARRAY) {                          // cosh/sinh(dt |Omega|), dt real
          const double2 om = p.cpl[idx];
          const double r = hypot(om.x, om.y);
          double2 sh;
          ccoshsinh(p.dr * r, 0.0, &cosom, &sh);
          const double2 ph = r > 0.0 ? make_double2(om.x / r, om.y / r) : make_double2(1.0, 0.0);
          soff = k ? cmulc(sh, ph) : cmul(sh, ph);
        } else {
          cosom = p.cosom;
          soff = k ? p.sinomm : p.sinomp;
        }
        v[m] = cscale(csub(cmul(cosom, v[m]), cmul(soff, o)), f);   // GPEtaskSpinor 70-71
      }
    }
    fft_pencil<PL, -1, C2, COLFAST>(v, t, c, ldsk, twl, swz);
    const unsigned bs = opaque(base);   // the store offset is formed here, not kept live
    if (active) {
#pragma unroll
      for (int m = 0; m < E; ++m) psik[bs + (unsigned)m * step] = v[m];
    }
  } else if constexpr (SEQ == SP_ITPC) {
    fft_pencil<PL, +1, C2, COLFAST>(v, t, c, ldsk, twl, swz);
    double acc = 0.0;
#pragma unroll
    for (int m = 0; m < E; ++m) acc += cabs2(v[m]);
    const unsigned bs = opaque(base);   // the store offset is formed here, not kept live
    if (active) {
#pragma unroll
      for (int m = 0; m < E; ++m) psik[bs + (unsigned)m * step] = v[m];
    }
    if (!active) acc = 0.0;
    const double r1 = block_sum(k ? 0.0 : acc, red);
    const double r2 = block_sum(k ? acc : 0.0, red);
    if (tid == 0) {
      p.partials[SL_A * p.maxwg + blockIdx.x] = r1;
      p.partials[SL_B * p.maxwg + blockIdx.x] = r2;
    }
  } else {
    double accA = 0.0, accB = 0.0, accC1 = 0.0, accC12 = 0.0, accC2 = 0.0, accX = 0.0, k1 = 0.0, k2 = 0.0;
    const double own1 = k ? 0.0 : 1.0;   // sums that belong to component 1's lane only
    if constexpr (SEQ == SP_KEXS) {
      // k_i^2/2/n |F_x psi|^2 = k_i^2/2/n n^2 |S|^2; S holds index k_i = t + T m
      constexpr double NN = (double)N * (double)N;
#pragma unroll
      for (int m = 0; m < E; ++m) {
        const double e = (p.keax[t + T * m] * NN) * cabs2(v[m]);
        k1 += own1 * e;
        k2 += (1.0 - own1) * e;
      }
      fft_pencil<PL, +1, C2, COLFAST>(v, t, c, ldsk, twl, swz);   // the real-space state
    }
    if (p.apply_sc) {
      const double sk = p.st->sc[k];
#pragma unroll
      for (int m = 0; m < E; ++m) v[m] = cscale(v[m], sk);
    }
#pragma unroll
    for (int m = 0; m < E; ++m) {
      const unsigned idx = base + (unsigned)m * step;
      const double2 o = xchg(v[m]);
      const double nown = cabs2(v[m]), noth = cabs2(o);
      const double2 om = p.cpl ? p.cpl[idx] : p.cpl_s;
      accA += nown;
      accB += Vk[idx] * nown;
      accC1 += own1 * (nown * nown);
      accC12 += own1 * (nown * noth);
      accC2 += (1.0 - own1) * (nown * nown);
      const double2 ow = cmul(om, o);
      accX += own1 * (v[m].x * ow.x + v[m].y * ow.y);   // Re conj(psi1) Omega psi2
    }
    if constexpr (SEQ == SP_KEX) {
      fft_pencil<PL, -1, C2, COLFAST>(v, t, c, ldsk, twl, swz);
#pragma unroll
      for (int m = 0; m < E; ++m) {
        const double e = p.keax[t + T * m] * cabs2(v[m]);
        k1 += own1 * e;
        k2 += (1.0 - own1) * e;
      }
    }
    if (!active) accA = accB = accC1 = accC12 = accC2 = accX = k1 = k2 = 0.0;
    const double vals[8] = {accA, accB, accC1, accC12, accC2, accX, k1, k2};
    const int slots[8] = {SL_A, SL_B, SL_C, SL_C12, SL_C2, SL_X, SL_KX, SL_KX2};
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      const double r = block_sum(vals[i], red);
      if (tid == 0) p.partials[slots[i] * p.maxwg + blockIdx.x] = r;
    }
  }
}

template <int N, int C2, int SEQ, int CK>
__global__ void __launch_bounds__(256) spinor_kernel(const SpinorParams p) {
  spinor_body<N, C2, SEQ, CK>(p);
}

// SP_KEXS unconstrained takes more than 256 registers (one wave per SIMD): its potential
// and coupling loads are hoisted above the inverse transform.  Two waves per SIMD.
template <int N, int C2>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2)))
spinor_kexs_kernel(const SpinorParams p) {
  spinor_body<N, C2, SP_KEXS, CK_SCALAR>(p);
}

// The transposed step has the largest live state (about 276 registers when the
// compiler is unconstrained: one wave per SIMD).  Two waves per SIMD with a few
// spilled registers run it 1.4x faster at 192^3 (255 vs 349 us per sweep);
// GPE_ADJ_WAVES sets the target.
template <int N, int C2>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(GPE_ADJ_WAVES)))
spinor_adj_kernel(const SpinorParams p) {
  spinor_body<N, C2, SP_ADJ, CK_SCALAR>(p);
}

#define GPE_CAT2(a, b) a##b
#define GPE_CAT(a, b) GPE_CAT2(a, b)

hipError_t GPE_CAT(launch_spinor_sweep_, GPE_N)(int seq, const SpinorParams& p, hipStream_t s, int* nblocks) {
  constexpr int N = GPE_N;
  constexpr int C2 = 128 / Plan<N>::T;   // columns per block, two lanes per column position
  const long long nb = ((long long)p.ncols + C2 - 1) / C2;
  if (nblocks) *nblocks = (int)nb;
  if (nb <= 0) return hipSuccess;
  const dim3 g((unsigned)nb), b(256);
  if (seq == SP_PR) {
    if (p.nlincpl != 0.0)
      hipLaunchKernelGGL((spinor_kernel<N, C2, SP_PR, CK_NONLIN>), g, b, 0, s, p);
    else if (p.cpl && p.dr == 0.0 && p.di * p.di * p.cpl_max2 <= 0.6)
      hipLaunchKernelGGL((spinor_kernel<N, C2, SP_PR, CK_ARRAY_TAYLOR>), g, b, 0, s, p);
    else if (p.cpl && p.dr == 0.0)
      hipLaunchKernelGGL((spinor_kernel<N, C2, SP_PR, CK_ARRAY_REAL>), g, b, 0, s, p);
    else if (p.cpl)
      hipLaunchKernelGGL((spinor_kernel<N, C2, SP_PR, CK_ARRAY>), g, b, 0, s, p);
    else if (p.dr == 0.0)
      hipLaunchKernelGGL((spinor_kernel<N, C2, SP_PR, CK_SCALAR_REAL>), g, b, 0, s, p);
    else
      hipLaunchKernelGGL((spinor_kernel<N, C2, SP_PR, CK_SCALAR>), g, b, 0, s, p);
  } else if (seq == SP_KEX) {
    hipLaunchKernelGGL((spinor_kernel<N, C2, SP_KEX, CK_SCALAR>), g, b, 0, s, p);
  } else if (seq == SP_KEXS) {
    hipLaunchKernelGGL((spinor_kexs_kernel<N, C2>), g, b, 0, s, p);
  } else if (seq == SP_ADJ) {
    hipLaunchKernelGGL((spinor_adj_kernel<N, C2>), g, b, 0, s, p);
  } else if (seq == SP_ITPA) {
    hipLaunchKernelGGL((spinor_kernel<N, C2, SP_ITPA, CK_NONE>), g, b, 0, s, p);
  } else if (seq == SP_ITPB) {
    if (p.mixture)
      hipLaunchKernelGGL((spinor_kernel<N, C2, SP_ITPB, CK_NONE>), g, b, 0, s, p);
    else if (p.cpl)
      hipLaunchKernelGGL((spinor_kernel<N, C2, SP_ITPB, CK_ARRAY>), g, b, 0, s, p);
    else
      hipLaunchKernelGGL((spinor_kernel<N, C2, SP_ITPB, CK_SCALAR>), g, b, 0, s, p);
  } else if (seq == SP_ITPC) {
    hipLaunchKernelGGL((spinor_kernel<N, C2, SP_ITPC, CK_NONE>), g, b, 0, s, p);
  } else {
    return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace gpe
