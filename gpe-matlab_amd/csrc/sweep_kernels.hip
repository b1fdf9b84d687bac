// Pencil sweep kernels for one FFT length.  This file is compiled once per
// supported length with -DGPE_N=<n>; each object exports launch_sweep_<n>.
#include "fft_dev.h"
#include "reduce_dev.h"
#include "sweep.h"

#ifndef GPE_N
#error "compile with -DGPE_N=<pencil length>"
#endif

namespace gpe {

// v * exp(E), E = (er, ei), built as MATLAB does: w = exp(E) first, then w .* v.
__device__ __forceinline__ double2 mul_cexp(double2 v, double er, double ei, bool has_re) {
  double s, c;
  sincos(ei, &s, &c);
  const double m = has_re ? exp(er) : 1.0;
  const double2 w = make_double2(m * c, m * s);
  return cmul(w, v);
}

template <int N, int C, bool COLFAST, int SEQ>
__global__ void __launch_bounds__(256) sweep_kernel(const SweepParams p) {
  using PL = Plan<N>;
  constexpr int T = PL::T, E = PL::E;
  static_assert(C * T == 256, "work-group is 256 threads");
  __shared__ double lds[Lds<N, C, COLFAST>::SIZE];
  __shared__ double red[4 * 4];

  const int tid = threadIdx.x;
  const int c = COLFAST ? tid % C : tid / T;
  const int t = COLFAST ? tid / C : tid % T;
  const long long cg = (long long)blockIdx.x * C + c;
  const bool active = cg < p.ncols;
  const long long base = (cg % p.A) * p.sa + (cg / p.A) * p.sb;
  const long long stride = p.stride;

  double2 v[E];
#pragma unroll
  for (int m = 0; m < E; ++m)
    v[m] = active ? p.psi[base + (long long)(t + T * m) * stride] : make_double2(0.0, 0.0);

  const bool has_re = p.dr != 0.0;
  double accA = 0.0, accB = 0.0, accC = 0.0, accK = 0.0;

  if constexpr (SEQ == S_OPEN) {
    // solve_split.m:55-56 with the pending renormalisation (solve_split.m:72-73) folded in
    const double s = p.st->scale, mu = p.st->mu;
#pragma unroll
    for (int m = 0; m < E; ++m) {
      const double vv = active ? p.V[base + (long long)(t + T * m) * stride] - mu : 0.0;
      double2 x = cscale(v[m], s);
      const double n = cabs2(x);
      const double a = -vv - p.g * n;
      v[m] = mul_cexp(x, a * p.dr * 0.5, a * p.di * 0.5, has_re);
    }
    fft_pencil<PL, -1, C, COLFAST>(v, t, c, lds, p.tw);
  } else if constexpr (SEQ == S_MID || SEQ == S_CLOSE) {
    fft_pencil<PL, +1, C, COLFAST>(v, t, c, lds, p.tw);
    const double mu = p.st->mu;
#pragma unroll
    for (int m = 0; m < E; ++m) {
      const double vv = active ? p.V[base + (long long)(t + T * m) * stride] - mu : 0.0;
      const double a = -vv - p.g * cabs2(v[m]);               // solve_split.m:60
      double2 x = mul_cexp(v[m], a * p.dr, a * p.di, has_re);  // solve_split.m:61
      if constexpr (SEQ == S_CLOSE) {
        const double b = vv + p.g * cabs2(x);                   // solve_split.m:63
        x = mul_cexp(x, b * p.dr * 0.5, b * p.di * 0.5, has_re);  // solve_split.m:64
      }
      v[m] = x;
    }
    if constexpr (SEQ == S_MID) fft_pencil<PL, -1, C, COLFAST>(v, t, c, lds, p.tw);
  } else if constexpr (SEQ == S_FWD) {
    fft_pencil<PL, -1, C, COLFAST>(v, t, c, lds, p.tw);
  } else if constexpr (SEQ == S_INV) {
    fft_pencil<PL, +1, C, COLFAST>(v, t, c, lds, p.tw);
  } else if constexpr (SEQ == S_KIN) {
    fft_pencil<PL, -1, C, COLFAST>(v, t, c, lds, p.tw);
    const double2 kc = active ? p.kcol[cg] : make_double2(0.0, 0.0);
#pragma unroll
    for (int m = 0; m < E; ++m) v[m] = cmul(cmul(kc, p.kax[t + T * m]), v[m]);
    fft_pencil<PL, +1, C, COLFAST>(v, t, c, lds, p.tw);
  } else if constexpr (SEQ == S_KEX || SEQ == S_KE) {
    if constexpr (SEQ == S_KEX) {
#pragma unroll
      for (int m = 0; m < E; ++m) {
        const double n = cabs2(v[m]);
        const double vv = active ? p.V[base + (long long)(t + T * m) * stride] : 0.0;
        accA += n;
        accB += vv * n;
        accC += n * n;
      }
    }
    fft_pencil<PL, -1, C, COLFAST>(v, t, c, lds, p.tw);
#pragma unroll
    for (int m = 0; m < E; ++m) accK += p.keax[t + T * m] * cabs2(v[m]);
  } else if constexpr (SEQ == S_ITPA) {
    fft_pencil<PL, +1, C, COLFAST>(v, t, c, lds, p.tw);
#pragma unroll
    for (int m = 0; m < E; ++m) {
      const double nl = active ? p.nop[base + (long long)(t + T * m) * stride] : 0.0;
      v[m] = cscale(v[m], nl);                                 // groundstate_itp.m:42
      accA += cabs2(v[m]);                                     // groundstate_itp.m:44
    }
  } else if constexpr (SEQ == S_ITPB) {
    const double mu = p.st->mu_itp;
    const double mu2 = mu * mu;
    const double hdt = -p.dr * 0.5;
#pragma unroll
    for (int m = 0; m < E; ++m) {
      const long long idx = base + (long long)(t + T * m) * stride;
      const double vv = active ? p.V[idx] : 0.0;
      const double n = cabs2(v[m]);                            // nlinop before renormalisation
      const double nl = exp(hdt * (n * p.g * mu2 + vv));       // groundstate_itp.m:76 (36 with mu = 1)
      if (active) p.nop[idx] = nl;
      v[m] = cscale(cscale(v[m], mu), nl);                     // groundstate_itp.m:52 then 40
    }
    fft_pencil<PL, -1, C, COLFAST>(v, t, c, lds, p.tw);
  }

  // ---- write back
  constexpr bool STORE = !(SEQ == S_KEX || SEQ == S_KE);
  if constexpr (STORE) {
    if (active) {
#pragma unroll
      for (int m = 0; m < E; ++m) p.psi[base + (long long)(t + T * m) * stride] = v[m];
    }
  }

  // ---- deterministic block reductions
  if constexpr (SEQ == S_KEX) {
    double* out = p.partials;
    const double a = block_sum(accA, red);
    const double b = block_sum(accB, red);
    const double cc = block_sum(accC, red);
    const double k = block_sum(accK, red);
    if (tid == 0) {
      out[SL_A * p.maxwg + blockIdx.x] = a;
      out[SL_B * p.maxwg + blockIdx.x] = b;
      out[SL_C * p.maxwg + blockIdx.x] = cc;
      out[p.slot_ke * p.maxwg + blockIdx.x] = k;
    }
  } else if constexpr (SEQ == S_KE) {
    const double k = block_sum(accK, red);
    if (tid == 0) p.partials[p.slot_ke * p.maxwg + blockIdx.x] = k;
  } else if constexpr (SEQ == S_ITPA) {
    const double a = block_sum(accA, red);
    if (tid == 0) p.partials[SL_A * p.maxwg + blockIdx.x] = a;
  }
}

#define GPE_CAT2(a, b) a##b
#define GPE_CAT(a, b) GPE_CAT2(a, b)

hipError_t GPE_CAT(launch_sweep_, GPE_N)(int seq, bool colfast, const SweepParams& p, hipStream_t s,
                                          int* nblocks) {
  constexpr int N = GPE_N;
  constexpr int C = 256 / Plan<N>::T;
  const long long nb = (p.ncols + C - 1) / C;
  if (nblocks) *nblocks = (int)nb;
  if (nb <= 0) return hipSuccess;
  dim3 grid((unsigned)nb), block(256);
#define GPE_L(CF, S)                                                     \
  case S:                                                                \
    hipLaunchKernelGGL((sweep_kernel<N, C, CF, S>), grid, block, 0, s, p); \
    break;
  if (colfast) {
    switch (seq) {
      GPE_L(true, S_FWD) GPE_L(true, S_INV) GPE_L(true, S_KIN) GPE_L(true, S_OPEN) GPE_L(true, S_MID)
      GPE_L(true, S_CLOSE) GPE_L(true, S_KEX) GPE_L(true, S_KE) GPE_L(true, S_ITPA) GPE_L(true, S_ITPB)
      default: return hipErrorInvalidValue;
    }
  } else {
    switch (seq) {
      GPE_L(false, S_FWD) GPE_L(false, S_INV) GPE_L(false, S_KIN) GPE_L(false, S_KE)
      GPE_L(false, S_KEX)
      default: return hipErrorInvalidValue;
    }
  }
#undef GPE_L
  return hipGetLastError();
}

}  // namespace gpe
