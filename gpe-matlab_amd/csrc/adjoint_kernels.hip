// Adjoint x-sweeps for one FFT length (-DGPE_N=<n>).
//
// The backward pass carries two arrays through the same conjugated steps: the
// state psi, reconstructed backwards (every forward map is exactly invertible
// for gamma = 0), and the adjoint lambda = 2 dL/d conj(psi).  For a phase map
// w = e^{-i theta} u, theta = tau (V - mu + g |u|^2) (|u| = |w|):
//     u        = e^{i theta} w
//     lambda_u = e^{i theta} (lambda_w - 2 tau g Im(conj(w) lambda_w) w)
//     dL/dV   += tau Im(conj(lambda_w) w)
// A_CLOSE undoes the block-close and last full phase (solve_split.m:60-64) and
// applies the forward x-FFT; A_MID does inverse x-FFT, undoes one full phase,
// forward x-FFT; A_OPEN does inverse x-FFT and undoes the block-open phase
// (solve_split.m:55-56).  dL/dV is accumulated in gV (read-modify-write).
#include <type_traits>

#include "adjoint.h"
#include "fft_dev.h"
#include "phase_dev.h"
#include "reduce_dev.h"

#ifndef GPE_N
#error "compile with -DGPE_N=<pencil length>"
#endif

namespace gpe {

__device__ __forceinline__ void undo_phase(double2& w, double2& lam, double& gv, double V, double mu, double g,
                                           double tau) {
  const double th = tau * ((V - mu) + g * cabs2(w));
  double s, c;
  sincos_cw(th, &s, &c);
  gv += tau * (lam.x * w.y - lam.y * w.x);            // tau Im(conj(lambda) w)
  const double xi = w.x * lam.y - w.y * lam.x;         // Im(conj(w) lambda)
  const double f = 2.0 * tau * g * xi;
  const double2 l = make_double2(lam.x - f * w.x, lam.y - f * w.y);
  lam = make_double2(c * l.x - s * l.y, c * l.y + s * l.x);
  w = make_double2(c * w.x - s * w.y, c * w.y + s * w.x);
}

// Plan of the adjoint x sweep: Plan<N>, or with var = 1 the length's second plan
// (AltPlan, fewer elements per thread where it has one), chosen by the handle's x-axis
// plan variant.
template <int N, int VAR>
using AdjPlan = typename std::conditional<VAR == 1 && AltPlan<N>::has, typename AltPlan<N>::type, Plan<N>>::type;

template <int N, int C, int SEQ, int VAR>
__global__ void __launch_bounds__(256) adj_x_kernel(const AdjParams p) {
  using PL = AdjPlan<N, VAR>;
  constexpr int T = PL::T, E = PL::E;
  constexpr bool COLFAST = true;
  __shared__ double lds[Lds<N, C, COLFAST>::SIZE];
  __shared__ double2 twl[N];

  const int tid = threadIdx.x;
  const int c = tid % C;
  const int t = tid / C;
  for (int i = tid; i < N; i += 256) twl[i] = p.tw[i];

  const unsigned cgu = (C < 16 ? xcd_pair_tile(blockIdx.x, gridDim.x) : blockIdx.x) * (unsigned)C + (unsigned)c;
  const bool active = cgu < p.ncols;
  const unsigned cg = active ? cgu : p.ncols - 1u;
  const unsigned base = (cg % p.A) * p.sa + (cg / p.A) * p.sb + (unsigned)t * p.stride;
  const unsigned step = (unsigned)T * p.stride;

  double2 u[E], l[E];
#pragma unroll
  for (int m = 0; m < E; ++m) {
    u[m] = p.psi[base + (unsigned)m * step];
    l[m] = p.lam[base + (unsigned)m * step];
  }
  if constexpr (SEQ != A_CLOSE) {
    fft_pencil<PL, +1, C, COLFAST>(u, t, c, lds, twl);
    fft_pencil<PL, +1, C, COLFAST>(l, t, c, lds, twl);
  }
  const double mu = p.st->mu;
  const double h = p.h;
#pragma unroll
  for (int m = 0; m < E; ++m) {
    const unsigned idx = base + (unsigned)m * step;
    const double V = p.V[idx];
    double gv = 0.0;
    if constexpr (SEQ == A_CLOSE) {
      undo_phase(u[m], l[m], gv, V, mu, p.g, -0.5 * h);   // block-close phase, tau = -h/2
      undo_phase(u[m], l[m], gv, V, mu, p.g, h);          // last full phase
    } else if constexpr (SEQ == A_MID) {
      undo_phase(u[m], l[m], gv, V, mu, p.g, h);
    } else {
      undo_phase(u[m], l[m], gv, V, mu, p.g, 0.5 * h);    // block-open phase
    }
    if (active) p.gV[idx] += gv;
  }
  if constexpr (SEQ != A_OPEN) {
    fft_pencil<PL, -1, C, COLFAST>(u, t, c, lds, twl);
    fft_pencil<PL, -1, C, COLFAST>(l, t, c, lds, twl);
  }
  const unsigned bs = opaque(base);   // the store offset is formed here, not kept live
  if (active) {
#pragma unroll
    for (int m = 0; m < E; ++m) {
      p.psi[bs + (unsigned)m * step] = u[m];
      p.lam[bs + (unsigned)m * step] = l[m];
    }
  }
}

#define GPE_CAT2(a, b) a##b
#define GPE_CAT(a, b) GPE_CAT2(a, b)

template <int N, int VAR>
static hipError_t launch_adj_var(int seq, const AdjParams& p, hipStream_t s, int* nblocks) {
  constexpr int C = 256 / AdjPlan<N, VAR>::T;
  const long long nb = ((long long)p.ncols + C - 1) / C;
  if (nblocks) *nblocks = (int)nb;
  if (nb <= 0) return hipSuccess;
  switch (seq) {
    case A_OPEN: hipLaunchKernelGGL((adj_x_kernel<N, C, A_OPEN, VAR>), dim3((unsigned)nb), dim3(256), 0, s, p); break;
    case A_MID: hipLaunchKernelGGL((adj_x_kernel<N, C, A_MID, VAR>), dim3((unsigned)nb), dim3(256), 0, s, p); break;
    case A_CLOSE: hipLaunchKernelGGL((adj_x_kernel<N, C, A_CLOSE, VAR>), dim3((unsigned)nb), dim3(256), 0, s, p); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t GPE_CAT(launch_adj_x_, GPE_N)(int seq, int var, const AdjParams& p, hipStream_t s, int* nblocks) {
  constexpr int N = GPE_N;
  if constexpr (AltPlan<N>::has) {
    if (var == 1) return launch_adj_var<N, 1>(seq, p, s, nblocks);
  }
  return launch_adj_var<N, 0>(seq, p, s, nblocks);
}

}  // namespace gpe
