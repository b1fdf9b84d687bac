// Second-stage reductions and block-boundary scalar updates, shared by the
// single-grid and the batched ensemble drivers.
#pragma once
#include "reduce_dev.h"
#include "state.h"

namespace gpe {

// Sum `count` partials of one slot in a fixed order (valid on thread 0).
__device__ __forceinline__ double slot_sum(const double* part, int maxwg, int slot, int count, double* red) {
  double acc = 0.0;
  for (int i = threadIdx.x; i < count; i += 256) acc += part[slot * maxwg + i];
  return block_sum(acc, red);
}

__device__ __forceinline__ double slot_max(const double* part, int maxwg, int slot, int count, double* red) {
  double acc = 0.0;
  for (int i = threadIdx.x; i < count; i += 256) acc = fmax(acc, part[slot * maxwg + i]);
  return block_max(acc, red);
}

// Second-stage reductions and the block-boundary scalar update of one grid
// (one 256-thread work-group; `red` is a 4-double __shared__ scratch).
__device__ __forceinline__ void finalize_core(const FinParams& f, double* red) {
  DevState* st = f.st;
  if (f.mode == FIN_MAX) {
    const double m = slot_max(f.partials, f.maxwg, SL_MAX, f.cnt[SL_MAX], red);
    if (threadIdx.x == 0) st->delta = m / (f.dt * 10.0);   // groundstate_itp.m:61
    return;
  }
  // one pass over the partials with every used slot's loads in flight together
  constexpr int NS = 12;
  const int slots[NS] = {SL_A, SL_B, SL_C, SL_KX, SL_KY, SL_KZ, SL_C12, SL_C2, SL_X, SL_KX2, SL_KY2, SL_KZ2};
  double acc[NS];
  int maxc = 0;
#pragma unroll
  for (int q = 0; q < NS; ++q) {
    acc[q] = 0.0;
    maxc = f.cnt[slots[q]] > maxc ? f.cnt[slots[q]] : maxc;
  }
  // four strides of loads in flight per pass; each thread still adds its partials in
  // index order (t, t + 256, t + 512, ...)
  for (int i0 = threadIdx.x; i0 < maxc; i0 += 4 * 256) {
    double v[4][NS];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int i = i0 + u * 256;
#pragma unroll
      for (int q = 0; q < NS; ++q) v[u][q] = i < f.cnt[slots[q]] ? f.partials[slots[q] * f.maxwg + i] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int i = i0 + u * 256;
#pragma unroll
      for (int q = 0; q < NS; ++q)
        if (i < f.cnt[slots[q]]) acc[q] += v[u][q];
    }
  }
  double tot[NS];
#pragma unroll
  for (int q = 0; q < NS; ++q) tot[q] = f.cnt[slots[q]] ? block_sum(acc[q], red) : 0.0;
  const double sA = tot[0], sB = tot[1], sC = tot[2], kx = tot[3], ky = tot[4], kz = tot[5];
  const double c12 = tot[6], c2 = tot[7], sx = tot[8], kx2 = tot[9], ky2 = tot[10], kz2 = tot[11];
  if (threadIdx.x != 0) return;
  if (f.mode == FIN_ITP2) {
    // GPEtaskSpinor/groundstate_itp.m:82-88, GPEtask2comp/groundstate_itp.m:82-90
    if (f.mixture) {
      const double mu1 = sqrt(f.n1t / (sA * f.w)), mu2 = sqrt(f.n2t / (sB * f.w));
      st->sc[0] = mu1;
      st->sc[1] = mu2;
      f.MU[f.iter] = log(mu1) / f.dt;
      f.MU2[f.iter] = log(mu2) / f.dt;
    } else {
      const double mu = sqrt(f.n1t / ((sA + sB) * f.w));
      st->sc[0] = mu;
      st->sc[1] = mu;
      f.MU[f.iter] = log(mu) / f.dt;
    }
    return;
  }
  if (f.mode == FIN_ITP2M) {
    // real(inner(phi, applyham(phi))) and get_energy (g/2) over Ntotal
    // (GPEtaskSpinor/groundstate_itp.m:91-100); x sums come from the scaled state,
    // the y/z kinetic sums from the unscaled one
    const double q1 = st->sc[0] * st->sc[0], q2 = st->sc[1] * st->sc[1];
    const double sKs = (kx + q1 * (ky + kz)) + (kx2 + q2 * (ky2 + kz2));
    const double sCf = (f.g11 * sC + (f.g12 + f.g21) * c12) + f.g22 * c2;
    const double base = (sB + sKs) + 2.0 * sx;
    if (f.which & 1) f.MU2[f.iter] = (base + sCf) * f.w / f.Ntotal;
    if (f.which & 2) f.EE[f.iter] = (base + 0.5 * sCf) * f.w / f.Ntotal;
    return;
  }
  const double sK = ((kx + ky) + kz) + ((kx2 + ky2) + kz2);
  st->sA = sA; st->sB = sB; st->sK = sK;
  // scalar runs: sC = sum |psi|^4; spinor runs: sC = sum_jk g_jk n_j n_k, sX = sum Re conj(psi1) Omega psi2
  st->sC = f.mode == FIN_SPINOR || f.cnt[SL_C12] ? (f.g11 * sC + (f.g12 + f.g21) * c12) + f.g22 * c2 : sC;
  st->pad[0] = sx;
  const double w = f.w;
  if (f.mode == FIN_BLOCK) {
    // solve_split.m:65-78
    const double ncur = sA * w;
    double muc;
    if (f.renorm) {
      double NNN = st->NNN;
      if (f.tau > 0) NNN = f.NN0 * exp(-f.time2 / f.tau);
      const double q = NNN / ncur;
      st->scale = sqrt(q);
      // real(inner(phi_s, applyham(phi_s)))/NNN with phi_s = sqrt(q) phi
      muc = ((q * sB + q * sK) + f.g * (q * q) * sC) * w / NNN;
      st->mu = muc;
      st->NNN = NNN;
    } else {
      st->scale = 1.0;
      muc = ((sB + sK) + f.g * sC) * w / ncur;
    }
    st->muc = muc;
    st->ncur = ncur;
  } else if (f.mode == FIN_SPINOR) {
    // lib/@GPEtaskSpinor/solve_split.m:100-115: ncur, optional renormalisation, mu = <phi|H phi>/NNN
    const double ncur = sA * w;
    double NNN = st->NNN;
    double q = 1.0;
    if (f.renorm) {
      if (f.tau > 0) NNN = f.NN0 * exp(-f.time2 / f.tau);
      q = NNN / ncur;
    }
    st->scale = f.renorm ? sqrt(q) : 1.0;
    const double mu = (q * ((sB + sK) + 2.0 * sx) + (q * q) * st->sC) * w / NNN;
    st->mu = mu;
    st->muc = mu;
    st->ncur = ncur;
    st->NNN = NNN;
  } else if (f.mode == FIN_ITP) {
    // groundstate_itp.m:44-52
    double mu;
    if (f.Ntotal > 0) {
      mu = sqrt(f.Ntotal / (sA * w));
      if (f.MU) f.MU[f.iter] = log(mu) / f.dt;
    } else {
      mu = exp(f.mu_init * f.dt);
      if (f.MU) f.MU[f.iter] = (sA * w) * (mu * mu);
    }
    st->mu_itp = mu;
  }
}

}  // namespace gpe
