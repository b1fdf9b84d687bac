// Stochastic GPE: SGPEtask.solve_split (lib/@SGPEtask/solve_split.m:1-81), omega = 0.
//
// Per block (52-76): block-open half phase with the stored density, n_rec inner
// steps (noise of step i added in real space -- 59 --, kinetic factor exp(-kk dt),
// masked to kk < ecut when ecut > 0 and T > 0 -- 39-44, 60 --, full phase -- 69 --),
// closing half phase (72) and ncur = integrate(|psi|^2) (73-75).  mu = mu_init is
// fixed and the state is never renormalised.  The noise of global step s at grid
// point p is the complex normal variate of Philox counter (p, s) under the task's
// noise seed (philox_dev.h), so a run is reproducible and resumable.  On a 3D grid with
// a 256-point y axis a step is the two sweeps of the y-split step (S_YSOPEN / S_YSMID
// with the noise, S_YKIN / S_YKINM), as for solve_split; elsewhere the four sweeps.
#include <cmath>
#include <vector>

#include "handle.h"

using namespace gpe_detail;

namespace {

SweepParams sg_params(gpe_handle* h, int axis) {
  SweepParams p = base_params(h, axis, h->psi[0], h->V[0]);
  p.noise_var = h->sg_var;
  p.noise_seed = h->sg_seed;
  p.kmax = h->sg_kmax;
  return p;
}

// The block of sg_block on the y-split step (blk_in / blk_out as for run_block).
int sg_block_ysplit(gpe_handle* h, long long step0, int* nb, bool blk_in, bool blk_out) {
  int rc;
  if ((rc = ysplit_refresh_v(h))) return rc;
  const bool noisy = h->sg_var != 0.0;
  SweepParams px = ysplit_params(h);
  px.yblk_in = blk_in ? 1 : 0;
  px.yblk_out = blk_out ? 1 : 0;
  px.noise_var = h->sg_var;
  px.noise_seed = h->sg_seed;
  SweepParams pz = ysplit_zparams(h);
  pz.kmax = h->sg_kmaxB;
  const int kin = h->sg_mask ? gpe::S_YKINM : gpe::S_YKIN;
  for (int i = 0; i < h->n_rec; ++i) {
    px.noise_step = (unsigned long long)(step0 + i);
    const int xs = i == 0 ? (noisy ? gpe::S_YSOPEN : gpe::S_YOPEN) : (noisy ? gpe::S_YSMID : gpe::S_YMID);
    if ((rc = ysplit_xsweep(h, xs, px, nullptr))) return rc;
    if ((rc = sweep(h, 2, kin, pz))) return rc;
  }
  return ysplit_xsweep(h, gpe::S_YCLOSEM, px, nb);   // 72-75 (only the norm is used)
}

int sg_block(gpe_handle* h, long long step0, double* ncur, bool blk_in = false, bool blk_out = false) {
  int rc;
  if (ysplit_active(h)) {
    int nb = 0;
    if ((rc = sg_block_ysplit(h, step0, &nb, blk_in, blk_out))) return rc;
    FinParams f = fin_params(h, gpe::FIN_SUMS);
    f.cnt[gpe::SL_A] = nb;
    HIPCHK(h, gpe::launch_finalize(f, h->stream));
    if (ncur) {
      DevState s;
      if ((rc = read_state(h, &s))) return rc;
      if (!std::isfinite(s.sA)) return fail(h, GPE_ENONFINITE, "SGPE block norm is not finite (%g)", s.sA);
      *ncur = s.sA * h->w;
    }
    return GPE_OK;
  }
  const bool noisy = h->sg_var != 0.0;
  const int kin = h->sg_mask ? gpe::S_KINM : gpe::S_KIN;
  for (int i = 0; i < h->n_rec; ++i) {
    SweepParams px = sg_params(h, 0);
    px.noise_step = (unsigned long long)(step0 + i);
    const int xs = i == 0 ? (noisy ? gpe::S_SOPEN : gpe::S_OPEN) : (noisy ? gpe::S_SMID : gpe::S_MID);
    if ((rc = sweep(h, 0, xs, px))) return rc;
    if (h->ndims == 3) {
      if ((rc = sweep(h, 1, gpe::S_FWD, sg_params(h, 1)))) return rc;
      if ((rc = sweep(h, 2, kin, sg_params(h, 2)))) return rc;
      if ((rc = sweep(h, 1, gpe::S_INV, sg_params(h, 1)))) return rc;
    } else {
      if ((rc = sweep(h, 1, kin, sg_params(h, 1)))) return rc;
    }
  }
  int nb = 0;
  FinParams f = fin_params(h, gpe::FIN_SUMS);
  if ((rc = sweep(h, 0, gpe::S_CLOSEM, sg_params(h, 0), &nb))) return rc;   // 72-75
  f.cnt[gpe::SL_A] = nb;
  HIPCHK(h, gpe::launch_finalize(f, h->stream));
  if (ncur) {
    DevState s;
    if ((rc = read_state(h, &s))) return rc;
    if (!std::isfinite(s.sA)) return fail(h, GPE_ENONFINITE, "SGPE block norm is not finite (%g)", s.sA);
    *ncur = s.sA * h->w;
  }
  return GPE_OK;
}

}  // namespace

extern "C" {

int gpe_sgpe_begin(gpe_handle* h, const gpe_sgpe_args* a) {
  int rc;
  if ((rc = ensure_ready(h))) return rc;
  if (!a) return fail(h, GPE_EINVAL, "null SGPE arguments");
  if (h->prm.ncomp != 1) return fail(h, GPE_EINVAL, "the stochastic GPE has one component");
  if (h->omega != 0.0) return fail(h, GPE_EINVAL, "the stochastic GPE runs without rotation (omega = 0)");
  if (a->niter_inner < 1 || h->prm.n_recalc < 1) return fail(h, GPE_EINVAL, "niter_inner and n_recalc must be >= 1");
  if (!(a->T >= 0) || !std::isfinite(a->T) || !std::isfinite(a->ecut)) return fail(h, GPE_EINVAL, "bad T or ecut");
  if ((rc = materialize(h))) return rc;
  h->n_rec = h->prm.n_recalc;                          // no n_rec fallback (52 truncates)
  h->niter_inner = a->niter_inner;
  h->ddt = a->ddt;
  const std::complex<double> dt = cdiv({0.0, a->ddt}, {1.0, h->prm.gamma});   // 29
  h->dr = dt.real();
  h->di = dt.imag();
  if ((rc = ysplit_prepare(h))) return rc;
  if ((rc = build_kinetic(h, dt, h->kcol, h->kax, 1.0, 0.0, ysplit_eligible(h) ? h->kcolB : nullptr)))
    return rc;                                                                   // 41 / 43
  h->sg_mask = a->ecut > 0 && a->T > 0;                                         // 39
  h->sg_var = std::sqrt(a->T * h->prm.gamma * a->ddt / h->w);                 // 47
  h->sg_seed = a->seed;
  if (h->sg_mask) {
    const int ma = mid_axis(h);
    const Axis& m = h->ax[ma];
    std::vector<double> col((size_t)m.ncols), ax(m.n);
    if (h->ndims == 3) {
      for (int ix = 0; ix < h->nx; ++ix)
        for (int iy = 0; iy < h->ny; ++iy)
          col[(size_t)iy + (size_t)h->ny * ix] = h->ax[0].k[ix] * h->ax[0].k[ix] + h->ax[1].k[iy] * h->ax[1].k[iy];
    } else {
      for (int ix = 0; ix < h->nx; ++ix) col[ix] = h->ax[0].k[ix] * h->ax[0].k[ix];
    }
    for (int i = 0; i < m.n; ++i) ax[i] = m.k[i] * m.k[i];
    // k_{n-a} = -k_a, and k_a^2 grows with a up to the Nyquist index n/2, so the mask of a
    // column is |k| <= kmax: the longest prefix a = 0, 1, ... with kk < ecut
    for (int q = 1; q < m.n - q; ++q)
      if (ax[q] != ax[m.n - q]) return fail(h, GPE_EINVAL, "asymmetric k table on the mid axis");
    std::vector<int> kmax((size_t)m.ncols);
    for (size_t c = 0; c < col.size(); ++c) {
      int q = -1;
      while (q < m.n / 2 && (col[c] + ax[q + 1]) / 2 < a->ecut) ++q;
      kmax[c] = q;
    }
    if (!h->sg_kmax) HIPCHK(h, hipMalloc(&h->sg_kmax, kmax.size() * sizeof(int)));
    HIPCHK(h, hipMemcpy(h->sg_kmax, kmax.data(), kmax.size() * sizeof(int), hipMemcpyHostToDevice));
    if (ysplit_eligible(h)) {
      // the y-split z sweep's column order (the kcolB order of build_kinetic): entry
      // pos + 256 ix belongs to ky = bitrev4(pos / 16) + 16 (pos mod 16)
      std::vector<int> kb(kmax.size());
      for (int ix = 0; ix < h->nx; ++ix)
        for (int pos = 0; pos < 256; ++pos) {
          const int r = pos >> 4;
          const int ky = (((r & 1) << 3) | ((r & 2) << 1) | ((r & 4) >> 1) | ((r & 8) >> 3)) + 16 * (pos & 15);
          kb[(size_t)pos + 256 * (size_t)ix] = kmax[(size_t)ky + 256 * (size_t)ix];
        }
      if (!h->sg_kmaxB) HIPCHK(h, hipMalloc(&h->sg_kmaxB, kb.size() * sizeof(int)));
      HIPCHK(h, hipMemcpy(h->sg_kmaxB, kb.data(), kb.size() * sizeof(int), hipMemcpyHostToDevice));
    }
  }
  DevState s;
  if ((rc = read_state(h, &s))) return rc;
  s.mu = h->prm.mu_init;                                // 50
  s.scale = 1.0;
  if ((rc = write_state(h, s))) return rc;
  h->sg_begun = true;
  return GPE_OK;
}

int gpe_sgpe_block(gpe_handle* h, long long step0, double* ncur) {
  int rc;
  if ((rc = ensure_ready(h))) return rc;
  if (!h->sg_begun) return fail(h, GPE_ESTATE, "gpe_sgpe_begin has not been called");
  if (step0 < 0) return fail(h, GPE_EINVAL, "negative step");
  return sg_block(h, step0, ncur);
}

int gpe_sgpe_outer(gpe_handle* h, int j, double* time2_out, double* ncur) {
  int rc;
  if ((rc = ensure_ready(h))) return rc;
  if (!h->sg_begun) return fail(h, GPE_ESTATE, "gpe_sgpe_begin has not been called");
  if (j < 1) return fail(h, GPE_EINVAL, "j must be >= 1");
  const double dt_outer = h->ddt * h->niter_inner;
  const double time = (double)(j - 1) * dt_outer;       // 51
  double time2 = time;
  const int nblocks = h->niter_inner / h->n_rec;
  for (int jj = 1; jj <= nblocks; ++jj) {
    time2 = time + (double)(jj - 1) * h->ddt * (double)h->n_rec;   // 53
    const long long step0 = (long long)(j - 1) * h->niter_inner + (long long)(jj - 1) * h->n_rec;
    if ((rc = sg_block(h, step0, jj == nblocks ? ncur : nullptr, jj > 1, jj < nblocks))) return rc;
  }
  if (time2_out) *time2_out = time2;
  return GPE_OK;
}

}  // extern "C"
