// Two-component (GPEtaskSpinor) split-step driver.
//
// Per block of n_rec steps (lib/@GPEtaskSpinor/solve_split.m:67-115):
//   open  : each component -> exp(-kk dt/2) in k space          (73-75)
//   steps : both components -> potential/interaction phase and  (76-96)
//           2x2 Rabi rotation in real space, then exp(-kk dt)
//   close : exp(+kk dt/2) folded into the last step's kinetic    (97-99)
//           multiply (table exp(-kk dt) exp(kk dt/2)), inverse x
//   reduce: ncur, optional renormalisation, mu = <phi|H phi>/NNN (100-115)
// In 3D the last kinetic sweep of a block (S_KINE) also sums kk |.|^2 over the spectrum
// it has just multiplied, which is the block-end state's: the kinetic energy needs no
// sweeps of its own.
// Between the sweeps of a step the state is the x spectrum S = F_x psi / nx.  In 3D,
// block ends inside an outer iteration keep it: the sums are taken from S (SP_KEXS and
// kinetic sums with nx-scaled weights), and the next block opens from S with its
// kinetic half step (the pending renormalisation in the y transform, nx in the kz
// table) instead of an inverse and a forward x sweep.
// 3D sweep order per step: x (both components, SP_PR), then for each component
// y forward, z forward.kinetic.inverse, y inverse.  2D: x (SP_PR), y kinetic.
#include "handle.h"
#include "spinor.h"

using namespace gpe_detail;

namespace gpe_spinor {

gpe::SpinorParams spinor_params(gpe_handle* h) {
  gpe::SpinorParams p{};
  const Axis& a = h->ax[0];
  p.psi1 = h->psi[0];
  p.psi2 = h->psi[1];
  p.V1 = h->V[0];
  p.V2 = h->V[1];
  p.cpl = h->cpl_array ? h->cpl : nullptr;
  p.cpl_max2 = h->cpl_max2;
  p.tw = a.tw;
  p.partials = h->partials;
  p.st = h->st;
  p.ncols = (unsigned)a.ncols;
  p.A = (unsigned)a.A;
  p.sa = (unsigned)a.sa;
  p.sb = (unsigned)a.sb;
  p.stride = (unsigned)a.stride;
  p.maxwg = h->maxwg;
  // g is column-major {g11, g21, g12, g22}
  p.g11 = h->prm.g[0];
  p.g21 = h->prm.g[1];
  p.g12 = h->prm.g[2];
  p.g22 = h->prm.g[3];
  p.dr = h->dr;
  p.di = h->di;
  p.cpl_s = make_double2(h->cpl_re, h->cpl_im);
  p.cosom = make_double2(h->cosom[0], h->cosom[1]);
  p.sinomm = make_double2(h->sinomm[0], h->sinomm[1]);
  p.sinomp = make_double2(h->sinomp[0], h->sinomp[1]);
  p.nlincpl = h->prm.nlincpl;
  p.keax = a.keax;
  p.npts = (unsigned)h->npts;
  return p;
}

int spinor_sweep(gpe_handle* h, int seq, const gpe::SpinorParams& p, int* nb) {
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (h->prof) {
    e0 = prof_event(h);
    e1 = prof_event(h);
    if (e0) hipEventRecord(e0, h->stream);
  }
  hipError_t e = gpe::launch_spinor_sweep(h->ax[0].n, seq, p, h->stream, nb);
  if (h->prof && e0 && e1) {
    hipEventRecord(e1, h->stream);
    // spinor sweeps are reported in the (seq, axis) slots of S_MID / S_KEX / S_CLOSE on x
    const int slot = seq == gpe::SP_PR ? gpe::S_MID : seq == gpe::SP_KEX || seq == gpe::SP_KEXS ? gpe::S_KEX : gpe::S_CLOSE;
    h->prof_recs.push_back({slot * 3 + 0, {e0, e1}});
  }
  if (e != hipSuccess) return fail(h, GPE_EHIP, "spinor sweep launch: %s", hipGetErrorString(e));
  return GPE_OK;
}

// Scalar Rabi coefficients for the current coupling (solve_split.m:61-65).
void rabi_scalars(gpe_handle* h) {
  const std::complex<double> dt(h->dr, h->di);
  const std::complex<double> om(h->cpl_re, h->cpl_im);
  const double r = std::abs(om);
  const double ang = std::arg(om);
  const std::complex<double> ch = std::cosh(dt * r);
  const std::complex<double> sh = std::sinh(dt * r);
  const std::complex<double> m = sh * std::exp(std::complex<double>(0.0, -ang));
  const std::complex<double> pl = sh * std::exp(std::complex<double>(0.0, ang));
  h->cosom[0] = ch.real(); h->cosom[1] = ch.imag();
  h->sinomm[0] = m.real(); h->sinomm[1] = m.imag();
  h->sinomp[0] = pl.real(); h->sinomp[1] = pl.imag();
}

// Reductions for ncur and mu on the real-space state of both components
// (scaled: each component multiplied by st->sc[k] in the x-sweep), or, when
// `spectral`, on the stored x spectrum S = F_x psi / nx of a block end inside an outer
// iteration (3D; y and z kinetic sums of S weighted by nx k^2/2/n, Parseval along x).
int spinor_reduce(gpe_handle* h, FinParams* f, bool scaled, bool spectral, int ke_groups) {
  int rc, nb = 0;
  gpe::SpinorParams p = spinor_params(h);
  p.apply_sc = scaled ? 1 : 0;
  if (h->use_vmu && h->Vmu[0] && h->Vmu[1]) {   // applyham at task.current_time (see gpe_set_potential_mu)
    p.V1 = h->Vmu[0];
    p.V2 = h->Vmu[1];
  }
  if ((rc = spinor_sweep(h, spectral ? gpe::SP_KEXS : gpe::SP_KEX, p, &nb))) return rc;
  for (int s : {gpe::SL_A, gpe::SL_B, gpe::SL_C, gpe::SL_C12, gpe::SL_C2, gpe::SL_X, gpe::SL_KX, gpe::SL_KX2})
    f->cnt[s] = nb;
  if (ke_groups > 0) {
    // the whole kinetic sum of each component is in SL_KZ / SL_KZ2 (S_KINE); the x sums
    // of the sweep above are left out
    f->cnt[gpe::SL_KX] = f->cnt[gpe::SL_KX2] = 0;
    f->cnt[gpe::SL_KZ] = f->cnt[gpe::SL_KZ2] = ke_groups;
  }
  for (int comp = 0; comp < 2 && ke_groups <= 0; ++comp)
    for (int a = 1; a < h->ndims; ++a) {
      SweepParams q = base_params(h, a, h->psi[comp], h->V[comp]);
      if (spectral) q.keax = h->keax_nx[a];
      q.slot_ke = (comp == 0 ? gpe::SL_KX : gpe::SL_KX2) + a;
      if ((rc = sweep(h, a, gpe::S_KE, q, &nb))) return rc;
      f->cnt[q.slot_ke] = nb;
    }
  f->g11 = h->prm.g[0];
  f->g21 = h->prm.g[1];
  f->g12 = h->prm.g[2];
  f->g22 = h->prm.g[3];
  return GPE_OK;
}

// y forward, z kinetic, y inverse (3D) or y kinetic (2D) for one array
// with the kinetic table pair (kcol, kax).
int spinor_rest_of_step(gpe_handle* h, double2* psi, double2* kcol, double2* kax) {
  int rc;
  if (h->ndims == 3) {
    if ((rc = sweep(h, 1, gpe::S_FWD, base_params(h, 1, psi, nullptr)))) return rc;
    SweepParams k = base_params(h, 2, psi, nullptr);
    k.kcol = kcol;
    k.kax = kax;
    if ((rc = sweep(h, 2, gpe::S_KIN, k))) return rc;
    if ((rc = sweep(h, 1, gpe::S_INV, base_params(h, 1, psi, nullptr)))) return rc;
  } else {
    SweepParams k = base_params(h, 1, psi, nullptr);
    k.kcol = kcol;
    k.kax = kax;
    if ((rc = sweep(h, 1, gpe::S_KIN, k))) return rc;
  }
  return GPE_OK;
}

// spinor_rest_of_step for both components, one launch per sweep: the second
// component runs on the second half of each grid (same operations per component).
// ke_groups (3D): the kinetic sweep also sums each component's kinetic energy of the
// result into SL_KZ / SL_KZ2; *ke_groups receives its work-groups per component.
static int spinor_rest_of_step_both(gpe_handle* h, double2* kcol, double2* kax, int* ke_groups = nullptr) {
  int rc;
  const int ka = h->ndims == 3 ? 2 : 1;
  if (h->ndims == 3) {
    SweepParams f = base_params(h, 1, h->psi[0], nullptr);
    f.psi2 = h->psi[1];
    if ((rc = sweep(h, 1, gpe::S_FWD, f))) return rc;
  }
  SweepParams k = base_params(h, ka, h->psi[0], nullptr);
  k.psi2 = h->psi[1];
  k.kcol = kcol;
  k.kax = kax;
  const bool ke = ke_groups && h->ndims == 3;
  if (ke) {
    k.kecol = h->kecolS;
    k.keax = h->kezS;
    k.slot_ke = gpe::SL_KZ;
    k.slot_ke2 = gpe::SL_KZ2;
  }
  int nb = 0;
  if ((rc = sweep(h, ka, ke ? gpe::S_KINE : gpe::S_KIN, k, &nb))) return rc;
  if (ke) *ke_groups = nb / 2;
  if (h->ndims == 3) {
    SweepParams v = base_params(h, 1, h->psi[0], nullptr);
    v.psi2 = h->psi[1];
    if ((rc = sweep(h, 1, gpe::S_INV, v))) return rc;
  }
  return GPE_OK;
}

// blk_in / blk_out (3D, reduce): the previous / next block of the same outer iteration
// hands the state over as the stored x spectrum (see the top of this file).
int spinor_block(gpe_handle* h, double time2, double2* tape, bool reduce, bool blk_in, bool blk_out) {
  int rc, ke_groups = 0;
  if (!blk_in) {
    // open: x forward of both components (with the pending renormalisation), then
    // exp(-kk dt/2)
    SweepParams p = base_params(h, 0, h->psi[0], nullptr);
    p.psi2 = h->psi[1];
    p.apply_scale = 1;
    if ((rc = sweep(h, 0, gpe::S_FWD, p))) return rc;
    if ((rc = spinor_rest_of_step_both(h, h->kcol_sp[0], h->kax_sp[0]))) return rc;
  } else {
    // open from S = F_x psi / nx: the y transform applies the pending renormalisation
    // and the kz table carries the nx an x forward transform would have supplied
    SweepParams f = base_params(h, 1, h->psi[0], nullptr);
    f.psi2 = h->psi[1];
    f.apply_scale = 1;
    if ((rc = sweep(h, 1, gpe::S_FWD, f))) return rc;
    SweepParams k = base_params(h, 2, h->psi[0], nullptr);
    k.psi2 = h->psi[1];
    k.kcol = h->kcol_sp[0];
    k.kax = h->kax_sp2;
    if ((rc = sweep(h, 2, gpe::S_KIN, k))) return rc;
    SweepParams v = base_params(h, 1, h->psi[0], nullptr);
    v.psi2 = h->psi[1];
    if ((rc = sweep(h, 1, gpe::S_INV, v))) return rc;
  }
  for (int i = 0; i < h->n_rec; ++i) {
    gpe::SpinorParams p = spinor_params(h);
    if (tape) {
      p.ust1 = tape + (size_t)(2 * i) * h->npts;
      p.ust2 = tape + (size_t)(2 * i + 1) * h->npts;
    }
    if ((rc = spinor_sweep(h, gpe::SP_PR, p, nullptr))) return rc;
    const bool last = i + 1 == h->n_rec;
    if ((rc = spinor_rest_of_step_both(h, last ? h->kcol_sp[1] : h->kcol, last ? h->kax_sp[1] : h->kax,
                                       last && reduce && h->ndims == 3 ? &ke_groups : nullptr)))
      return rc;
  }
  if (!blk_out) {
    SweepParams p = base_params(h, 0, h->psi[0], nullptr);
    p.psi2 = h->psi[1];
    if ((rc = sweep(h, 0, gpe::S_INV, p))) return rc;
  }
  if (!reduce) return GPE_OK;
  FinParams f = fin_params(h, gpe::FIN_SPINOR);
  if ((rc = spinor_reduce(h, &f, false, blk_out, ke_groups))) return rc;
  f.renorm = h->prm.gamma > 0 ? 1 : 0;
  f.NN0 = h->NN0;
  f.tau = h->prm.decay_rate;
  f.time2 = time2;
  HIPCHK(h, gpe::launch_finalize(f, h->stream));
  return GPE_OK;
}

}  // namespace gpe_spinor

using namespace gpe_spinor;

namespace gpe {

int spinor_measure(gpe_handle* h, double out[5]) {
  int rc;
  FinParams f = fin_params(h, FIN_SUMS);
  if ((rc = spinor_reduce(h, &f, false))) return rc;
  HIPCHK(h, launch_finalize(f, h->stream));
  DevState s;
  if ((rc = read_state(h, &s))) return rc;
  const double q = s.scale * s.scale;
  out[0] = q * s.sA * h->w;
  out[1] = q * s.sB * h->w;
  out[2] = q * q * s.sC * h->w;   // sum_jk g_jk |psi_j|^2 |psi_k|^2 w
  out[3] = q * s.sK * h->w;
  out[4] = q * 2.0 * s.pad[0] * h->w;
  return GPE_OK;
}

int spinor_split_begin(gpe_handle* h, double ddt, int niter_inner, double* muc_out) {
  int rc;
  if (!(h->prm.Ntotal > 0))
    return fail(h, GPE_EINVAL, "GPEtaskSpinor.solve_split needs Ntotal > 0 (spinor solve_split.m:30-38)");
  if (niter_inner / h->prm.n_recalc < 1)
    return fail(h, GPE_EINVAL, "niter_inner < n_recalc runs no block (spinor solve_split.m:67)");
  if ((rc = materialize(h))) return rc;
  for (int i = 0; i < 2; ++i) {
    const int ma = mid_axis(h);
    if (!h->kcol_sp[i]) HIPCHK(h, hipMalloc(&h->kcol_sp[i], (size_t)h->ax[ma].ncols * sizeof(double2)));
    if (!h->kax_sp[i]) HIPCHK(h, hipMalloc(&h->kax_sp[i], h->ax[ma].n * sizeof(double2)));
  }
  h->NN0 = h->prm.Ntotal;
  h->n_rec = h->prm.n_recalc;
  h->niter_inner = niter_inner;
  h->ddt = ddt;
  const std::complex<double> dt = cdiv({0.0, ddt}, {1.0, h->prm.gamma});   // solve_split.m:42
  h->dr = dt.real();
  h->di = dt.imag();
  if ((rc = build_kinetic(h, dt, h->kcol, h->kax))) return rc;                 // ekk  (43)
  if ((rc = build_kinetic(h, dt, h->kcol_sp[0], h->kax_sp[0], 0.5))) return rc;   // ekkp (44)
  if ((rc = build_kinetic(h, dt, h->kcol_sp[1], h->kax_sp[1], 1.0, -0.5))) return rc;  // ekk * ekkm (43, 45)
  if (h->ndims == 3) {
    // tables of the block ends inside an outer iteration (spinor_block): kax_sp[0] x nx
    // (kcol_sp[0] is rebuilt unchanged) and the y and z k^2/2/n tables x nx
    if (!h->kax_sp2) HIPCHK(h, hipMalloc(&h->kax_sp2, h->ax[2].n * sizeof(double2)));
    if ((rc = build_kinetic(h, dt, h->kcol_sp[0], h->kax_sp2, 0.5, 0.0, nullptr, (double)h->nx))) return rc;
    if (!h->kecolS) {
      const std::vector<double>&kx = h->ax[0].k, &ky = h->ax[1].k, &kz = h->ax[2].k;
      const double np = (double)h->npts;
      std::vector<double> wc((size_t)h->nx * h->ny), wz(h->nz);
      for (int ix = 0; ix < h->nx; ++ix)
        for (int iy = 0; iy < h->ny; ++iy)
          wc[(size_t)iy + (size_t)h->ny * ix] = (kx[ix] * kx[ix] + ky[iy] * ky[iy]) / 2 * np;
      for (int iz = 0; iz < h->nz; ++iz) wz[iz] = kz[iz] * kz[iz] / 2 * np;
      HIPCHK(h, hipMalloc(&h->kecolS, wc.size() * sizeof(double)));
      HIPCHK(h, hipMemcpy(h->kecolS, wc.data(), wc.size() * sizeof(double), hipMemcpyHostToDevice));
      HIPCHK(h, hipMalloc(&h->kezS, wz.size() * sizeof(double)));
      HIPCHK(h, hipMemcpy(h->kezS, wz.data(), wz.size() * sizeof(double), hipMemcpyHostToDevice));
    }
    for (int a = 1; a < 3; ++a) {
      if (h->keax_nx[a]) continue;
      const Axis& ax = h->ax[a];
      std::vector<double> t(ax.n);
      for (int i = 0; i < ax.n; ++i) t[i] = ax.k[i] * ax.k[i] / 2 / (double)ax.n * (double)h->nx;
      HIPCHK(h, hipMalloc(&h->keax_nx[a], t.size() * sizeof(double)));
      HIPCHK(h, hipMemcpy(h->keax_nx[a], t.data(), t.size() * sizeof(double), hipMemcpyHostToDevice));
    }
  }
  double m[5];
  if ((rc = spinor_measure(h, m))) return rc;
  // 52-57: mu = sum_j real(inner(phi_j, H phi_j)) / NN0
  const double mu = (((m[1] + m[3]) + m[4]) + m[2]) / h->NN0;
  DevState s;
  if ((rc = read_state(h, &s))) return rc;
  s.mu = mu;
  s.scale = 1.0;
  s.NNN = h->NN0;
  s.muc = mu;
  s.ncur = h->NN0;
  if ((rc = write_state(h, s))) return rc;
  if (muc_out) *muc_out = mu;
  h->begun = true;
  return GPE_OK;
}

int spinor_split_block(gpe_handle* h, double time2, double* muc, double* ncur) {
  int rc;
  rabi_scalars(h);
  if ((rc = spinor_block(h, time2, nullptr, true, false, false))) return rc;
  DevState s;
  if ((rc = read_state(h, &s))) return rc;
  if ((rc = check_finite(h, s, "gpe_split_block"))) return rc;
  if (muc) *muc = s.muc;
  if (ncur) *ncur = s.ncur;
  return GPE_OK;
}

int spinor_split_outer(gpe_handle* h, int j, double* time2_out, double* muc, double* ncur) {
  int rc;
  rabi_scalars(h);                                            // solve_split.m:61-65
  const double dt_outer = h->ddt * h->niter_inner;
  const double time = (double)(j - 1) * dt_outer;             // 66
  double time2 = time;
  const int nblocks = h->niter_inner / h->n_rec;              // 67 (colon truncates)
  // graph replay as in gpe_split_outer; the coupling scalars, its array's magnitude
  // bound (the Taylor choice) and the block-end potential are baked into the launches
  const double key[8] = {h->cpl_array ? 1.0 : 0.0, h->cpl_re, h->cpl_im, h->cpl_max2, h->use_vmu ? 1.0 : 0.0,
                         (double)reinterpret_cast<uintptr_t>(h->cpl), 0.0, 0.0};
  rc = run_outer(h, h->prm.decay_rate == 0, key, [&] {
    for (int jj = 1; jj <= nblocks; ++jj) {
      const bool d3 = h->ndims == 3;
      const int r = spinor_block(h, time + (double)(jj - 1) * h->ddt * (double)h->n_rec, nullptr, true,
                                 d3 && jj > 1, d3 && jj < nblocks);   // 69
      if (r) return r;
    }
    return GPE_OK;
  });
  if (rc) return rc;
  time2 = time + (double)(nblocks - 1) * h->ddt * (double)h->n_rec;
  DevState s;
  if ((rc = read_state(h, &s))) return rc;
  if ((rc = check_finite(h, s, "gpe_split_outer"))) return rc;
  if (time2_out) *time2_out = time2;
  if (muc) *muc = s.muc;
  if (ncur) *ncur = s.ncur;
  return GPE_OK;
}

}  // namespace gpe
