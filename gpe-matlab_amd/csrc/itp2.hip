// Two-component imaginary-time ground states: GPEtaskSpinor.groundstate_itp
// (lib/@GPEtaskSpinor/groundstate_itp.m:1-145) and GPEtask2comp.groundstate_itp
// (lib/@GPEtask2comp/groundstate_itp.m:1-157), omega = 0.
//
// One iteration, per component k (3D; 2D replaces the y/z sweeps by one y sweep):
//   SP_ITPA (x)   renormalise by the last factor, store P_k = exp(-(V_k + sum_j g_kj n_j) dt)
//                 from that density (49-54), forward x-FFT
//   y/z sweeps    rest of the kinetic half step exp(-kk dt / 2 / M_k)      (67-68)
//   SP_ITPB (x)   inverse x-FFT, P_k (R psi)_k, forward x-FFT              (70-71)
//   y/z sweeps    second kinetic half step                                 (78-79)
//   SP_ITPC (x)   inverse x-FFT, norms of both components                  (82)
//   finalize      renormalisation factors, MU / MU2 histories              (83-88)
//   [SP_KEX + y/z kinetic sums, finalize: applyham / energy histories]     (91-100)
// Every 10 iterations the host reads the convergence series and applies the
// step-size cascade dt -> dt / 1.5 (102-123).
#include <algorithm>
#include <cmath>
#include <complex>
#include <vector>

#include "handle.h"
#include "spinor.h"

using namespace gpe_detail;
using namespace gpe_spinor;

namespace {

int alloc_tables(gpe_handle* h) {
  const int ma = mid_axis(h);
  for (int i = 0; i < 2; ++i) {
    if (!h->kcol_sp[i]) HIPCHK(h, hipMalloc(&h->kcol_sp[i], (size_t)h->ax[ma].ncols * sizeof(double2)));
    if (!h->kax_sp[i]) HIPCHK(h, hipMalloc(&h->kax_sp[i], h->ax[ma].n * sizeof(double2)));
  }
  return GPE_OK;
}

// exp(-kk dt / 2 / M_k) for each component (GPEtaskSpinor: M = 1 for both) and the
// Rabi factors for the real step dt (43-46, 112-119).
int set_step(gpe_handle* h, double dt, const gpe_itp2_args* a) {
  int rc;
  h->dr = dt;
  h->di = 0.0;
  const double m1 = a->mixture ? a->M1 : 1.0, m2 = a->mixture ? a->M2 : 1.0;
  if ((rc = build_kinetic(h, {dt, 0.0}, h->kcol_sp[0], h->kax_sp[0], 0.5 / m1))) return rc;
  if ((rc = build_kinetic(h, {dt, 0.0}, h->kcol_sp[1], h->kax_sp[1], 0.5 / m2))) return rc;
  rabi_scalars(h);
  return GPE_OK;
}

int x_sweep(gpe_handle* h, int seq, bool mixture, int* nb) {
  gpe::SpinorParams p = spinor_params(h);
  p.pbuf = h->itp2_P;
  p.mixture = mixture ? 1 : 0;
  if (mixture) p.cpl = nullptr;
  return spinor_sweep(h, seq, p, nb);
}

}  // namespace

extern "C" int gpe_itp2(gpe_handle* h, const gpe_itp2_args* a, double* MU, double* MU2, double* EE, int* n_iter,
                        double* dt_final, int* converged) {
  int rc;
  if ((rc = ensure_ready(h))) return rc;
  if (!a || !MU || a->max_iter < 1) return fail(h, GPE_EINVAL, "bad two-component ITP arguments");
  if (h->prm.ncomp != 2) return fail(h, GPE_EINVAL, "gpe_itp2 needs two components");
  if (h->omega != 0.0) return fail(h, GPE_EINVAL, "the rotating frame is supported for one component");
  if (a->mixture && !(a->M1 > 0 && a->M2 > 0)) return fail(h, GPE_EINVAL, "masses M1, M2 must be positive");
  if (!(a->dt > 0) || !std::isfinite(a->dt)) return fail(h, GPE_EINVAL, "dt must be positive");
  const bool mix = a->mixture != 0;
  const bool want_mu2 = !mix && a->nargout >= 3;   // spinor MU2 from applyham (91-95)
  const bool want_ee = a->nargout >= 4;            // get_energy (96-100)
  if ((rc = materialize(h))) return rc;
  if ((rc = alloc_tables(h))) return rc;
  if (!h->itp2_P) HIPCHK(h, hipMalloc(&h->itp2_P, 2 * (size_t)h->npts * sizeof(double)));
  if (h->itp2_cap < a->max_iter) {
    for (double*& b : h->itp2_hist) {
      (void)hipFree(b);
      b = nullptr;
    }
    for (double*& b : h->itp2_hist) HIPCHK(h, hipMalloc(&b, (size_t)a->max_iter * sizeof(double)));
    h->itp2_cap = a->max_iter;
  }
  double* dMU = h->itp2_hist[0];
  double* dMU2 = h->itp2_hist[1];
  double* dEE = h->itp2_hist[2];
  // the convergence series (96-101 / 98-102): energy, else MU2 (spinor: applyham or MU)
  const double* series = want_ee ? dEE : (mix || want_mu2 ? dMU2 : dMU);
  {
    DevState s;
    if ((rc = read_state(h, &s))) return rc;
    s.sc[0] = s.sc[1] = 1.0;
    if ((rc = write_state(h, s))) return rc;
  }
  double dt = a->dt;
  if ((rc = set_step(h, dt, a))) return rc;
  int i = 1, iswitch = 20, conv = 1;   // 39-40
  std::vector<double> ebuf(10);
  for (;;) {
    if ((rc = x_sweep(h, gpe::SP_ITPA, mix, nullptr))) return rc;
    for (int comp = 0; comp < 2; ++comp)
      if ((rc = spinor_rest_of_step(h, h->psi[comp], h->kcol_sp[comp], h->kax_sp[comp]))) return rc;
    if ((rc = x_sweep(h, gpe::SP_ITPB, mix, nullptr))) return rc;
    for (int comp = 0; comp < 2; ++comp)
      if ((rc = spinor_rest_of_step(h, h->psi[comp], h->kcol_sp[comp], h->kax_sp[comp]))) return rc;
    int nb = 0;
    if ((rc = x_sweep(h, gpe::SP_ITPC, mix, &nb))) return rc;
    FinParams f = fin_params(h, gpe::FIN_ITP2);
    f.cnt[gpe::SL_A] = f.cnt[gpe::SL_B] = nb;
    f.mixture = mix ? 1 : 0;
    f.n1t = mix ? a->N1 : h->prm.Ntotal;
    f.n2t = a->N2;
    f.dt = dt;
    f.MU = dMU;
    f.MU2 = dMU2;
    f.EE = dEE;
    f.iter = i - 1;
    HIPCHK(h, gpe::launch_finalize(f, h->stream));
    if (want_mu2 || want_ee) {
      FinParams m = fin_params(h, gpe::FIN_ITP2M);
      if (mix) {   // GPEtask2comp.applyham1c has no coupling term (GPEtask2comp.m:51-64)
        const bool arr = h->cpl_array;
        const double cr = h->cpl_re, ci = h->cpl_im;
        h->cpl_array = false;
        h->cpl_re = h->cpl_im = 0.0;
        rc = spinor_reduce(h, &m, true);
        h->cpl_array = arr;
        h->cpl_re = cr;
        h->cpl_im = ci;
      } else {
        rc = spinor_reduce(h, &m, true);
      }
      if (rc) return rc;
      m.Ntotal = h->prm.Ntotal;
      m.MU2 = dMU2;
      m.EE = dEE;
      m.which = (want_mu2 ? 1 : 0) | (want_ee ? 2 : 0);
      m.iter = i - 1;
      HIPCHK(h, gpe::launch_finalize(m, h->stream));
    }
    if ((i - iswitch) > 10 && i % 10 == 0) {   // 102-123
      HIPCHK(h, hipMemcpyAsync(ebuf.data(), series + (i - 10), 10 * sizeof(double), hipMemcpyDeviceToHost,
                               h->stream));
      HIPCHK(h, hipStreamSynchronize(h->stream));
      const double e_i = ebuf[9], e_i1 = ebuf[8], e_i9 = ebuf[0];
      const double delta = (std::fabs(e_i - e_i9) / 9 + std::fabs(e_i - e_i1)) / dt;
      if (!std::isfinite(delta))
        return fail(h, GPE_ENONFINITE, "gpe_itp2: iteration %d is not finite (convergence series %g)", i, e_i);
      if (delta < a->eps) {
        if (dt < a->eps || dt < 5e-6) break;
        dt = dt / 1.5;
        if ((rc = set_step(h, dt, a))) return rc;
        iswitch = i;
      }
    }
    if (i >= a->max_iter) {   // 124-127
      conv = 0;
      break;
    }
    ++i;
  }
  // phi = mu * stored state (86-88 of the last iteration)
  for (int comp = 0; comp < 2; ++comp)
    HIPCHK(h, gpe::launch_scale(h->psi[comp], h->st, comp == 0 ? 2 : 3, h->npts, elem_blocks(h->npts), h->stream,
                                false));
  // histories hold entries 1..i-1 (131-142)
  const size_t nh = (size_t)(i - 1) * sizeof(double);
  if (nh) {
    HIPCHK(h, hipMemcpyAsync(MU, dMU, nh, hipMemcpyDeviceToHost, h->stream));
    if (MU2 && (mix || want_mu2)) HIPCHK(h, hipMemcpyAsync(MU2, dMU2, nh, hipMemcpyDeviceToHost, h->stream));
    if (EE && want_ee) HIPCHK(h, hipMemcpyAsync(EE, dEE, nh, hipMemcpyDeviceToHost, h->stream));
  }
  HIPCHK(h, hipStreamSynchronize(h->stream));
  {
    DevState s;
    if ((rc = read_state(h, &s))) return rc;
    s.sc[0] = s.sc[1] = 1.0;
    if ((rc = write_state(h, s))) return rc;
  }
  if (n_iter) *n_iter = i;
  if (dt_final) *dt_final = dt;
  if (converged) *converged = conv;
  h->begun = false;
  if (nh && !std::isfinite(MU[i - 2])) return fail(h, GPE_ENONFINITE, "gpe_itp2: chemical potential is not finite");
  return GPE_OK;
}
