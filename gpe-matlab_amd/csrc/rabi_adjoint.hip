// Rabi-coupling adjoint: gradient of the final population imbalance
// L = (N1 - N2) / Ntotal of a two-component run (GPEtaskSpinor.solve_split with
// gamma = 0, lib/@GPEtaskSpinor/solve_split.m:1-121) with respect to (om0, kw) of
// the coupling Omega(x) = om0 exp(i kw x).
//
// The spinor step is not explicitly invertible (the Rabi rotation mixes the
// components before the density-dependent phase), so the backward pass does not
// reconstruct the state.  The run is cut into segments of `checkpoint_blocks`
// blocks (about sqrt(T) steps).  The forward pass stores the real-space state
// before each step ("tape") for as many trailing segments as the tape budget
// holds, a checkpoint of both components at the start of every other segment, and
// the chemical potential of every block.  The backward pass walks the segments in
// reverse, recomputing the tape of a segment without one from its checkpoint, and
// runs the transposed steps of each segment in reverse.  The chemical potential
// multiplies both components by a common phase and does not change L, so it is
// replayed from the forward run and not differentiated.
#include <algorithm>
#include <cmath>

#include "handle.h"
#include "reduce_dev.h"
#include "spinor.h"

using namespace gpe_detail;
using namespace gpe_spinor;

namespace gpe {

// Omega(x) = om0 exp(i kw x) on the [ny, nx, nz] grid
__global__ void __launch_bounds__(256) rabi_coupling_kernel(double2* __restrict__ cpl, const double* __restrict__ xc,
                                                            double om0, double kw, long long ny, long long nx,
                                                            long long n) {
  for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long long)gridDim.x * 256) {
    double s, c;
    sincos(kw * xc[(i / ny) % nx], &s, &c);
    cpl[i] = make_double2(om0 * c, om0 * s);
  }
}

// Partials of sum (|psi1|^2 - |psi2|^2) (slot A, differenced per point, so a small
// imbalance of two large populations keeps its digits) and sum (|psi1|^2 + |psi2|^2)
// (slot B); with lam1 set, also the final adjoint lam1 = f psi1, lam2 = -f psi2
// (2 dL/d conj(psi)).
__global__ void __launch_bounds__(256) rabi_final_kernel(const double2* __restrict__ p1, const double2* __restrict__ p2,
                                                         double2* __restrict__ lam1, double2* __restrict__ lam2,
                                                         double f, long long n, double* partials, int maxwg) {
  __shared__ double red[4];
  double s1 = 0.0, s2 = 0.0;
  for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long long)gridDim.x * 256) {
    const double2 a = p1[i], b = p2[i];
    const double n1 = a.x * a.x + a.y * a.y, n2 = b.x * b.x + b.y * b.y;
    s1 += n1 - n2;
    s2 += n1 + n2;
    if (lam1) {
      lam1[i] = make_double2(f * a.x, f * a.y);
      lam2[i] = make_double2(-f * b.x, -f * b.y);
    }
  }
  const double r1 = block_sum(s1, red);
  const double r2 = block_sum(s2, red);
  if (threadIdx.x == 0) {
    partials[SL_A * maxwg + blockIdx.x] = r1;
    partials[SL_B * maxwg + blockIdx.x] = r2;
  }
}

// out[k] = sum of gacc[k * maxwg + i], i < nb, in a fixed order
__global__ void __launch_bounds__(256) rabi_grad_finalize_kernel(const double* gacc, int maxwg, int nb, double* out) {
  __shared__ double red[4];
  for (int k = 0; k < 2; ++k) {
    double acc = 0.0;
    for (int i = threadIdx.x; i < nb; i += 256) acc += gacc[k * maxwg + i];
    const double r = block_sum(acc, red);
    if (threadIdx.x == 0) out[k] = r;
  }
}

}  // namespace gpe

namespace {

struct Plan {
  int B = 0;     // blocks of n_recalc steps
  int cpb = 1;   // blocks per checkpoint segment
  int nck = 0;   // segments
  int keep = 1;  // trailing segments whose pre-step states the forward pass stores
};

int check_args(gpe_handle* h, const gpe_rabi_args* a, Plan* c) {
  if (!a) return fail(h, GPE_EINVAL, "null Rabi control args");
  if (!h->have_params) return fail(h, GPE_ESTATE, "gpe_set_params has not been called");
  if (h->prm.ncomp != 2) return fail(h, GPE_EINVAL, "the Rabi adjoint needs two components");
  if (h->prm.gamma != 0.0) return fail(h, GPE_EINVAL, "the Rabi adjoint needs gamma = 0");
  if (h->omega != 0.0) return fail(h, GPE_EINVAL, "the Rabi adjoint does not support the rotating frame");
  if (h->prm.nlincpl != 0.0) return fail(h, GPE_EINVAL, "the Rabi adjoint needs nlincpl = 0");
  if (!(h->prm.Ntotal > 0)) return fail(h, GPE_EINVAL, "the Rabi adjoint needs Ntotal > 0");
  if (!std::isfinite(a->ddt) || !std::isfinite(a->omega0) || !std::isfinite(a->k))
    return fail(h, GPE_EINVAL, "non-finite ddt, omega0 or k");
  const int nr = h->prm.n_recalc;
  if (nr < 1 || a->nsteps <= 0 || a->nsteps % nr)
    return fail(h, GPE_EINVAL, "nsteps must be a positive multiple of n_recalc");
  c->B = a->nsteps / nr;
  c->cpb = a->checkpoint_blocks > 0 ? std::min(a->checkpoint_blocks, c->B)
                                    : std::max(1, (int)std::lround(std::sqrt((double)a->nsteps) / nr));
  c->nck = (c->B + c->cpb - 1) / c->cpb;
  return GPE_OK;
}

template <class T>
int grow(gpe_handle* h, T** p, long long* cap, long long need) {
  if (*cap >= need) return GPE_OK;
  (void)hipFree(*p);
  *p = nullptr;
  *cap = 0;
  if (hipMalloc(p, (size_t)need * sizeof(T)) != hipSuccess)
    return fail(h, GPE_ENOMEM, "Rabi adjoint: cannot allocate %lld bytes", need * (long long)sizeof(T));
  *cap = need;
  return GPE_OK;
}

// Segments whose tape fits the budget: tape_gib > 0 GiB, 0 = half the free device
// memory, < 0 = one segment (the minimum the recomputation needs).
int plan_tape(gpe_handle* h, const gpe_rabi_args* a, Plan* c) {
  const double seg = (double)c->cpb * h->prm.n_recalc * 2.0 * (double)h->npts * sizeof(double2);
  double budget = 0.0;
  if (a->tape_gib > 0) {
    budget = a->tape_gib * 1073741824.0;
  } else if (a->tape_gib == 0) {
    size_t fr = 0, tot = 0;
    HIPCHK(h, hipMemGetInfo(&fr, &tot));
    budget = 0.5 * ((double)fr + (double)h->rb_ntape * sizeof(double2));
  }
  c->keep = (int)std::min<double>(c->nck, std::max(1.0, std::floor(budget / seg)));
  return GPE_OK;
}

int ensure_buffers(gpe_handle* h, Plan& c, const gpe_rabi_args* a, bool grad) {
  int rc;
  const long long n = h->npts;
  if ((rc = grow(h, &h->rb_ckpt, &h->rb_nckpt, (grad ? c.nck : 1) * 2 * n))) return rc;
  if ((rc = grow(h, &h->rb_mu, &h->rb_nmu, (long long)c.B))) return rc;
  if (!h->cpl) HIPCHK(h, hipMalloc(&h->cpl, n * sizeof(double2)));
  if (!h->rb_xc) {
    HIPCHK(h, hipMalloc(&h->rb_xc, h->nx * sizeof(double)));
    HIPCHK(h, hipMemcpy(h->rb_xc, h->ax[0].x.data(), h->nx * sizeof(double), hipMemcpyHostToDevice));
  }
  if (!grad) return GPE_OK;
  if ((rc = plan_tape(h, a, &c))) return rc;
  const long long seg = (long long)c.cpb * h->prm.n_recalc * 2 * n;
  if ((rc = grow(h, &h->rb_tape, &h->rb_ntape, c.keep * seg))) {
    if (rc != GPE_ENOMEM || c.keep == 1) return rc;
    c.keep = 1;   // the budget did not fit after all: store one segment, recompute the rest
    if ((rc = grow(h, &h->rb_tape, &h->rb_ntape, seg))) return rc;
  }
  for (int i = 0; i < 2; ++i)
    if (!h->rb_lam[i]) HIPCHK(h, hipMalloc(&h->rb_lam[i], n * sizeof(double2)));
  if (!h->rb_gacc) HIPCHK(h, hipMalloc(&h->rb_gacc, 2 * (size_t)h->maxwg * sizeof(double)));
  if (!h->rb_out) HIPCHK(h, hipMalloc(&h->rb_out, 2 * sizeof(double)));
  const int ma = mid_axis(h);
  for (int i = 0; i < 3; ++i) {
    if (!h->kcol_rc[i]) HIPCHK(h, hipMalloc(&h->kcol_rc[i], (size_t)h->ax[ma].ncols * sizeof(double2)));
    if (!h->kax_rc[i]) HIPCHK(h, hipMalloc(&h->kax_rc[i], h->ax[ma].n * sizeof(double2)));
  }
  return GPE_OK;
}

int copy2(gpe_handle* h, double2* dst, double2* const src[2], bool to_pair) {
  const size_t bytes = h->npts * sizeof(double2);
  for (int comp = 0; comp < 2; ++comp) {
    double2* d = to_pair ? dst + comp * h->npts : src[comp];
    const double2* s = to_pair ? src[comp] : dst + comp * h->npts;
    HIPCHK(h, hipMemcpyAsync(d, s, bytes, hipMemcpyDeviceToDevice, h->stream));
  }
  return GPE_OK;
}
int save_state(gpe_handle* h, double2* slot) { return copy2(h, slot, h->psi, true); }
int load_state(gpe_handle* h, double2* slot) { return copy2(h, slot, h->psi, false); }

int set_mu(gpe_handle* h, int b) {
  HIPCHK(h, hipMemcpyAsync(&h->st->mu, h->rb_mu + b, sizeof(double), hipMemcpyDeviceToDevice, h->stream));
  return GPE_OK;
}

// Coupling array, initial chemical potential and kinetic tables (spinor solve_split.m:40-65).
int begin(gpe_handle* h, const gpe_rabi_args* a) {
  int rc;
  hipLaunchKernelGGL(gpe::rabi_coupling_kernel, dim3(elem_blocks(h->npts)), dim3(256), 0, h->stream, h->cpl,
                     h->rb_xc, a->omega0, a->k, (long long)h->ny, (long long)h->nx, (long long)h->npts);
  HIPCHK(h, hipGetLastError());
  h->cpl_array = true;
  h->cpl_max2 = a->omega0 * a->omega0;
  if ((rc = gpe::spinor_split_begin(h, a->ddt, h->prm.n_recalc, nullptr))) return rc;
  rabi_scalars(h);
  return GPE_OK;
}

// Tape of block b: the last c.keep segments have their own slots; the segments
// before them are recomputed into slot 0 once those are consumed.
double2* tape_of(gpe_handle* h, const Plan& c, int b) {
  const int seg = b / c.cpb;
  const int slot = seg >= c.nck - c.keep ? seg - (c.nck - c.keep) : 0;
  return h->rb_tape + ((size_t)slot * c.cpb + (size_t)(b % c.cpb)) * h->prm.n_recalc * 2 * h->npts;
}

// Forward run; with grad, keeps the checkpoints of the segments it does not store
// and the tape of the last c.keep segments.
int forward(gpe_handle* h, const Plan& c, bool grad) {
  int rc;
  for (int b = 0; b < c.B; ++b) {
    const int seg = b / c.cpb;
    const bool stored = grad && seg >= c.nck - c.keep;
    if (grad && !stored && b % c.cpb == 0 && b > 0)
      if ((rc = save_state(h, h->rb_ckpt + (size_t)seg * 2 * h->npts))) return rc;
    HIPCHK(h, hipMemcpyAsync(h->rb_mu + b, &h->st->mu, sizeof(double), hipMemcpyDeviceToDevice, h->stream));
    if ((rc = spinor_block(h, 0.0, stored ? tape_of(h, c, b) : nullptr, true))) return rc;
  }
  return GPE_OK;
}

// loss (and, with lam, the final adjoint) from the state at T
int final_loss(gpe_handle* h, bool lam, double* loss) {
  int rc;
  const int nb = elem_blocks(h->npts);
  const double f = 2.0 * h->w / h->prm.Ntotal;
  hipLaunchKernelGGL(gpe::rabi_final_kernel, dim3(nb), dim3(256), 0, h->stream, h->psi[0], h->psi[1],
                     lam ? h->rb_lam[0] : nullptr, lam ? h->rb_lam[1] : nullptr, f, (long long)h->npts, h->partials,
                     h->maxwg);
  HIPCHK(h, hipGetLastError());
  FinParams fp = fin_params(h, gpe::FIN_SUMS);
  fp.cnt[gpe::SL_A] = fp.cnt[gpe::SL_B] = nb;
  HIPCHK(h, gpe::launch_finalize(fp, h->stream));
  DevState s;
  if ((rc = read_state(h, &s))) return rc;
  if (loss) *loss = s.sA * h->w / h->prm.Ntotal;
  return GPE_OK;
}

int adj_sweep(gpe_handle* h, const gpe_rabi_args* a, double2* tape, int* nb) {
  gpe::SpinorParams p = spinor_params(h);
  p.psi1 = h->rb_lam[0];
  p.psi2 = h->rb_lam[1];
  p.ust1 = tape;
  p.ust2 = tape + h->npts;
  p.xc = h->rb_xc;
  p.om0 = a->omega0;
  p.kw = a->k;
  p.gacc = h->rb_gacc;
  return spinor_sweep(h, gpe::SP_ADJ, p, nb);
}

// Transposed block: the forward block's sweeps in reverse with conjugate kinetic
// tables (the unnormalised x transforms are each other's adjoints).
int adjoint_block(gpe_handle* h, const gpe_rabi_args* a, double2* tape, int* nb) {
  int rc;
  const int nr = h->prm.n_recalc;
  for (int comp = 0; comp < 2; ++comp)
    if ((rc = sweep(h, 0, gpe::S_FWD, base_params(h, 0, h->rb_lam[comp], nullptr)))) return rc;
  for (int i = nr - 1; i >= 0; --i) {
    const int t = i + 1 == nr ? 2 : 1;
    for (int comp = 0; comp < 2; ++comp)
      if ((rc = spinor_rest_of_step(h, h->rb_lam[comp], h->kcol_rc[t], h->kax_rc[t]))) return rc;
    if ((rc = adj_sweep(h, a, tape + (size_t)(2 * i) * h->npts, nb))) return rc;
  }
  for (int comp = 0; comp < 2; ++comp) {
    if ((rc = spinor_rest_of_step(h, h->rb_lam[comp], h->kcol_rc[0], h->kax_rc[0]))) return rc;
    if ((rc = sweep(h, 0, gpe::S_INV, base_params(h, 0, h->rb_lam[comp], nullptr)))) return rc;
  }
  return GPE_OK;
}

int finish(gpe_handle* h) {
  int rc;
  if ((rc = load_state(h, h->rb_ckpt))) return rc;
  h->begun = false;   // the next solve_split starts with gpe_split_begin
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return GPE_OK;
}

}  // namespace

extern "C" {

int gpe_rabi_forward(gpe_handle* h, const gpe_rabi_args* a, double* loss, double* psi_out) {
  int rc;
  Plan c;
  if (!h) return GPE_EINVAL;
  if ((rc = ensure_ready(h))) return rc;
  if ((rc = check_args(h, a, &c))) return rc;
  if ((rc = ensure_buffers(h, c, a, false))) return rc;
  if ((rc = materialize(h))) return rc;
  if ((rc = save_state(h, h->rb_ckpt))) return rc;
  if ((rc = begin(h, a))) return rc;
  if ((rc = forward(h, c, false))) return rc;
  if ((rc = final_loss(h, false, loss))) return rc;
  if (psi_out)
    for (int comp = 0; comp < 2; ++comp)
      HIPCHK(h, hipMemcpyAsync((double2*)psi_out + (size_t)comp * h->npts, h->psi[comp], h->npts * sizeof(double2),
                               hipMemcpyDeviceToHost, h->stream));
  return finish(h);
}

int gpe_rabi_grad(gpe_handle* h, const gpe_rabi_args* a, double* loss, double* grad) {
  int rc;
  Plan c;
  if (!h || !grad) return h ? fail(h, GPE_EINVAL, "null grad") : GPE_EINVAL;
  if ((rc = ensure_ready(h))) return rc;
  if ((rc = check_args(h, a, &c))) return rc;
  if ((rc = ensure_buffers(h, c, a, true))) return rc;
  if ((rc = materialize(h))) return rc;
  if ((rc = save_state(h, h->rb_ckpt))) return rc;
  if ((rc = begin(h, a))) return rc;
  const std::complex<double> cdt(0.0, -a->ddt);   // conjugate of dt = i ddt
  if ((rc = build_kinetic(h, cdt, h->kcol_rc[0], h->kax_rc[0], 0.5))) return rc;
  if ((rc = build_kinetic(h, cdt, h->kcol_rc[1], h->kax_rc[1]))) return rc;
  if ((rc = build_kinetic(h, cdt, h->kcol_rc[2], h->kax_rc[2], 1.0, -0.5))) return rc;
  if ((rc = forward(h, c, true))) return rc;
  if ((rc = final_loss(h, true, loss))) return rc;
  HIPCHK(h, hipMemsetAsync(h->rb_gacc, 0, 2 * (size_t)h->maxwg * sizeof(double), h->stream));
  int nb = 0;
  for (int j = c.nck - 1; j >= 0; --j) {
    const int b0 = j * c.cpb, b1 = std::min(c.B, b0 + c.cpb);
    if (j < c.nck - c.keep) {   // recompute the segment from its checkpoint, keeping the pre-step states
      if ((rc = load_state(h, h->rb_ckpt + (size_t)j * 2 * h->npts))) return rc;
      for (int b = b0; b < b1; ++b) {
        if ((rc = set_mu(h, b))) return rc;
        if ((rc = spinor_block(h, 0.0, tape_of(h, c, b), false))) return rc;
      }
    }
    for (int b = b1 - 1; b >= b0; --b) {
      if ((rc = set_mu(h, b))) return rc;
      if ((rc = adjoint_block(h, a, tape_of(h, c, b), &nb))) return rc;
    }
  }
  hipLaunchKernelGGL(gpe::rabi_grad_finalize_kernel, dim3(1), dim3(256), 0, h->stream, h->rb_gacc, h->maxwg, nb,
                     h->rb_out);
  HIPCHK(h, hipGetLastError());
  HIPCHK(h, hipMemcpyAsync(grad, h->rb_out, 2 * sizeof(double), hipMemcpyDeviceToHost, h->stream));
  return finish(h);
}

}  // extern "C"
