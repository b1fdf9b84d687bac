// Adjoint split-step driver: gradient of 1 - |<psi_target|psi(T)>|^2 with respect
// to piecewise-constant amplitudes of a control potential.
//
// Forward: the solve_split block sequence (gamma = 0, fixed mu) with the state
// copied to a checkpoint every `checkpoint_blocks` blocks (about sqrt(T) steps).
// Backward: conjugated steps in reverse on two arrays at once -- the state,
// reconstructed by inverting each map, and the adjoint -- re-anchored to the
// stored checkpoint at every checkpoint boundary.  dL/dV is accumulated per
// point and reduced against the basis fields at every interval boundary.
#include <algorithm>

#include "adjoint.h"
#include "handle.h"
#include "reduce_dev.h"

using namespace gpe_detail;

namespace gpe {

#define GPE_DECL(N) hipError_t launch_adj_x_##N(int, int, const AdjParams&, hipStream_t, int*);
GPE_DECL(8) GPE_DECL(12) GPE_DECL(16) GPE_DECL(24) GPE_DECL(32) GPE_DECL(48) GPE_DECL(64) GPE_DECL(96)
GPE_DECL(128) GPE_DECL(192) GPE_DECL(256) GPE_DECL(384) GPE_DECL(512)
#undef GPE_DECL

hipError_t launch_adj_x(int n, int seq, int var, const AdjParams& p, hipStream_t s, int* nblocks) {
  switch (n) {
#define GPE_C(N) case N: return launch_adj_x_##N(seq, var, p, s, nblocks);
    GPE_C(8) GPE_C(12) GPE_C(16) GPE_C(24) GPE_C(32) GPE_C(48) GPE_C(64) GPE_C(96) GPE_C(128) GPE_C(192)
    GPE_C(256) GPE_C(384) GPE_C(512)
#undef GPE_C
    default:
      return hipErrorInvalidValue;
  }
}

constexpr int MAXBASIS = 16;
struct Coef {
  double c[MAXBASIS];
};

// V = V0 + c_0 G_0 + c_1 G_1 + ... (left to right, as Vtrap + Vtd)
__global__ void __launch_bounds__(256) build_v_kernel(double* __restrict__ V, const double* __restrict__ V0,
                                                      const double* __restrict__ G, Coef c, int nb, long long n) {
  for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long long)gridDim.x * 256) {
    double v = V0[i];
    for (int k = 0; k < nb; ++k) v = v + c.c[k] * G[k * n + i];
    V[i] = v;
  }
}

// partials of sum conj(a) b (slots A: real, B: imaginary) or, with diff != 0,
// of sum |a - b|^2 (A) and sum |b|^2 (B)
__global__ void __launch_bounds__(256) pair_reduce_kernel(const double2* __restrict__ a, const double2* __restrict__ b,
                                                          long long n, int diff, double* partials, int maxwg) {
  __shared__ double red[4];
  double s0 = 0.0, s1 = 0.0;
  for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long long)gridDim.x * 256) {
    const double2 x = a[i], y = b[i];
    if (diff) {
      const double dx = x.x - y.x, dy = x.y - y.y;
      s0 += dx * dx + dy * dy;
      s1 += y.x * y.x + y.y * y.y;
    } else {
      s0 += x.x * y.x + x.y * y.y;
      s1 += x.x * y.y - x.y * y.x;
    }
  }
  const double r0 = block_sum(s0, red);
  const double r1 = block_sum(s1, red);
  if (threadIdx.x == 0) {
    partials[SL_A * maxwg + blockIdx.x] = r0;
    partials[SL_B * maxwg + blockIdx.x] = r1;
  }
}

__global__ void __launch_bounds__(256) lam_init_kernel(double2* __restrict__ lam, const double2* __restrict__ t,
                                                       double2 coef, long long n) {
  for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long long)gridDim.x * 256) {
    const double2 x = t[i];
    lam[i] = make_double2(coef.x * x.x - coef.y * x.y, coef.x * x.y + coef.y * x.x);
  }
}

// partials of sum_x G_k(x) gV(x) for every basis field (slot k)
__global__ void __launch_bounds__(256) basis_kernel(const double* __restrict__ G, const double* __restrict__ gV,
                                                    int nb, long long n, double* partials, int maxwg) {
  __shared__ double red[4];
  double acc[MAXBASIS];
#pragma unroll
  for (int k = 0; k < MAXBASIS; ++k) acc[k] = 0.0;
  for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long long)gridDim.x * 256) {
    const double g = gV[i];
#pragma unroll
    for (int k = 0; k < MAXBASIS; ++k)
      if (k < nb) acc[k] += G[k * n + i] * g;
  }
#pragma unroll
  for (int k = 0; k < MAXBASIS; ++k) {
    if (k >= nb) break;
    const double r = block_sum(acc[k], red);
    if (threadIdx.x == 0) partials[k * maxwg + blockIdx.x] = r;
  }
}

__global__ void __launch_bounds__(256) basis_finalize_kernel(const double* partials, int maxwg, int nblk, int nb,
                                                             double* grad, int col) {
  __shared__ double red[4];
  for (int k = 0; k < nb; ++k) {
    double acc = 0.0;
    for (int i = threadIdx.x; i < nblk; i += 256) acc += partials[k * maxwg + i];
    const double r = block_sum(acc, red);
    if (threadIdx.x == 0) grad[k + nb * col] += r;
  }
}

}  // namespace gpe

namespace {

struct Ctl {
  int B = 0;        // blocks
  int cpb = 1;      // blocks per checkpoint
  int nck = 0;      // checkpoints
};

int check_args(gpe_handle* h, const gpe_control_args* a, Ctl* c) {
  if (!a) return fail(h, GPE_EINVAL, "null control args");
  if (h->prm.ncomp != 1) return fail(h, GPE_EINVAL, "the adjoint supports one component");
  if (h->prm.gamma != 0.0) return fail(h, GPE_EINVAL, "the adjoint needs gamma = 0 (exactly invertible steps)");
  if (h->omega != 0.0) return fail(h, GPE_EINVAL, "the adjoint does not support the rotating frame");
  if (h->adj_nbasis <= 0) return fail(h, GPE_ESTATE, "gpe_control_set has not been called");
  const int nr = h->prm.n_recalc;
  if (a->nsteps <= 0 || a->nsteps % nr || a->steps_per_interval <= 0 || a->steps_per_interval % nr)
    return fail(h, GPE_EINVAL, "nsteps and steps_per_interval must be positive multiples of n_recalc");
  if ((long long)a->nintervals * a->steps_per_interval < a->nsteps)
    return fail(h, GPE_EINVAL, "nintervals * steps_per_interval < nsteps");
  c->B = a->nsteps / nr;
  c->cpb = a->checkpoint_blocks > 0 ? a->checkpoint_blocks
                                    : std::max(1, (int)std::lround(std::sqrt((double)a->nsteps) / nr));
  c->nck = (c->B + c->cpb - 1) / c->cpb;
  return GPE_OK;
}

int ensure_buffers(gpe_handle* h, const Ctl& c, int ncoef) {
  const size_t n = (size_t)h->npts;
  if (!h->adj_lam) HIPCHK(h, hipMalloc(&h->adj_lam, n * sizeof(double2)));
  if (!h->adj_gV) HIPCHK(h, hipMalloc(&h->adj_gV, n * sizeof(double)));
  if (!h->adj_target) HIPCHK(h, hipMalloc(&h->adj_target, n * sizeof(double2)));
  if (h->adj_nckpt < c.nck) {
    (void)hipFree(h->adj_ckpt);
    h->adj_ckpt = nullptr;
    HIPCHK(h, hipMalloc(&h->adj_ckpt, (size_t)c.nck * n * sizeof(double2)));
    h->adj_nckpt = c.nck;
  }
  if (h->adj_gradcap < ncoef) {
    (void)hipFree(h->adj_grad);
    h->adj_grad = nullptr;
    HIPCHK(h, hipMalloc(&h->adj_grad, ncoef * sizeof(double)));
    h->adj_gradcap = ncoef;
  }
  const int ma = mid_axis(h);
  if (!h->kcol_c) HIPCHK(h, hipMalloc(&h->kcol_c, (size_t)h->ax[ma].ncols * sizeof(double2)));
  if (!h->kax_c) HIPCHK(h, hipMalloc(&h->kax_c, h->ax[ma].n * sizeof(double2)));
  return GPE_OK;
}

int build_v(gpe_handle* h, const double* coeffs, int m) {
  gpe::Coef c{};
  for (int k = 0; k < h->adj_nbasis; ++k) c.c[k] = coeffs[k + h->adj_nbasis * m];
  hipLaunchKernelGGL(gpe::build_v_kernel, dim3(elem_blocks(h->npts)), dim3(256), 0, h->stream, h->V[0], h->adj_V0,
                     h->adj_G, c, h->adj_nbasis, (long long)h->npts);
  HIPCHK(h, hipGetLastError());
  h->have_V[0] = true;
  ++h->v_gen;
  return GPE_OK;
}

// sum conj(a) b (diff = 0) or (sum |a-b|^2, sum |b|^2) (diff = 1), read back to the host
int pair_reduce(gpe_handle* h, const double2* a, const double2* b, int diff, double* r0, double* r1) {
  const int nb = elem_blocks(h->npts);
  hipLaunchKernelGGL(gpe::pair_reduce_kernel, dim3(nb), dim3(256), 0, h->stream, a, b, (long long)h->npts, diff,
                     h->partials, h->maxwg);
  HIPCHK(h, hipGetLastError());
  FinParams f = fin_params(h, gpe::FIN_SUMS);
  f.cnt[gpe::SL_A] = f.cnt[gpe::SL_B] = nb;
  HIPCHK(h, gpe::launch_finalize(f, h->stream));
  DevState s;
  int rc;
  if ((rc = read_state(h, &s))) return rc;
  *r0 = s.sA;
  *r1 = s.sB;
  return GPE_OK;
}

// mu of solve_split.m:41-46 for the initial state under the interval-0 potential
int control_begin(gpe_handle* h, const gpe_control_args* a, const double* coeffs) {
  int rc;
  h->ddt = a->ddt;
  h->dr = 0.0;
  h->di = a->ddt;
  h->n_rec = h->prm.n_recalc;
  if ((rc = build_v(h, coeffs, 0))) return rc;
  double m[5];
  if ((rc = gpe_measure(h, m))) return rc;
  const double NN0 = h->prm.Ntotal > 0 ? h->prm.Ntotal : m[0];
  const double muc = ((m[1] + m[3]) + h->prm.g[0] * m[2]) / NN0;
  DevState s;
  if ((rc = read_state(h, &s))) return rc;
  s.mu = h->prm.mu_init > 0 ? h->prm.mu_init : muc;
  s.scale = 1.0;
  if ((rc = write_state(h, s))) return rc;
  if ((rc = build_kinetic(h, {0.0, a->ddt}, h->kcol, h->kax))) return rc;
  if ((rc = build_kinetic(h, {0.0, -a->ddt}, h->kcol_c, h->kax_c))) return rc;
  return GPE_OK;
}

// Forward block sweeps without the block-end reductions (mu is fixed for gamma = 0).
int forward_block(gpe_handle* h) {
  int rc;
  double2* psi = h->psi[0];
  const double* V = h->V[0];
  for (int i = 0; i < h->n_rec; ++i) {
    if ((rc = sweep(h, 0, i == 0 ? gpe::S_OPEN : gpe::S_MID, base_params(h, 0, psi, V)))) return rc;
    if (h->ndims == 3) {
      if ((rc = sweep(h, 1, gpe::S_FWD, base_params(h, 1, psi, V)))) return rc;
      if ((rc = sweep(h, 2, gpe::S_KIN, base_params(h, 2, psi, V)))) return rc;
      if ((rc = sweep(h, 1, gpe::S_INV, base_params(h, 1, psi, V)))) return rc;
    } else {
      if ((rc = sweep(h, 1, gpe::S_KIN, base_params(h, 1, psi, V)))) return rc;
    }
  }
  return sweep(h, 0, gpe::S_CLOSE, base_params(h, 0, psi, V));
}

int run_forward(gpe_handle* h, const gpe_control_args* a, const double* coeffs, const Ctl& c, bool keep) {
  int rc;
  int cur = -1;
  for (int b = 0; b < c.B; ++b) {
    if (keep && b % c.cpb == 0)
      HIPCHK(h, hipMemcpyAsync(h->adj_ckpt + (size_t)(b / c.cpb) * h->npts, h->psi[0], h->npts * sizeof(double2),
                               hipMemcpyDeviceToDevice, h->stream));
    const int m = (b * h->n_rec) / a->steps_per_interval;
    if (m != cur) {
      if ((rc = build_v(h, coeffs, m))) return rc;
      cur = m;
    }
    if ((rc = forward_block(h))) return rc;
  }
  return GPE_OK;
}

// inverse kinetic step (conjugate table) on one array: y/z sweeps of a step
int inverse_kinetic(gpe_handle* h, double2* arr) {
  int rc;
  if (h->ndims == 3) {
    if ((rc = sweep(h, 1, gpe::S_FWD, base_params(h, 1, arr, nullptr)))) return rc;
    SweepParams k = base_params(h, 2, arr, nullptr);
    k.kcol = h->kcol_c;
    k.kax = h->kax_c;
    if ((rc = sweep(h, 2, gpe::S_KIN, k))) return rc;
    return sweep(h, 1, gpe::S_INV, base_params(h, 1, arr, nullptr));
  }
  SweepParams k = base_params(h, 1, arr, nullptr);
  k.kcol = h->kcol_c;
  k.kax = h->kax_c;
  return sweep(h, 1, gpe::S_KIN, k);
}

int adj_x(gpe_handle* h, int seq) {
  gpe::AdjParams p{};
  const Axis& a = h->ax[0];
  p.psi = h->psi[0];
  p.lam = h->adj_lam;
  p.V = h->V[0];
  p.gV = h->adj_gV;
  p.tw = a.tw;
  p.st = h->st;
  p.ncols = (unsigned)a.ncols;
  p.A = (unsigned)a.A;
  p.sa = (unsigned)a.sa;
  p.sb = (unsigned)a.sb;
  p.stride = (unsigned)a.stride;
  p.g = h->prm.g[0];
  p.h = h->ddt;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (h->prof) {
    e0 = prof_event(h);
    e1 = prof_event(h);
    if (e0) hipEventRecord(e0, h->stream);
  }
  // the x-axis plan of the forward sweeps (8 elements per thread for the 8.8.8 plan of 512)
  hipError_t e = gpe::launch_adj_x(a.n, seq, h->var[0], p, h->stream, nullptr);
  if (h->prof && e0 && e1) {
    hipEventRecord(e1, h->stream);
    h->prof_recs.push_back({(seq == gpe::A_MID ? gpe::S_MID : seq == gpe::A_OPEN ? gpe::S_OPEN : gpe::S_CLOSE) * 3,
                            {e0, e1}});
  }
  if (e != hipSuccess) return fail(h, GPE_EHIP, "adjoint sweep: %s", hipGetErrorString(e));
  return GPE_OK;
}

int reduce_basis(gpe_handle* h, int m) {
  const int nb = elem_blocks(h->npts);
  hipLaunchKernelGGL(gpe::basis_kernel, dim3(nb), dim3(256), 0, h->stream, h->adj_G, h->adj_gV, h->adj_nbasis,
                     (long long)h->npts, h->partials, h->maxwg);
  hipLaunchKernelGGL(gpe::basis_finalize_kernel, dim3(1), dim3(256), 0, h->stream, h->partials, h->maxwg, nb,
                     h->adj_nbasis, h->adj_grad, m);
  HIPCHK(h, hipGetLastError());
  HIPCHK(h, hipMemsetAsync(h->adj_gV, 0, h->npts * sizeof(double), h->stream));
  return GPE_OK;
}

}  // namespace

extern "C" {

int gpe_control_set(gpe_handle* h, int nbasis, const double* V0, const double* G) {
  if (!h || !V0 || !G || nbasis < 1 || nbasis > gpe::MAXBASIS)
    return h ? fail(h, GPE_EINVAL, "need 1..%d basis fields", gpe::MAXBASIS) : GPE_EINVAL;
  HIPCHK(h, hipSetDevice(h->device));
  const size_t n = (size_t)h->npts;
  (void)hipFree(h->adj_V0);
  (void)hipFree(h->adj_G);
  h->adj_V0 = nullptr;
  h->adj_G = nullptr;
  HIPCHK(h, hipMalloc(&h->adj_V0, n * sizeof(double)));
  HIPCHK(h, hipMalloc(&h->adj_G, (size_t)nbasis * n * sizeof(double)));
  HIPCHK(h, hipMemcpy(h->adj_V0, V0, n * sizeof(double), hipMemcpyHostToDevice));
  HIPCHK(h, hipMemcpy(h->adj_G, G, (size_t)nbasis * n * sizeof(double), hipMemcpyHostToDevice));
  h->adj_nbasis = nbasis;
  return GPE_OK;
}

int gpe_control_forward(gpe_handle* h, const gpe_control_args* a, const double* coeffs, double* psi_out) {
  int rc;
  Ctl c;
  if (!h || !coeffs) return GPE_EINVAL;
  if (!h->have_params) return fail(h, GPE_ESTATE, "gpe_set_params has not been called");
  HIPCHK(h, hipSetDevice(h->device));
  if ((rc = check_args(h, a, &c))) return rc;
  if ((rc = ensure_buffers(h, c, h->adj_nbasis * a->nintervals))) return rc;
  if ((rc = materialize(h))) return rc;
  HIPCHK(h, hipMemcpyAsync(h->adj_ckpt, h->psi[0], h->npts * sizeof(double2), hipMemcpyDeviceToDevice, h->stream));
  if ((rc = control_begin(h, a, coeffs))) return rc;
  if ((rc = run_forward(h, a, coeffs, c, false))) return rc;
  if (psi_out)
    HIPCHK(h, hipMemcpyAsync(psi_out, h->psi[0], h->npts * sizeof(double2), hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipMemcpyAsync(h->psi[0], h->adj_ckpt, h->npts * sizeof(double2), hipMemcpyDeviceToDevice, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return GPE_OK;
}

int gpe_adjoint_grad(gpe_handle* h, const gpe_control_args* a, const double* coeffs, const double* psi_target,
                     double* loss, double* grad, double* recon_err) {
  int rc;
  Ctl c;
  if (!h || !coeffs || !psi_target || !grad) return GPE_EINVAL;
  if (!h->have_params) return fail(h, GPE_ESTATE, "gpe_set_params has not been called");
  HIPCHK(h, hipSetDevice(h->device));
  if ((rc = check_args(h, a, &c))) return rc;
  const int ncoef = h->adj_nbasis * a->nintervals;
  if ((rc = ensure_buffers(h, c, ncoef))) return rc;
  if ((rc = materialize(h))) return rc;
  HIPCHK(h, hipMemcpyAsync(h->adj_target, psi_target, h->npts * sizeof(double2), hipMemcpyHostToDevice, h->stream));
  if ((rc = control_begin(h, a, coeffs))) return rc;
  if ((rc = run_forward(h, a, coeffs, c, true))) return rc;
  // loss = 1 - |o|^2, o = <psi_target|psi_T>; lambda_T = -2 o dV psi_target
  double ore, oim;
  if ((rc = pair_reduce(h, h->adj_target, h->psi[0], 0, &ore, &oim))) return rc;
  ore *= h->w;
  oim *= h->w;
  if (loss) *loss = 1.0 - (ore * ore + oim * oim);
  const double2 coef = make_double2(-2.0 * ore * h->w, -2.0 * oim * h->w);
  hipLaunchKernelGGL(gpe::lam_init_kernel, dim3(elem_blocks(h->npts)), dim3(256), 0, h->stream, h->adj_lam,
                     h->adj_target, coef, (long long)h->npts);
  HIPCHK(h, hipGetLastError());
  HIPCHK(h, hipMemsetAsync(h->adj_gV, 0, h->npts * sizeof(double), h->stream));
  HIPCHK(h, hipMemsetAsync(h->adj_grad, 0, ncoef * sizeof(double), h->stream));
  int cur = -1;
  double rerr = 0.0;
  for (int b = c.B - 1; b >= 0; --b) {
    const int m = (b * h->n_rec) / a->steps_per_interval;
    if (m != cur) {
      if (cur >= 0 && (rc = reduce_basis(h, cur))) return rc;
      if ((rc = build_v(h, coeffs, m))) return rc;
      cur = m;
    }
    if ((rc = adj_x(h, gpe::A_CLOSE))) return rc;
    for (int i = h->n_rec - 1; i >= 0; --i) {
      if ((rc = inverse_kinetic(h, h->psi[0]))) return rc;
      if ((rc = inverse_kinetic(h, h->adj_lam))) return rc;
      if ((rc = adj_x(h, i > 0 ? gpe::A_MID : gpe::A_OPEN))) return rc;
    }
    if (b % c.cpb == 0) {
      const double2* ck = h->adj_ckpt + (size_t)(b / c.cpb) * h->npts;
      if (b == 0 && recon_err) {
        double d2, n2;
        if ((rc = pair_reduce(h, h->psi[0], ck, 1, &d2, &n2))) return rc;
        rerr = std::sqrt(d2 / n2);
      }
      HIPCHK(h, hipMemcpyAsync(h->psi[0], ck, h->npts * sizeof(double2), hipMemcpyDeviceToDevice, h->stream));
    }
  }
  if ((rc = reduce_basis(h, cur))) return rc;
  HIPCHK(h, hipMemcpyAsync(grad, h->adj_grad, ncoef * sizeof(double), hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  if (recon_err) *recon_err = rerr;
  return GPE_OK;
}

}  // extern "C"
