// Device-evaluated time-dependent potential: V_j = Vtrap_j + (c_0 G_0 + c_1 G_1 + ...),
// the value of getVtotal(time) = Vtrap + Vtd(task, time) (GPEtask.m:62-71,
// GPEtaskSpinor.m:24-44) for a Vtd that is a linear combination of fixed fields
// with time-dependent coefficients.  solve_split evaluates the potential once per
// block (solve_split.m:53-54); with the basis resident in HBM a block costs one
// small coefficient upload and one pass over the basis instead of a host
// evaluation and a full-grid host-to-device copy.
#include "handle.h"

using namespace gpe_detail;

namespace gpe {

constexpr int VTD_MAXBASIS = 16;
struct VtdCoef {
  double c[VTD_MAXBASIS];
};

// Vtd is summed left to right first and then added to Vtrap, as Vtrap + Vtd(...)
__global__ void __launch_bounds__(256) vtd_build_kernel(double* __restrict__ V, const double* __restrict__ V0,
                                                        const double* __restrict__ G, VtdCoef c, int nb,
                                                        long long n) {
  for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long long)gridDim.x * 256) {
    double acc = c.c[0] * G[i];
    for (int k = 1; k < nb; ++k) acc = acc + c.c[k] * G[k * n + i];
    V[i] = V0[i] + acc;
  }
}

}  // namespace gpe

extern "C" {

int gpe_vtd_set(gpe_handle* h, int nbasis, const double* G, const double* Vtrap) {
  if (!h || !G || !Vtrap) return h ? fail(h, GPE_EINVAL, "null basis or Vtrap") : GPE_EINVAL;
  if (nbasis < 1 || nbasis > gpe::VTD_MAXBASIS) return fail(h, GPE_EINVAL, "need 1..%d basis fields", gpe::VTD_MAXBASIS);
  if (!h->have_params) return fail(h, GPE_ESTATE, "gpe_set_params has not been called");
  HIPCHK(h, hipSetDevice(h->device));
  const size_t n = (size_t)h->npts;
  const int nc = h->prm.ncomp;
  (void)hipFree(h->vtd_G);
  (void)hipFree(h->vtd_V0);
  h->vtd_G = nullptr;
  h->vtd_V0 = nullptr;
  h->vtd_nb = 0;
  HIPCHK(h, hipMalloc(&h->vtd_G, (size_t)nbasis * n * sizeof(double)));
  HIPCHK(h, hipMalloc(&h->vtd_V0, (size_t)nc * n * sizeof(double)));
  HIPCHK(h, hipMemcpy(h->vtd_G, G, (size_t)nbasis * n * sizeof(double), hipMemcpyHostToDevice));
  HIPCHK(h, hipMemcpy(h->vtd_V0, Vtrap, (size_t)nc * n * sizeof(double), hipMemcpyHostToDevice));
  h->vtd_nb = nbasis;
  h->vtd_nc = nc;
  return GPE_OK;
}

int gpe_vtd_coeffs(gpe_handle* h, const double* c, int target) {
  if (!h || !c) return h ? fail(h, GPE_EINVAL, "null coefficients") : GPE_EINVAL;
  if (h->vtd_nb < 1) return fail(h, GPE_ESTATE, "gpe_vtd_set has not been called");
  if (h->vtd_nc != h->prm.ncomp) return fail(h, GPE_ESTATE, "gpe_vtd_set was called for another component count");
  HIPCHK(h, hipSetDevice(h->device));
  gpe::VtdCoef k{};
  for (int i = 0; i < h->vtd_nb; ++i) {
    if (!std::isfinite(c[i])) return fail(h, GPE_EINVAL, "non-finite coefficient %d", i);
    k.c[i] = c[i];
  }
  if (target != 0 && target != 1) return fail(h, GPE_EINVAL, "target must be 0 (step) or 1 (reductions)");
  for (int comp = 0; comp < h->vtd_nc; ++comp) {
    double** dst = target ? &h->Vmu[comp] : &h->V[comp];
    if (!*dst) HIPCHK(h, hipMalloc(dst, (size_t)h->npts * sizeof(double)));
    hipLaunchKernelGGL(gpe::vtd_build_kernel, dim3(elem_blocks(h->npts)), dim3(256), 0, h->stream, *dst,
                       h->vtd_V0 + (size_t)comp * h->npts, h->vtd_G, k, h->vtd_nb, (long long)h->npts);
    HIPCHK(h, hipGetLastError());
    if (target) h->use_vmu = true;
    else h->have_V[comp] = true;
    if (!target && comp == 0) ++h->v_gen;
  }
  return GPE_OK;
}

}  // extern "C"
