// Array-valued Hamiltonian: applyham / applyh0 of GPEtask (GPEtask.m:73-95) and
// applyham / applyham1c / applyh0 of GPEtaskSpinor (GPEtaskSpinor.m:45-86).
//
// The kinetic term lap(phi) = ifftn(kk .* fftn(phi)) with kk = (kx^2 + ky^2 + kz^2)/2
// (grid3d/lap.m:5-6) is separable: it equals the sum over the axes of
// ifft_a(k_a^2/2 .* fft_a(phi)).  Each axis term is one S_KIN sweep (forward,
// multiply, inverse along that axis) of a copy of phi, added to the output.  With a
// scalar omega the rotating-frame operator of grid2d/lap.m:8-9 and grid3d/lap.m:8-10
// adds -omega y ifftx(kx fftx(phi)) + omega x iffty(ky ffty(phi)): two more first-
// derivative sweeps, weighted by the coordinate when they are added.
//
// The rest is pointwise: V_j phi_j, the interaction (g |phi|^2 phi, or
// sum_k g_jk |phi_k|^2 phi_j) and the Rabi coupling (Omega phi_2 for component 1,
// conj(Omega) phi_1 for component 2).  The call is not on the time-stepping path; it
// allocates its work buffers for the call and frees them before it returns.
#include "handle.h"

using namespace gpe_detail;

namespace {

struct PointParams {
  const double2* phi[2];
  const double* V;
  double2* out;
  long long n;
  int j;                // component written
  int ncomp;
  int h0;               // applyh0: potential term only
  double gj[2];         // g_j1, g_j2 (one component: g, 0)
  const double2* cpl;   // coupling array, or null for the scalar cpl_s
  double2 cpl_s;
};

// out = V phi_j [+ (sum_k g_jk |phi_k|^2) phi_j] [+ coupling term]
__global__ void __launch_bounds__(256) ham_point_kernel(const PointParams p) {
  for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < p.n; i += (long long)gridDim.x * 256) {
    const double2 f = p.phi[p.j][i];
    const double v = p.V[i];
    double2 r = make_double2(v * f.x, v * f.y);
    if (!p.h0) {
      double a = 0.0;
      for (int k = 0; k < p.ncomp; ++k) {
        const double2 q = p.phi[k][i];
        a += p.gj[k] * (q.x * q.x + q.y * q.y);
      }
      r.x += a * f.x;
      r.y += a * f.y;
      if (p.ncomp == 2) {
        const double2 om = p.cpl ? p.cpl[i] : p.cpl_s;
        const double2 o = p.phi[1 - p.j][i];
        // component 1: Omega phi_2; component 2: conj(Omega) phi_1 (GPEtaskSpinor.m:66-70)
        const double oy = p.j == 0 ? om.y : -om.y;
        r.x += om.x * o.x - oy * o.y;
        r.y += om.x * o.y + oy * o.x;
      }
    }
    p.out[i] = r;
  }
}

// out += w t, with w = 1 (mode 0), c * y[i mod ny] (mode 1) or c * x[(i / ny) mod nx]
// (mode 2) on the column-major [ny, nx, nz] grid.
__global__ void __launch_bounds__(256) ham_acc_kernel(double2* __restrict__ out, const double2* __restrict__ t,
                                                      long long n, int mode, const double* __restrict__ coord,
                                                      double c, unsigned ny, unsigned nx) {
  for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long long)gridDim.x * 256) {
    const double2 x = t[i];
    double2 o = out[i];
    if (mode == 0) {
      o.x += x.x;
      o.y += x.y;
    } else {
      const unsigned r = mode == 1 ? (unsigned)(i % ny) : (unsigned)((i / ny) % nx);
      const double w = c * coord[r];
      o.x += w * x.x;
      o.y += w * x.y;
    }
    out[i] = o;
  }
}

__global__ void __launch_bounds__(256) ham_scale_kernel(double2* __restrict__ x, long long n, double s) {
  for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long long)gridDim.x * 256) {
    const double2 v = x[i];
    x[i] = make_double2(v.x * s, v.y * s);
  }
}

// S_KIN tables of the kinetic term: kcol = 1 for every pencil, then per axis k^2/2/n_a
// (off2) and, for the rotating frame, k/n_a of x and y (off1).
std::vector<double2> kinetic_tables(const gpe_handle* h, bool rot, long long off2[3], long long off1[3]) {
  long long maxcols = 0;
  for (int a = 0; a < h->ndims; ++a) maxcols = std::max(maxcols, h->ax[a].ncols);
  std::vector<double2> tab((size_t)maxcols, make_double2(1.0, 0.0));
  for (int a = 0; a < 3; ++a) off2[a] = off1[a] = 0;
  for (int a = 0; a < h->ndims; ++a) {
    const Axis& A = h->ax[a];
    off2[a] = (long long)tab.size();
    for (int i = 0; i < A.n; ++i) tab.push_back(make_double2(A.k[i] * A.k[i] / 2 / A.n, 0.0));
    off1[a] = (long long)tab.size();
    if (rot && a < 2)
      for (int i = 0; i < A.n; ++i) tab.push_back(make_double2(A.k[i] / A.n, 0.0));
  }
  return tab;
}

// out += lap(phi, omega) (grid2d/lap.m:5-9, grid3d/lap.m:5-10) on the device: one S_KIN
// sweep per axis of a copy of phi in tmp (and per first derivative when omega != 0).
int add_kinetic(gpe_handle* h, const double2* phi, double2* out, double2* tmp, const double2* dtab,
                const long long off2[3], const long long off1[3], double omega) {
  int rc;
  const long long n = h->npts;
  const int nb = elem_blocks(n);
  const bool rot = omega != 0.0;
  for (int a = 0; a < h->ndims; ++a) {
    const int passes = rot && a < 2 ? 2 : 1;
    for (int q = 0; q < passes; ++q) {
      HIPCHK(h, hipMemcpyAsync(tmp, phi, (size_t)n * sizeof(double2), hipMemcpyDeviceToDevice, h->stream));
      SweepParams sp = base_params(h, a, tmp, nullptr);
      sp.kcol = dtab;
      sp.kax = dtab + (q == 0 ? off2[a] : off1[a]);
      sp.trow = 0;
      sp.rtab = nullptr;
      if ((rc = sweep(h, a, gpe::S_KIN, sp))) return rc;
      // grid3d/lap.m:8-10: -omega y on the x derivative, +omega x on the y derivative
      const int mode = q == 0 ? 0 : (a == 0 ? 1 : 2);
      const double cw = a == 0 ? -omega : omega;
      const double* coord = a == 0 ? h->ax[1].xc : h->ax[0].xc;
      hipLaunchKernelGGL(ham_acc_kernel, dim3(nb), dim3(256), 0, h->stream, out, (const double2*)tmp, n, mode,
                         coord, cw, (unsigned)h->ny, (unsigned)h->nx);
      HIPCHK(h, hipGetLastError());
    }
  }
  return GPE_OK;
}

}  // namespace

extern "C" int gpe_applyham(gpe_handle* h, const double* phi, double* out, int h0_only) {
  int rc;
  if ((rc = ensure_ready(h))) return rc;
  if (!out) return fail(h, GPE_EINVAL, "gpe_applyham: null output");
  const int nc = h->prm.ncomp;
  if (nc == 2 && h->omega != 0.0) return fail(h, GPE_EINVAL, "the rotating frame is supported for one component");
  for (int c = 0; c < nc; ++c)
    if (!phi && !h->psi[c]) return fail(h, GPE_ESTATE, "no state of component %d on the device", c);
  const long long n = h->npts;
  const int nd = h->ndims;
  const bool rot = h->omega != 0.0;
  long long off2[3], off1[3];
  const std::vector<double2> tab = kinetic_tables(h, rot, off2, off1);
  // work buffers: phi, the output and one sweep copy
  const size_t cb = (size_t)n * sizeof(double2);
  const size_t nphi = (size_t)nc;
  char* buf = nullptr;
  const size_t bytes = (nphi + (size_t)nc + 1) * cb + tab.size() * sizeof(double2);
  HIPCHK(h, hipMalloc(&buf, bytes));
  auto run = [&]() -> int {
    const int nb = elem_blocks(n);
    double2* ph[2] = {nullptr, nullptr};
    for (int c = 0; c < nc; ++c) {
      ph[c] = (double2*)(buf + c * cb);
      if (phi) {
        HIPCHK(h, hipMemcpyAsync(ph[c], phi + 2 * (size_t)n * c, cb, hipMemcpyHostToDevice, h->stream));
      } else {
        // the stored state times its pending renormalisation (solve_split.m:72), which
        // stays pending for the run
        HIPCHK(h, hipMemcpyAsync(ph[c], h->psi[c], cb, hipMemcpyDeviceToDevice, h->stream));
        HIPCHK(h, gpe::launch_scale(ph[c], h->st, 0, n, nb, h->stream, false));
      }
    }
    double2* ho = (double2*)(buf + nphi * cb);
    double2* tmp = (double2*)(buf + (nphi + nc) * cb);
    double2* dtab = (double2*)(buf + (nphi + nc + 1) * cb);
    HIPCHK(h, hipMemcpyAsync(dtab, tab.data(), tab.size() * sizeof(double2), hipMemcpyHostToDevice, h->stream));
    for (int j = 0; j < nc; ++j) {
      double2* oj = ho + (size_t)n * j;
      PointParams pp{};
      pp.phi[0] = ph[0];
      pp.phi[1] = ph[1];
      pp.V = h->V[j];
      pp.out = oj;
      pp.n = n;
      pp.j = j;
      pp.ncomp = nc;
      pp.h0 = h0_only ? 1 : 0;
      pp.gj[0] = nc == 2 ? h->prm.g[j] : h->prm.g[0];
      pp.gj[1] = nc == 2 ? h->prm.g[j + 2] : 0.0;
      pp.cpl = h->cpl_array ? h->cpl : nullptr;
      pp.cpl_s = make_double2(h->cpl_re, h->cpl_im);
      hipLaunchKernelGGL(ham_point_kernel, dim3(nb), dim3(256), 0, h->stream, pp);
      HIPCHK(h, hipGetLastError());
      // kinetic terms, one sweep per axis (and per derivative when rotating)
      if ((rc = add_kinetic(h, ph[j], oj, tmp, dtab, off2, off1, h->omega))) return rc;
    }
    for (int j = 0; j < nc; ++j)
      HIPCHK(h, hipMemcpyAsync(out + 2 * (size_t)n * j, ho + (size_t)n * j, cb, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    return GPE_OK;
  };
  rc = run();
  (void)hipStreamSynchronize(h->stream);
  (void)hipFree(buf);
  return rc;
}

extern "C" int gpe_grid_transform(gpe_handle* h, int op, double omega, const double* in, double* out) {
  if (!h) return GPE_EINVAL;
  if (!in || !out) return fail(h, GPE_EINVAL, "gpe_grid_transform: null array");
  if (op != GPE_GRID_FFT && op != GPE_GRID_IFFT && op != GPE_GRID_LAP)
    return fail(h, GPE_EINVAL, "gpe_grid_transform: unknown op %d", op);
  if (!std::isfinite(omega)) return fail(h, GPE_EINVAL, "gpe_grid_transform: omega is not finite");
  HIPCHK(h, hipSetDevice(h->device));
  int rc;
  const long long n = h->npts;
  const bool lap = op == GPE_GRID_LAP;
  long long off2[3], off1[3];
  const std::vector<double2> tab =
      lap ? kinetic_tables(h, omega != 0.0, off2, off1) : std::vector<double2>();
  // work buffers: the array (transformed in place, or phi of lap), lap's sum and sweep copy
  const size_t cb = (size_t)n * sizeof(double2);
  char* buf = nullptr;
  const size_t bytes = (lap ? 3 : 1) * cb + tab.size() * sizeof(double2);
  HIPCHK(h, hipMalloc(&buf, bytes));
  auto run = [&]() -> int {
    double2* x = (double2*)buf;
    HIPCHK(h, hipMemcpyAsync(x, in, cb, hipMemcpyHostToDevice, h->stream));
    const double2* res = x;
    if (lap) {
      double2* acc = (double2*)(buf + cb);
      double2* tmp = (double2*)(buf + 2 * cb);
      double2* dtab = (double2*)(buf + 3 * cb);
      HIPCHK(h, hipMemcpyAsync(dtab, tab.data(), tab.size() * sizeof(double2), hipMemcpyHostToDevice, h->stream));
      HIPCHK(h, hipMemsetAsync(acc, 0, cb, h->stream));
      if ((rc = add_kinetic(h, x, acc, tmp, dtab, off2, off1, omega))) return rc;
      res = acc;
    } else {
      // fftn / ifftn (grid2d/fft.m:11, ifft.m:11; grid3d/fft.m:11, ifft.m:11): one
      // transform sweep per axis; ifftn divides by the number of points
      for (int a = 0; a < h->ndims; ++a) {
        SweepParams sp = base_params(h, a, x, nullptr);
        sp.apply_scale = 0;
        if ((rc = sweep(h, a, op == GPE_GRID_FFT ? gpe::S_FWD : gpe::S_INV, sp))) return rc;
      }
      if (op == GPE_GRID_IFFT) {
        hipLaunchKernelGGL(ham_scale_kernel, dim3(elem_blocks(n)), dim3(256), 0, h->stream, x, n, 1.0 / (double)n);
        HIPCHK(h, hipGetLastError());
      }
    }
    HIPCHK(h, hipMemcpyAsync(out, res, cb, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    return GPE_OK;
  };
  rc = run();
  (void)hipStreamSynchronize(h->stream);
  (void)hipFree(buf);
  return rc;
}
