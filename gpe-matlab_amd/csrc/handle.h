// Private definition of the library handle and the helpers shared by the
// drivers (scalar GPEtask, spinor, adjoint, batch).
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>
#include <complex>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/gpe_hip.h"
#include "state.h"
#include "sweep.h"

using gpe::DevState;
using gpe::FinParams;
using gpe::SweepParams;

namespace gpe_detail {

struct Axis {
  int n = 1;
  long long stride = 1, ncols = 0, A = 1, sa = 0, sb = 0;
  double2* tw = nullptr;      // forward twiddles
  double* keax = nullptr;     // k^2/2/n
  std::vector<double> k;      // wave vectors (grid*/init.m k layout)
  std::vector<double> x;      // coordinates
  double* k1ax = nullptr;     // device k/n (rotating-frame kinetic sums)
  double* xc = nullptr;       // device coordinates
};

}  // namespace gpe_detail

struct gpe_handle {
  int device = 0;
  int ndims = 2;
  int nx = 1, ny = 1, nz = 1;
  long long npts = 0;
  double dx = 0, dy = 0, dz = 0, w = 0;
  gpe_detail::Axis ax[3];  // 0: x, 1: y, 2: z
  hipStream_t stream = nullptr;
  double2* psi[2] = {nullptr, nullptr};
  double* V[2] = {nullptr, nullptr};
  double2* cpl = nullptr;       // spinor coupling array (or null for scalar)
  void* stage = nullptr;        // C-order transfers: npts complex values (allocated on first use)
  double cpl_re = 0, cpl_im = 0;
  bool cpl_array = false;
  double cpl_max2 = 0;          // max |Omega|^2 of the coupling array (NaN propagates)
  double omega = 0;             // rotation frequency (GPEtask.m:7), scalar
  double2* rtab_x = nullptr;    // rotating frame: x half-step table [iy][kx]
  double2* rtab_y = nullptr;    // rotating frame, 2D: y step table [ix][ky]
  double2* kcol = nullptr;      // kinetic factor per mid-axis column
  double2* kax = nullptr;       // kinetic factor per mid-axis index (incl. 1/npts)
  double2* kcol_sp[2] = {nullptr, nullptr};  // spinor: exp(-kk dt/2) and exp(-kk dt) exp(kk dt/2)
  double2* kax_sp[2] = {nullptr, nullptr};
  // spinor block ends inside an outer iteration (3D): kax_sp[0] times nx (a block
  // opening from the stored x spectrum F_x psi / nx) and the y and z k^2/2/n tables
  // times nx (kinetic sums of that spectrum)
  double2* kax_sp2 = nullptr;
  double* keax_nx[3] = {nullptr, nullptr, nullptr};
  // 3D spinor block ends: kinetic-sum weights of the last kinetic sweep (S_KINE):
  // (kx^2 + ky^2)/2 npts per z pencil (kcol's order) and kz^2/2 npts per kz index
  double* kecolS = nullptr;
  double* kezS = nullptr;
  double cosom[2] = {1, 0}, sinomm[2] = {0, 0}, sinomp[2] = {0, 0};  // scalar Rabi coefficients
  double* nop = nullptr;        // ITP stored operator
  double2* psi_old = nullptr;   // ITP convergence reference
  double* partials = nullptr;
  int maxwg = 0;
  DevState* st = nullptr;
  double* MUdev = nullptr;
  int MUcap = 0;
  gpe_params prm{};
  bool have_params = false;
  bool have_V[2] = {false, false};
  // solve_split state
  bool begun = false;
  double ddt = 0, dr = 0, di = 0;
  int niter_inner = 0, n_rec = 1;
  double NN0 = 0;
  std::string err;
  // adjoint / control (gpe_control_set, gpe_adjoint_grad)
  int adj_nbasis = 0;
  double* adj_V0 = nullptr;
  double* adj_G = nullptr;        // nbasis x npts basis fields
  double2* adj_lam = nullptr;
  double* adj_gV = nullptr;
  double2* adj_ckpt = nullptr;
  int adj_nckpt = 0;
  double2* adj_target = nullptr;
  double* adj_grad = nullptr;
  int adj_gradcap = 0;
  double2* kcol_c = nullptr;      // conjugate kinetic tables (backward steps)
  double2* kax_c = nullptr;
  // stochastic GPE (gpe_sgpe_*)
  bool sg_begun = false, sg_mask = false;
  double sg_var = 0;
  unsigned long long sg_seed = 0;
  int* sg_kmax = nullptr;         // per mid-axis column: largest |k| index with kk < ecut
  int* sg_kmaxB = nullptr;        // the same in the y-split z sweep's column order
  // two-component block-end reductions: potential at task.current_time (gpe_set_potential_mu)
  double* Vmu[2] = {nullptr, nullptr};
  bool use_vmu = false;
  // device-evaluated time-dependent potential (gpe_vtd_set / gpe_vtd_coeffs)
  double* vtd_G = nullptr;        // nbasis x npts basis fields
  double* vtd_V0 = nullptr;       // ncomp x npts trap potentials
  int vtd_nb = 0, vtd_nc = 0;
  // two-component ITP (gpe_itp2)
  double* itp2_P = nullptr;       // stored phase factors, 2 x npts
  double* itp2_hist[3] = {nullptr, nullptr, nullptr};   // MU, MU2, EE device histories
  int itp2_cap = 0;
  // spinor Rabi adjoint (gpe_rabi_forward, gpe_rabi_grad)
  double2* rb_lam[2] = {nullptr, nullptr};
  double2* rb_ckpt = nullptr;     // checkpoints, 2 components each
  long long rb_nckpt = 0;
  double2* rb_tape = nullptr;     // pre-step states of one segment, 2 components per step
  long long rb_ntape = 0;
  double* rb_mu = nullptr;        // mu at each block start
  long long rb_nmu = 0;
  double* rb_xc = nullptr;        // x coordinates
  double* rb_gacc = nullptr;      // 2 x maxwg gradient accumulators
  double* rb_out = nullptr;       // 2 reduced gradient components
  double2* kcol_rc[3] = {nullptr, nullptr, nullptr};   // conj(exp(-kk dt/2)), conj(exp(-kk dt)),
  double2* kax_rc[3] = {nullptr, nullptr, nullptr};    // conj(exp(-kk dt) exp(kk dt/2))
  // vortex-core detection (gpe_detect_core*): one grown scratch allocation
  char* dc_buf = nullptr;
  size_t dc_bytes = 0;
  // FFT plan variant of the tunable sweeps along each axis, fixed when the handle is
  // created (from the plan wisdom) or tuned (gpe_tune)
  int var[3] = {0, 0, 0};
  // y-split 3D step (ny = 256): potential in the x sweeps' tile order, kinetic factor
  // per z pencil in the split's spectral order; v_gen counts writes of V[0]
  bool ysplit_off = false;      // GPE_YSPLIT=0 turns the two-sweep step off
  double* Vp = nullptr;
  double2* psiy = nullptr;       // y-split step: the state in block order inside an outer iteration
  double2* kcolB = nullptr;
  double* keB = nullptr;         // y-split block ends: kinetic-sum weights in the z sweep's order
  double2* twy = nullptr;       // 256-point twiddles of the y axis
  long long v_gen = 0, vp_gen = -1;
  // one outer iteration of gpe_split_outer captured as a graph (static potential,
  // decay_rate = 0); dropped by gpe_split_begin
  hipGraph_t og = nullptr;
  hipGraphExec_t ogx = nullptr;
  bool og_off = false;   // capture failed or disabled (GPE_NO_GRAPH)
  double og_key[8] = {};  // host values baked into the captured launches
  // per-sweep HIP-event profiling (gpe_profile_enable / gpe_profile_read)
  bool prof = false;
  std::vector<hipEvent_t> ev_pool;
  std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> prof_recs;
};

namespace gpe_detail {

inline int fail(gpe_handle* h, int code, const char* fmt, ...) {
  if (h) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    h->err = buf;
  }
  return code;
}

#define HIPCHK(h, x)                                                                         \
  do {                                                                                       \
    hipError_t e_ = (x);                                                                     \
    if (e_ != hipSuccess) return fail((h), GPE_EHIP, "%s: %s", #x, hipGetErrorString(e_));   \
  } while (0)

// numpy/MATLAB linspace: x_i = start + i*step, last point set to stop.
inline std::vector<double> linspace(double a, double b, int n) {
  std::vector<double> x(n);
  const double step = (b - a) / (n - 1);
  for (int i = 0; i < n; ++i) x[i] = (double)i * step + a;
  x[n - 1] = b;
  return x;
}

// grid2d/init.m:27: k = [0:n/2, -(n/2-1):-1] * 2 pi / (x(end) - x(1) + dx)
inline std::vector<double> kvec(const std::vector<double>& x, double dx) {
  const int n = (int)x.size();
  std::vector<double> k(n);
  const double L = x[n - 1] - x[0] + dx;
  const double s = 2.0 * M_PI / L;
  for (int i = 0; i < n; ++i) {
    const int idx = i <= n / 2 ? i : i - n;
    k[i] = (double)idx * s;
  }
  return k;
}

// Twiddles e^{-2 pi i j/n} in extended precision, rounded once.
inline std::vector<double2> twiddles(int n) {
  std::vector<double2> t(n);
  const long double pi = 3.141592653589793238462643383279502884L;
  for (int j = 0; j < n; ++j) {
    const long double a = -2.0L * pi * (long double)j / (long double)n;
    t[j] = make_double2((double)cosl(a), (double)sinl(a));
  }
  return t;
}

// MATLAB/CPython complex quotient (Smith's algorithm): a / b.
inline std::complex<double> cdiv(std::complex<double> a, std::complex<double> b) {
  const double br = b.real(), bi = b.imag();
  if (std::fabs(br) >= std::fabs(bi)) {
    const double ratio = bi / br;
    const double denom = br + bi * ratio;
    return {(a.real() + a.imag() * ratio) / denom, (a.imag() - a.real() * ratio) / denom};
  }
  const double ratio = br / bi;
  const double denom = br * ratio + bi;
  return {(a.real() * ratio + a.imag()) / denom, (a.imag() * ratio - a.real()) / denom};
}

inline int elem_blocks(long long n) {
  long long nb = (n + 255) / 256;
  if (nb > 2048) nb = 2048;
  if (nb < 1) nb = 1;
  return (int)nb;
}

inline SweepParams base_params(gpe_handle* h, int axis, double2* psi, const double* V) {
  SweepParams p{};
  const Axis& a = h->ax[axis];
  p.psi = psi;
  p.V = V;
  p.tw = a.tw;
  p.keax = a.keax;
  p.kcol = h->kcol;
  p.kax = h->kax;
  p.nop = h->nop;
  p.partials = h->partials;
  p.st = h->st;
  p.ncols = (unsigned)a.ncols;
  p.A = (unsigned)a.A;
  p.sa = (unsigned)a.sa;
  p.sb = (unsigned)a.sb;
  p.stride = (unsigned)a.stride;
  p.maxwg = h->maxwg;
  p.slot_ke = gpe::SL_KX + axis;
  p.g = h->prm.g[0];
  p.dr = h->dr;
  p.di = h->di;
  if (h->omega != 0.0 && axis < 2) {
    // lap(phi, omega) (grid2d/lap.m:8-9, grid3d/lap.m:8-10): x pencils carry -kx y omega,
    // y pencils +ky x omega; the row of a pencil is its y (x pencils) or x (y pencils) index
    p.trow = (unsigned)(axis == 0 ? h->ny : h->nx);
    p.rcoord = h->ax[axis == 0 ? 1 : 0].xc;
    p.rot_w = axis == 0 ? -h->omega : h->omega;
    p.k1ax = a.k1ax;
    p.rtab = axis == 0 ? h->rtab_x : h->rtab_y;
  }
  return p;
}

inline hipEvent_t prof_event(gpe_handle* h) {
  hipEvent_t e = nullptr;
  if (!h->ev_pool.empty()) {
    e = h->ev_pool.back();
    h->ev_pool.pop_back();
  } else if (hipEventCreate(&e) != hipSuccess) {
    e = nullptr;
  }
  return e;
}

inline int sweep(gpe_handle* h, int axis, int seq, const SweepParams& p, int* nb = nullptr) {
  const bool colfast = axis != 1;  // y is contiguous: threads run along the pencil
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (h->prof) {
    e0 = prof_event(h);
    e1 = prof_event(h);
    if (e0) hipEventRecord(e0, h->stream);
  }
  hipError_t e = gpe::launch_sweep(h->ax[axis].n, seq, colfast, p, h->stream, nb, h->var[axis]);
  if (h->prof && e0 && e1) {
    hipEventRecord(e1, h->stream);
    h->prof_recs.push_back({seq * 3 + axis, {e0, e1}});
  }
  if (e != hipSuccess) return fail(h, GPE_EHIP, "sweep launch (axis %d, seq %d): %s", axis, seq, hipGetErrorString(e));
  return GPE_OK;
}

inline int mid_axis(const gpe_handle* h) { return h->ndims == 3 ? 2 : 1; }

// Kinetic tables for complex dt: kcol[col] * kax[i] = exp(-kk dt) / npts
// (set_kinop.m:9) split into the mid-spectrum axis and its columns.
inline int build_kinetic(gpe_handle* h, std::complex<double> dt, double2* kcol, double2* kax, double frac = 1.0,
                         double frac2 = 0.0, double2* kcolB = nullptr, double axmul = 1.0) {
  const int ma = mid_axis(h);
  const Axis& m = h->ax[ma];
  std::vector<double2> col, axv(m.n);
  if (h->ndims == 3) {
    col.resize((size_t)h->nx * h->ny);
    for (int ix = 0; ix < h->nx; ++ix)
      for (int iy = 0; iy < h->ny; ++iy) {
        const double kk = (h->ax[0].k[ix] * h->ax[0].k[ix] + h->ax[1].k[iy] * h->ax[1].k[iy]) / 2;
        std::complex<double> e = std::exp(-kk * dt * frac);
        if (frac2 != 0.0) e *= std::exp(-kk * dt * frac2);
        col[(size_t)iy + (size_t)h->ny * ix] = make_double2(e.real(), e.imag());
      }
  } else {
    col.resize(h->nx);
    for (int ix = 0; ix < h->nx; ++ix) {
      const double kk = h->ax[0].k[ix] * h->ax[0].k[ix] / 2;
      std::complex<double> e = std::exp(-kk * dt * frac);
      if (frac2 != 0.0) e *= std::exp(-kk * dt * frac2);
      col[ix] = make_double2(e.real(), e.imag());
    }
  }
  const double inv = axmul / (double)h->npts;
  for (int i = 0; i < m.n; ++i) {
    const double kk = m.k[i] * m.k[i] / 2;
    std::complex<double> e = std::exp(-kk * dt * frac);
    if (frac2 != 0.0) e *= std::exp(-kk * dt * frac2);
    e *= inv;
    axv[i] = make_double2(e.real(), e.imag());
  }
  HIPCHK(h, hipMemcpyAsync(kcol, col.data(), col.size() * sizeof(double2), hipMemcpyHostToDevice, h->stream));
  HIPCHK(h, hipMemcpyAsync(kax, axv.data(), axv.size() * sizeof(double2), hipMemcpyHostToDevice, h->stream));
  std::vector<double2> colB;
  if (kcolB && h->ndims == 3 && h->ny == 256) {
    // y-split step: entry k2 + 16 s of a z tile (y block s) multiplies ky index
    // bitrev4(s) + 16 k2 (the x sweep leaves its factor in bit-reversed order; the z
    // sweep's factor over a runs in registers in natural order)
    auto br4 = [](int r) { return ((r & 1) << 3) | ((r & 2) << 1) | ((r & 4) >> 1) | ((r & 8) >> 3); };
    colB.resize(col.size());
    for (int ix = 0; ix < h->nx; ++ix)
      for (int pos = 0; pos < 256; ++pos) {
        const int ky = br4(pos >> 4) + 16 * (pos & 15);
        colB[(size_t)pos + 256 * (size_t)ix] = col[(size_t)ky + 256 * (size_t)ix];
      }
    HIPCHK(h, hipMemcpyAsync(kcolB, colB.data(), colB.size() * sizeof(double2), hipMemcpyHostToDevice, h->stream));
  }
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return GPE_OK;
}

// The y-split 3D step can run on this handle (its tables are set up by ysplit_prepare).
inline bool ysplit_eligible(const gpe_handle* h) {
  return !h->ysplit_off && h->ndims == 3 && h->ny == 256 && h->prm.ncomp == 1 &&
         gpe::ysplit_x_supported(h->nx) && gpe::ysplit_x_supported(h->nz) && h->nz % 8 == 0;
}

// Rotating-frame kinetic tables for complex dt (set_kinop.m:6-7):
//   x half step  rtab_x[iy][kx] = exp(-(kx^2/2 - kx y omega) dt/2) / nx
//   3D y/z step  kcol[iy + ny ix] = exp(-(ky^2/2 + ky x omega) dt), kax[kz] = exp(-kz^2/2 dt) / (ny nz)
//   2D y step    rtab_y[ix][ky] = exp(-(ky^2/2 + ky x omega) dt) / ny
inline int build_kinetic_rot(gpe_handle* h, std::complex<double> dt) {
  const double om = h->omega;
  const int nx = h->nx, ny = h->ny;
  const std::vector<double>& kx = h->ax[0].k;
  const std::vector<double>& ky = h->ax[1].k;
  const std::vector<double>& xs = h->ax[0].x;
  const std::vector<double>& ys = h->ax[1].x;
  if (!h->rtab_x) HIPCHK(h, hipMalloc(&h->rtab_x, (size_t)nx * ny * sizeof(double2)));
  std::vector<double2> tx((size_t)nx * ny);
  for (int iy = 0; iy < ny; ++iy)
    for (int ix = 0; ix < nx; ++ix) {
      const std::complex<double> e = std::exp(-(kx[ix] * kx[ix] / 2 - kx[ix] * ys[iy] * om) * dt / 2.0) / (double)nx;
      tx[(size_t)iy * nx + ix] = make_double2(e.real(), e.imag());
    }
  HIPCHK(h, hipMemcpyAsync(h->rtab_x, tx.data(), tx.size() * sizeof(double2), hipMemcpyHostToDevice, h->stream));
  std::vector<double2> ty((size_t)nx * ny);
  if (h->ndims == 3) {
    const int nz = h->nz;
    const std::vector<double>& kz = h->ax[2].k;
    for (int ix = 0; ix < nx; ++ix)
      for (int iy = 0; iy < ny; ++iy) {
        const std::complex<double> e = std::exp(-(ky[iy] * ky[iy] / 2 + ky[iy] * xs[ix] * om) * dt);
        ty[(size_t)iy + (size_t)ny * ix] = make_double2(e.real(), e.imag());
      }
    std::vector<double2> az(nz);
    for (int iz = 0; iz < nz; ++iz) {
      const std::complex<double> e = std::exp(-(kz[iz] * kz[iz] / 2) * dt) / ((double)ny * nz);
      az[iz] = make_double2(e.real(), e.imag());
    }
    HIPCHK(h, hipMemcpyAsync(h->kcol, ty.data(), ty.size() * sizeof(double2), hipMemcpyHostToDevice, h->stream));
    HIPCHK(h, hipMemcpyAsync(h->kax, az.data(), az.size() * sizeof(double2), hipMemcpyHostToDevice, h->stream));
    if (h->kcolB && ysplit_eligible(h)) {
      // the y-split z sweep's order of the same (ky, x) table (see build_kinetic)
      std::vector<double2> tb(ty.size());
      for (int ix = 0; ix < nx; ++ix)
        for (int pos = 0; pos < 256; ++pos) {
          const int r = pos >> 4;
          const int ky = (((r & 1) << 3) | ((r & 2) << 1) | ((r & 4) >> 1) | ((r & 8) >> 3)) + 16 * (pos & 15);
          tb[(size_t)pos + 256 * (size_t)ix] = ty[(size_t)ky + 256 * (size_t)ix];
        }
      HIPCHK(h, hipMemcpyAsync(h->kcolB, tb.data(), tb.size() * sizeof(double2), hipMemcpyHostToDevice, h->stream));
    }
  } else {
    if (!h->rtab_y) HIPCHK(h, hipMalloc(&h->rtab_y, (size_t)nx * ny * sizeof(double2)));
    for (int ix = 0; ix < nx; ++ix)
      for (int iy = 0; iy < ny; ++iy) {
        const std::complex<double> e = std::exp(-(ky[iy] * ky[iy] / 2 + ky[iy] * xs[ix] * om) * dt) / (double)ny;
        ty[(size_t)ix * ny + iy] = make_double2(e.real(), e.imag());
      }
    HIPCHK(h, hipMemcpyAsync(h->rtab_y, ty.data(), ty.size() * sizeof(double2), hipMemcpyHostToDevice, h->stream));
  }
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return GPE_OK;
}

// The kinetic tables of the current frame.

inline int build_step_tables(gpe_handle* h, std::complex<double> dt) {
  if (h->omega != 0.0) return build_kinetic_rot(h, dt);
  return build_kinetic(h, dt, h->kcol, h->kax, 1.0, 0.0, ysplit_eligible(h) ? h->kcolB : nullptr);
}

// The permuted potential of the y-split x sweeps, refreshed when V[0] was written.
inline int ysplit_refresh_v(gpe_handle* h) {
  if (h->vp_gen != h->v_gen) {
    HIPCHK(h, gpe::launch_ysplit_permute_v(h->V[0], h->Vp, (unsigned)h->nx, (unsigned)h->nz, h->stream));
    h->vp_gen = h->v_gen;
  }
  return GPE_OK;
}

// Buffers of the y-split step (allocated once) and the potential in row order.
inline int ysplit_prepare(gpe_handle* h) {
  if (!ysplit_eligible(h)) return GPE_OK;
  if (!h->twy) {
    std::vector<double2> t = twiddles(256);
    HIPCHK(h, hipMalloc(&h->twy, 256 * sizeof(double2)));
    HIPCHK(h, hipMemcpy(h->twy, t.data(), 256 * sizeof(double2), hipMemcpyHostToDevice));
  }
  if (!h->kcolB) HIPCHK(h, hipMalloc(&h->kcolB, (size_t)h->nx * h->ny * sizeof(double2)));
  if (!h->Vp) HIPCHK(h, hipMalloc(&h->Vp, (size_t)h->npts * sizeof(double)));
  if (!h->psiy) HIPCHK(h, hipMalloc(&h->psiy, (size_t)h->npts * sizeof(double2)));
  if (!h->keB) {
    // S_YKEZ weights: (kx^2 + ky^2)/2/npts per z pencil in kcolB's order (entry k2 + 16 s
    // of the pencils of one x is ky index bitrev4(s) + 16 k2), then kz^2/2/npts
    auto br4 = [](int r) { return ((r & 1) << 3) | ((r & 2) << 1) | ((r & 4) >> 1) | ((r & 8) >> 3); };
    const double inv = 1.0 / (double)h->npts;
    const std::vector<double>&kx = h->ax[0].k, &ky = h->ax[1].k, &kz = h->ax[2].k;
    std::vector<double> w((size_t)h->nx * 256 + h->nz);
    for (int ix = 0; ix < h->nx; ++ix)
      for (int pos = 0; pos < 256; ++pos) {
        const int iy = br4(pos >> 4) + 16 * (pos & 15);
        w[(size_t)pos + 256 * (size_t)ix] = (kx[ix] * kx[ix] + ky[iy] * ky[iy]) / 2 * inv;
      }
    for (int iz = 0; iz < h->nz; ++iz) w[(size_t)h->nx * 256 + iz] = kz[iz] * kz[iz] / 2 * inv;
    HIPCHK(h, hipMalloc(&h->keB, w.size() * sizeof(double)));
    HIPCHK(h, hipMemcpy(h->keB, w.data(), w.size() * sizeof(double), hipMemcpyHostToDevice));
  }
  return ysplit_refresh_v(h);
}

inline int read_state(gpe_handle* h, DevState* out) {
  HIPCHK(h, hipMemcpyAsync(out, h->st, sizeof(DevState), hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return GPE_OK;
}

// The block end divides by the norm (solve_split.m:67-73, phi*sqrt(NNN/ncur)) and the
// chemical potential feeds the next block's phase: once either is NaN or Inf the state
// is no longer finite, and the run is stopped with GPE_ENONFINITE instead of continuing.
inline int check_finite(gpe_handle* h, const DevState& s, const char* where) {
  if (!std::isfinite(s.ncur) || !std::isfinite(s.muc) || !std::isfinite(s.scale) || !std::isfinite(s.mu))
    return fail(h, GPE_ENONFINITE, "%s: block reduction is not finite (ncur %g, mu %g, scale %g)", where, s.ncur,
                s.muc, s.scale);
  return GPE_OK;
}

inline int write_state(gpe_handle* h, const DevState& s) {
  HIPCHK(h, hipMemcpyAsync(h->st, &s, sizeof(DevState), hipMemcpyHostToDevice, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return GPE_OK;
}

inline FinParams fin_params(gpe_handle* h, int mode) {
  FinParams f{};
  f.partials = h->partials;
  f.maxwg = h->maxwg;
  f.st = h->st;
  f.mode = mode;
  f.w = h->w;
  f.g = h->prm.g[0];
  return f;
}

// Reductions of the stored state: slots A, B, C and the kinetic term of every
// axis (Parseval per axis: sum_k kk |F psi|^2 / npts = sum_a sum k_a^2/2 |F_a psi|^2 / n_a).
inline int launch_measure(gpe_handle* h, int comp, FinParams* f) {
  int rc, nb = 0;
  SweepParams p = base_params(h, 0, h->psi[comp], h->V[comp]);
  if ((rc = sweep(h, 0, gpe::S_KEX, p, &nb))) return rc;
  f->cnt[gpe::SL_A] = f->cnt[gpe::SL_B] = f->cnt[gpe::SL_C] = f->cnt[gpe::SL_KX] = nb;
  for (int a = 1; a < h->ndims; ++a) {
    SweepParams q = base_params(h, a, h->psi[comp], h->V[comp]);
    if ((rc = sweep(h, a, gpe::S_KE, q, &nb))) return rc;
    f->cnt[gpe::SL_KX + a] = nb;
  }
  return GPE_OK;
}

// The y-split 3D step applies: one component, ny = 256, nx and nz in {64, 128, 256},
// nz a multiple of 8 (ysplit_eligible), with its tables and work arrays in place.
inline bool ysplit_active(const gpe_handle* h) {
  return ysplit_eligible(h) && h->kcolB && h->Vp && h->twy && h->psiy && h->keB;
}

// x-sweep parameters of the y-split step.
inline SweepParams ysplit_params(gpe_handle* h) {
  SweepParams p = base_params(h, 0, h->psi[0], h->Vp);
  p.psiy = h->psiy;
  p.twy = h->twy;
  p.gnx = (unsigned)h->nx;
  p.gny = (unsigned)h->ny;
  p.gnz = (unsigned)h->nz;
  return p;
}

// z-sweep parameters of the y-split step (kinetic factor in the split's order).
inline SweepParams ysplit_zparams(gpe_handle* h) {
  SweepParams p = base_params(h, 2, h->psi[0], h->V[0]);
  p.kcol = h->kcolB;
  p.psiy = h->psiy;
  p.twy = h->twy;
  return p;
}

// One y-split x sweep, timed like the other sweeps when profiling.
inline int ysplit_xsweep(gpe_handle* h, int seq, const SweepParams& px, int* nb) {
  if (16 * h->nz > h->maxwg)
    return fail(h, GPE_EINVAL, "y-split x sweep: %d work-groups exceed the partial-sum pitch %d", 16 * h->nz, h->maxwg);
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (h->prof) {
    e0 = prof_event(h);
    e1 = prof_event(h);
    if (e0) hipEventRecord(e0, h->stream);
  }
  hipError_t e = gpe::launch_ysplit_x(h->nx, seq, px, h->stream, nb);
  if (h->prof && e0 && e1) {
    hipEventRecord(e1, h->stream);
    h->prof_recs.push_back({seq * 3 + 0, {e0, e1}});
  }
  if (e != hipSuccess) return fail(h, GPE_EHIP, "y-split x sweep (seq %d): %s", seq, hipGetErrorString(e));
  return GPE_OK;
}

// One block of solve_split.m:53-78 as two sweeps per step (y factor over b with the x
// sweep, y factor over a with the z sweep): same operations as run_block, the y
// transform factored 16 x 16.  blk_in: the previous block of the same outer iteration
// left the state in psiy; blk_out: leave it there for the next one.  Without rotation
// the hand-over is the mixed form of the steps (S_YCLOSEB, S_YKEZ, S_YOPENB); in the
// rotating frame it is the real-space state in block order (its block-end kinetic sums
// read the y pencils through ysplit_pos, S_YKE, and the z pencils in their permuted
// column order, which leaves the sum over columns unchanged).
inline int run_block_ysplit(gpe_handle* h, double time2, bool blk_in, bool blk_out) {
  int rc;
  if ((rc = ysplit_refresh_v(h))) return rc;
  SweepParams px = ysplit_params(h);
  px.yblk_in = blk_in ? 1 : 0;
  px.yblk_out = blk_out ? 1 : 0;
  const SweepParams pz = ysplit_zparams(h);
  auto xs = [&](int seq, int* nb) { return ysplit_xsweep(h, seq, px, nb); };
  const bool rot = h->omega != 0.0;
  const int s_open = rot ? gpe::S_YROPEN : (blk_in ? gpe::S_YOPENB : gpe::S_YOPEN);
  const int s_mid = rot ? gpe::S_YRMID : gpe::S_YMID;
  for (int i = 0; i < h->n_rec; ++i) {
    if ((rc = xs(i > 0 ? s_mid : s_open, nullptr))) return rc;
    if ((rc = sweep(h, 2, gpe::S_YKIN, pz))) return rc;
  }
  FinParams f = fin_params(h, gpe::FIN_BLOCK);
  int nb = 0;
  if (!rot && blk_out) {
    if ((rc = xs(gpe::S_YCLOSEB, &nb))) return rc;
    f.cnt[gpe::SL_A] = f.cnt[gpe::SL_B] = f.cnt[gpe::SL_C] = nb;
    SweepParams pk = ysplit_zparams(h);
    pk.kecol = h->keB;
    pk.slot_ke = gpe::SL_KZ;
    if ((rc = sweep(h, 2, gpe::S_YKEZ, pk, &nb))) return rc;
    f.cnt[gpe::SL_KZ] = nb;   // the whole kinetic sum; the x and y slots stay empty
  } else {
    if ((rc = xs(rot ? gpe::S_YRCLOSEM : gpe::S_YCLOSEM, &nb))) return rc;
    f.cnt[gpe::SL_A] = f.cnt[gpe::SL_B] = f.cnt[gpe::SL_C] = f.cnt[gpe::SL_KX] = nb;
    double2* st = blk_out ? h->psiy : h->psi[0];
    for (int a = 1; a < 3; ++a) {
      const int seq = blk_out && a == 1 ? gpe::S_YKE : gpe::S_KE;
      if ((rc = sweep(h, a, seq, base_params(h, a, st, h->V[0]), &nb))) return rc;
      f.cnt[gpe::SL_KX + a] = nb;
    }
  }
  f.renorm = (h->prm.gamma > 0 && h->prm.mu_init == 0) ? 1 : 0;
  f.NN0 = h->NN0;
  f.tau = h->prm.decay_rate;
  f.time2 = time2;
  HIPCHK(h, gpe::launch_finalize(f, h->stream));
  return GPE_OK;
}

// blk_in / blk_out: the block is not the first / last of its outer iteration (the
// y-split step keeps the state in block order between such blocks).
inline int run_block(gpe_handle* h, double time2, bool blk_in = false, bool blk_out = false) {
  const bool rot = h->omega != 0.0;
  if (ysplit_active(h)) return run_block_ysplit(h, time2, blk_in, blk_out);
  int rc;
  double2* psi = h->psi[0];
  const double* V = h->V[0];
  // rotating frame: the x sweeps carry the x half steps of ssft_kin_step.m:23,25 around
  // the phases, and the 2D y step uses the (x, ky) table
  const int s_open = rot ? gpe::S_ROPEN : gpe::S_OPEN, s_mid = rot ? gpe::S_RMID : gpe::S_MID;
  const int s_close = rot ? gpe::S_RCLOSEM : gpe::S_CLOSEM, s_kin2d = rot ? gpe::S_RKIN : gpe::S_KIN;
  for (int i = 0; i < h->n_rec; ++i) {
    if ((rc = sweep(h, 0, i > 0 ? s_mid : s_open, base_params(h, 0, psi, V)))) return rc;
    if (h->ndims == 3) {
      if ((rc = sweep(h, 1, gpe::S_FWD, base_params(h, 1, psi, V)))) return rc;
      if ((rc = sweep(h, 2, gpe::S_KIN, base_params(h, 2, psi, V)))) return rc;
      if ((rc = sweep(h, 1, gpe::S_INV, base_params(h, 1, psi, V)))) return rc;
    } else {
      if ((rc = sweep(h, 1, s_kin2d, base_params(h, 1, psi, V)))) return rc;
    }
  }
  // closing phases fused with the x part of the block-end reductions
  FinParams f = fin_params(h, gpe::FIN_BLOCK);
  int nb = 0;
  if ((rc = sweep(h, 0, s_close, base_params(h, 0, psi, V), &nb))) return rc;
  f.cnt[gpe::SL_A] = f.cnt[gpe::SL_B] = f.cnt[gpe::SL_C] = f.cnt[gpe::SL_KX] = nb;
  for (int a = 1; a < h->ndims; ++a) {
    if ((rc = sweep(h, a, gpe::S_KE, base_params(h, a, psi, V), &nb))) return rc;
    f.cnt[gpe::SL_KX + a] = nb;
  }
  f.renorm = (h->prm.gamma > 0 && h->prm.mu_init == 0) ? 1 : 0;
  f.NN0 = h->NN0;
  f.tau = h->prm.decay_rate;
  f.time2 = time2;
  HIPCHK(h, gpe::launch_finalize(f, h->stream));
  return GPE_OK;
}

inline int ensure_ready(gpe_handle* h) {
  if (!h) return GPE_EINVAL;
  if (!h->have_params) return fail(h, GPE_ESTATE, "gpe_set_params has not been called");
  for (int c = 0; c < h->prm.ncomp; ++c)
    if (!h->have_V[c]) return fail(h, GPE_ESTATE, "potential of component %d not set", c);
  HIPCHK(h, hipSetDevice(h->device));
  return GPE_OK;
}

inline void drop_outer_graph(gpe_handle* h) {
  if (h->ogx) (void)hipGraphExecDestroy(h->ogx);
  if (h->og) (void)hipGraphDestroy(h->og);
  h->ogx = nullptr;
  h->og = nullptr;
}

// Runs `launches` (the kernels of one outer iteration, no host synchronisation) through
// a captured graph when `graphable`: the first call captures, later calls replay while
// the host values in key[0..7] that the launches bake in are unchanged (else recapture).
// Profiling, GPE_NO_GRAPH=1 and a failed capture run the launches directly.
template <class F>
inline int run_outer(gpe_handle* h, bool graphable, const double (&key)[8], F launches) {
  if (!graphable || h->og_off || h->prof) return launches();
  if (h->ogx && std::memcmp(h->og_key, key, sizeof key) != 0) drop_outer_graph(h);
  if (!h->ogx) {
    HIPCHK(h, hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
    const int crc = launches();
    hipGraph_t g = nullptr;
    const hipError_t ec = hipStreamEndCapture(h->stream, &g);
    if (!crc && ec == hipSuccess && g && hipGraphInstantiate(&h->ogx, g, nullptr, nullptr, 0) == hipSuccess) {
      h->og = g;
      std::memcpy(h->og_key, key, sizeof key);
    } else {
      if (g) (void)hipGraphDestroy(g);
      h->ogx = nullptr;
      h->og_off = true;
      (void)hipGetLastError();
      return crc ? crc : launches();
    }
  }
  HIPCHK(h, hipGraphLaunch(h->ogx, h->stream));
  return GPE_OK;
}

// Apply a pending renormalisation scale to the stored state(s).
inline int materialize(gpe_handle* h) {
  DevState s;
  int rc;
  if ((rc = read_state(h, &s))) return rc;
  if (s.scale != 1.0) {
    for (int c = 0; c < h->prm.ncomp; ++c)
      HIPCHK(h, gpe::launch_scale(h->psi[c], h->st, 0, h->npts, elem_blocks(h->npts), h->stream,
                                  c + 1 == h->prm.ncomp));
    HIPCHK(h, hipStreamSynchronize(h->stream));
  }
  return GPE_OK;
}

}  // namespace gpe_detail
