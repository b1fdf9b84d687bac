// gpe_mex -- MATLAB MEX gateway over the C ABI of libgpehip.so (include/gpe_hip.h).
//
//   out = gpe_mex(command, args...)
//
// The reference (GPE-MATLAB) dispatches MATLAB methods on handle classes; the
// wrapper class in ../matlab/@GPEtaskHIP keeps those methods (solve_split,
// groundstate_itp, ...) and replaces their bodies by calls through this gateway.
// Every command maps onto one or two gpe_* entry points; the comment at each case
// names the reference code the call replaces.
//
// Build (MATLAB R2018a or later, interleaved complex API):
//   mex -R2018a gpe_mex.cpp -I<repo>/include -L<repo>/gpe-matlab_amd/gpe_hip -lgpehip
//
// Handles.  'create' and 'batch_create' return a uint64 id; the gateway keeps the
// handle, its grid sizes and a registry of live ids (an unknown id is an error, not
// a fault), and destroys whatever is still open when MATLAB clears the MEX file.
//
// Arrays.  States and potentials are MATLAB arrays of the grid size [ny nx (nz)],
// column-major, which is the library's layout; real arrays given as states are
// widened to complex.  Every failing call raises a MATLAB error with the library's
// message (gpe_last_error); GPE_ENONFINITE raises gpe:nonfinite.
#include <array>
#include <cstdint>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "gpe_hip.h"
#include "mex.h"

namespace {

struct Task {
  gpe_handle* h = nullptr;
  int ndims = 2;
  int n[3] = {1, 1, 1};   // nx, ny, nz
  int ncomp = 0;          // components of the last accepted 'params' (0: none yet)
  size_t npts() const { return (size_t)n[0] * n[1] * n[2]; }
};

struct Batch {
  gpe_batch* b = nullptr;
  std::vector<std::array<int, 2>> n;   // nx, ny per trajectory
};

std::map<uint64_t, Task> tasks;
std::map<uint64_t, Batch> batches;
uint64_t next_id = 1;

void cleanup() {
  for (auto& kv : tasks) gpe_destroy(kv.second.h);
  for (auto& kv : batches) gpe_batch_destroy(kv.second.b);
  tasks.clear();
  batches.clear();
}

void need(int nrhs, int k, const char* cmd) {
  if (nrhs < k) mexErrMsgIdAndTxt("gpe:args", "gpe_mex('%s'): %d arguments needed", cmd, k - 1);
}

void check(int rc, const gpe_handle* h, const char* what) {
  if (rc == GPE_OK) return;
  const char* msg = h ? gpe_last_error(h) : "";
  if (rc == GPE_ENONFINITE) mexErrMsgIdAndTxt("gpe:nonfinite", "%s: %s", what, msg);
  mexErrMsgIdAndTxt("gpe:hip", "%s failed (%d): %s", what, rc, msg);
}

void checkb(int rc, const gpe_batch* b, const char* what) {
  if (rc == GPE_OK) return;
  const char* msg = b ? gpe_batch_last_error(b) : "";
  if (rc == GPE_ENONFINITE) mexErrMsgIdAndTxt("gpe:nonfinite", "%s: %s", what, msg);
  mexErrMsgIdAndTxt("gpe:hip", "%s failed (%d): %s", what, rc, msg);
}

uint64_t id_of(const mxArray* a) {
  if (!mxIsUint64(a) || mxGetNumberOfElements(a) != 1) mexErrMsgIdAndTxt("gpe:handle", "handle must be a uint64 id");
  return *static_cast<const uint64_t*>(mxGetData(a));
}

Task& task_of(const mxArray* a) {
  auto it = tasks.find(id_of(a));
  if (it == tasks.end()) mexErrMsgIdAndTxt("gpe:handle", "unknown or destroyed task handle");
  return it->second;
}

Batch& batch_of(const mxArray* a) {
  auto it = batches.find(id_of(a));
  if (it == batches.end()) mexErrMsgIdAndTxt("gpe:handle", "unknown or destroyed batch handle");
  return it->second;
}

mxArray* new_id(uint64_t id) {
  mxArray* r = mxCreateNumericMatrix(1, 1, mxUINT64_CLASS, mxREAL);
  *static_cast<uint64_t*>(mxGetData(r)) = id;
  return r;
}

int as_int(const mxArray* a) { return (int)mxGetScalar(a); }

// Real double array of exactly n elements.
const double* reals(const mxArray* a, size_t n, const char* what) {
  if (!mxIsDouble(a) || mxIsComplex(a) || mxGetNumberOfElements(a) != n)
    mexErrMsgIdAndTxt("gpe:args", "%s: real double array of %zu elements expected", what, n);
  return mxGetDoubles(a);
}

// Interleaved complex view of a double array of n elements (a real array is widened
// into `tmp`).
const double* complexes(const mxArray* a, size_t n, std::vector<double>& tmp, const char* what) {
  if (!mxIsDouble(a) || mxGetNumberOfElements(a) != n)
    mexErrMsgIdAndTxt("gpe:args", "%s: double array of %zu elements expected", what, n);
  if (mxIsComplex(a)) return reinterpret_cast<const double*>(mxGetComplexDoubles(a));
  const double* r = mxGetDoubles(a);
  tmp.assign(2 * n, 0.0);
  for (size_t i = 0; i < n; ++i) tmp[2 * i] = r[i];
  return tmp.data();
}

// Complex array of the grid's size: [ny nx] or [ny nx nz] (grid.size).
mxArray* grid_complex(const Task& t, int ncopies = 1) {
  mwSize d[4] = {(mwSize)t.n[1], (mwSize)t.n[0], (mwSize)t.n[2], (mwSize)ncopies};
  mwSize nd = t.ndims == 3 ? 3 : 2;
  if (ncopies > 1) {
    if (t.ndims == 2) d[2] = (mwSize)ncopies;
    ++nd;
  }
  return mxCreateNumericArray(nd, d, mxDOUBLE_CLASS, mxCOMPLEX);
}

double* cdata(mxArray* a) { return reinterpret_cast<double*>(mxGetComplexDoubles(a)); }

mxArray* column(const double* v, int n) {
  mxArray* r = mxCreateDoubleMatrix(n, 1, mxREAL);
  if (n > 0) std::memcpy(mxGetDoubles(r), v, (size_t)n * sizeof(double));
  return r;
}

void set_out(int nlhs, mxArray* plhs[], int i, mxArray* v) {
  if (i < nlhs || (i == 0 && nlhs == 0)) plhs[i] = v;
  else mxDestroyArray(v);
}

}  // namespace

void mexFunction(int nlhs, mxArray* plhs[], int nrhs, const mxArray* prhs[]) {
  static bool registered = false;
  if (!registered) {
    mexAtExit(cleanup);
    registered = true;
  }
  if (nrhs < 1 || !mxIsChar(prhs[0])) mexErrMsgIdAndTxt("gpe:args", "gpe_mex(command, ...)");
  char buf[64];
  mxGetString(prhs[0], buf, sizeof buf);
  const std::string cmd(buf);
  const char* c = buf;

  if (cmd == "create") {
    // grid2d/grid3d(xmax, nx, ...) + GPEtask(grid, V) (grid3d/init.m:5-46, GPEtask.m:41-60):
    // gpe_mex('create', [nx ny (nz)], [xmax ymax (zmax)])
    need(nrhs, 3, c);
    const size_t nd = mxGetNumberOfElements(prhs[1]);
    if ((nd != 2 && nd != 3) || mxGetNumberOfElements(prhs[2]) != nd)
      mexErrMsgIdAndTxt("gpe:args", "create: 2 or 3 sizes and extents expected");
    gpe_grid_desc d{};
    d.ndims = (int)nd;
    Task t;
    t.ndims = (int)nd;
    for (size_t a = 0; a < nd; ++a) {
      d.n[a] = t.n[a] = (int)mxGetDoubles(prhs[1])[a];
      d.xmax[a] = mxGetDoubles(prhs[2])[a];
    }
    check(gpe_create(&d, &t.h), nullptr, "gpe_create");
    const uint64_t id = next_id++;
    tasks[id] = t;
    set_out(nlhs, plhs, 0, new_id(id));
  } else if (cmd == "destroy") {
    need(nrhs, 2, c);
    auto it = tasks.find(id_of(prhs[1]));
    if (it != tasks.end()) {
      gpe_destroy(it->second.h);
      tasks.erase(it);
    }
  } else if (cmd == "tune") {
    // FFTW planning of the GPEtask constructor (GPEtask.m:28, 55-58): gpe_mex('tune', h)
    need(nrhs, 2, c);
    Task& t = task_of(prhs[1]);
    double us[6];
    check(gpe_tune(t.h, 0, us), t.h, "gpe_tune");
    set_out(nlhs, plhs, 0, column(us, 6));
  } else if (cmd == "forget_wisdom") {
    // fftw('dwisdom', [])
    gpe_wisdom_forget();
  } else if (cmd == "wisdom_export") {
    // fftw('wisdom')
    int needed = 0;
    check(gpe_wisdom_export(nullptr, 0, &needed), nullptr, "gpe_wisdom_export");
    std::vector<char> text((size_t)needed + 1, '\0');
    check(gpe_wisdom_export(text.data(), (int)text.size(), &needed), nullptr, "gpe_wisdom_export");
    set_out(nlhs, plhs, 0, mxCreateString(text.data()));
  } else if (cmd == "wisdom_import") {
    need(nrhs, 2, c);
    char* text = mxArrayToString(prhs[1]);
    const int rc = gpe_wisdom_import(text);
    mxFree(text);
    check(rc, nullptr, "gpe_wisdom_import");
  } else if (cmd == "params") {
    // task properties (GPEtask.m:4-34, GPEtaskSpinor.m:5-9):
    // gpe_mex('params', h, g, gamma, decay_rate, Ntotal, mu_init, n_recalc, nlincpl)
    need(nrhs, 9, c);
    Task& t = task_of(prhs[1]);
    gpe_params p{};
    const size_t ng = mxGetNumberOfElements(prhs[2]);
    if (ng != 1 && ng != 4) mexErrMsgIdAndTxt("gpe:args", "params: g is a scalar or a 2x2 matrix");
    p.ncomp = ng == 4 ? 2 : 1;
    std::memcpy(p.g, reals(prhs[2], ng, "g"), ng * sizeof(double));   // column-major {g11, g21, g12, g22}
    p.gamma = mxGetScalar(prhs[3]);
    p.decay_rate = mxGetScalar(prhs[4]);
    p.Ntotal = mxGetScalar(prhs[5]);
    p.mu_init = mxGetScalar(prhs[6]);
    p.n_recalc = as_int(prhs[7]);
    p.nlincpl = mxGetScalar(prhs[8]);
    check(gpe_set_params(t.h, &p), t.h, "gpe_set_params");
    t.ncomp = p.ncomp;
  } else if (cmd == "omega") {
    // the omega property (GPEtask.m:7) as set_kinop.m:5-7 reads it
    need(nrhs, 3, c);
    Task& t = task_of(prhs[1]);
    check(gpe_set_omega(t.h, mxGetScalar(prhs[2])), t.h, "gpe_set_omega");
  } else if (cmd == "potential" || cmd == "potential_mu") {
    // Vtrap / getVtotal(time) (GPEtask.m:62-71); potential_mu: the potential of the
    // two-component block-end applyham (GPEtaskSpinor/solve_split.m:109-114), [] resets
    need(nrhs, 4, c);
    Task& t = task_of(prhs[1]);
    const int comp = as_int(prhs[2]);
    if (cmd == "potential") {
      check(gpe_set_potential(t.h, comp, reals(prhs[3], t.npts(), "V")), t.h, "gpe_set_potential");
    } else {
      const double* V = mxIsEmpty(prhs[3]) ? nullptr : reals(prhs[3], t.npts(), "V");
      check(gpe_set_potential_mu(t.h, comp, V), t.h, "gpe_set_potential_mu");
    }
  } else if (cmd == "coupling") {
    // coupling property (GPEtaskSpinor.m:6): scalar or array, complex
    need(nrhs, 3, c);
    Task& t = task_of(prhs[1]);
    const size_t n = mxGetNumberOfElements(prhs[2]);
    std::vector<double> tmp;
    if (n == 1) {
      const double* w = complexes(prhs[2], 1, tmp, "coupling");
      check(gpe_set_coupling(t.h, w, 0), t.h, "gpe_set_coupling");
    } else {
      const double* w = complexes(prhs[2], t.npts(), tmp, "coupling");
      check(gpe_set_coupling(t.h, w, 1), t.h, "gpe_set_coupling");
    }
  } else if (cmd == "set_state") {
    // init_state / current_state upload (GPEtask.m:18-19): gpe_mex('set_state', h, comp, psi)
    need(nrhs, 4, c);
    Task& t = task_of(prhs[1]);
    std::vector<double> tmp;
    check(gpe_set_state(t.h, as_int(prhs[2]), complexes(prhs[3], t.npts(), tmp, "psi")), t.h, "gpe_set_state");
  } else if (cmd == "get_state") {
    // current_state download (GPEtask.m:119): psi = gpe_mex('get_state', h, comp)
    need(nrhs, 3, c);
    Task& t = task_of(prhs[1]);
    mxArray* out = grid_complex(t);
    check(gpe_get_state(t.h, as_int(prhs[2]), cdata(out)), t.h, "gpe_get_state");
    set_out(nlhs, plhs, 0, out);
  } else if (cmd == "split_begin") {
    // solve_split.m:13-46: muc = gpe_mex('split_begin', h, ddt, niter_inner)
    need(nrhs, 4, c);
    Task& t = task_of(prhs[1]);
    double muc = 0.0;
    check(gpe_split_begin(t.h, mxGetScalar(prhs[2]), as_int(prhs[3]), &muc), t.h, "gpe_split_begin");
    set_out(nlhs, plhs, 0, mxCreateDoubleScalar(muc));
  } else if (cmd == "split_outer") {
    // one outer iteration, solve_split.m:49-79: [time2, muc, ncur] = gpe_mex('split_outer', h, j)
    need(nrhs, 3, c);
    Task& t = task_of(prhs[1]);
    double t2 = 0, m = 0, n = 0;
    check(gpe_split_outer(t.h, as_int(prhs[2]), &t2, &m, &n), t.h, "gpe_split_outer");
    set_out(nlhs, plhs, 0, mxCreateDoubleScalar(t2));
    set_out(nlhs, plhs, 1, mxCreateDoubleScalar(m));
    set_out(nlhs, plhs, 2, mxCreateDoubleScalar(n));
  } else if (cmd == "split_block") {
    // one block, solve_split.m:53-78 (Vtd a host function): [muc, ncur] = gpe_mex('split_block', h, time2)
    need(nrhs, 3, c);
    Task& t = task_of(prhs[1]);
    double m = 0, n = 0;
    check(gpe_split_block(t.h, mxGetScalar(prhs[2]), &m, &n), t.h, "gpe_split_block");
    set_out(nlhs, plhs, 0, mxCreateDoubleScalar(m));
    set_out(nlhs, plhs, 1, mxCreateDoubleScalar(n));
  } else if (cmd == "measure") {
    // applyham/get_energy reductions (GPEtask.m:73-107): m = gpe_mex('measure', h)
    need(nrhs, 2, c);
    Task& t = task_of(prhs[1]);
    double m[5];
    check(gpe_measure(t.h, m), t.h, "gpe_measure");
    set_out(nlhs, plhs, 0, column(m, 5));
  } else if (cmd == "applyham") {
    // applyham / applyh0 as arrays (GPEtask.m:73-95, GPEtaskSpinor.m:45-86):
    // H = gpe_mex('applyham', h, phi, h0_only); phi = [] takes the current state.
    // Two components (a 2x2 g in 'params'): phi and H are [grid.size, 2], the component
    // along the last dimension.  The output is sized from the task's component count.
    need(nrhs, 4, c);
    Task& t = task_of(prhs[1]);
    const int nc = t.ncomp;
    if (nc != 1 && nc != 2) mexErrMsgIdAndTxt("gpe:args", "applyham: call 'params' first");
    std::vector<double> tmp;
    const double* in = mxIsEmpty(prhs[2]) ? nullptr : complexes(prhs[2], (size_t)nc * t.npts(), tmp, "phi");
    mxArray* out = grid_complex(t, nc);
    check(gpe_applyham(t.h, in, cdata(out), as_int(prhs[3])), t.h, "gpe_applyham");
    set_out(nlhs, plhs, 0, out);
  } else if (cmd == "grid_transform") {
    // the grid's fftn / ifftn / lap(phi, omega) on the device (grid2d/fft.m, ifft.m,
    // lap.m; grid3d/fft.m, ifft.m, lap.m): out = gpe_mex('grid_transform', h, op, omega, phi)
    // with op 0 = fft, 1 = ifft, 2 = lap
    need(nrhs, 5, c);
    Task& t = task_of(prhs[1]);
    std::vector<double> tmp;
    const double* in = complexes(prhs[4], (size_t)t.npts(), tmp, "phi");
    mxArray* out = grid_complex(t, 1);
    check(gpe_grid_transform(t.h, as_int(prhs[2]), mxGetScalar(prhs[3]), in, cdata(out)), t.h, "gpe_grid_transform");
    set_out(nlhs, plhs, 0, out);
  } else if (cmd == "itp") {
    // groundstate_itp.m:20-112 after process_init_state:
    // [MU, MU2, EE, n_iter, dt_final, converged] = gpe_mex('itp', h, dt, eps, max_iter, min_iter, adjust, nargout)
    need(nrhs, 8, c);
    Task& t = task_of(prhs[1]);
    gpe_itp_args a{};
    a.dt = mxGetScalar(prhs[2]);
    a.eps = mxGetScalar(prhs[3]);
    a.max_iter = as_int(prhs[4]);
    a.min_iter = as_int(prhs[5]);
    a.adjust_stepsize = as_int(prhs[6]);
    a.nargout = as_int(prhs[7]);
    if (a.max_iter < 1) mexErrMsgIdAndTxt("gpe:args", "itp: max_iter must be >= 1");
    std::vector<double> MU(a.max_iter), MU2(a.max_iter), EE(a.max_iter);
    int n_iter = 0, conv = 0;
    double dtf = 0;
    check(gpe_itp(t.h, &a, MU.data(), MU2.data(), EE.data(), &n_iter, &dtf, &conv), t.h, "gpe_itp");
    set_out(nlhs, plhs, 0, column(MU.data(), n_iter));
    set_out(nlhs, plhs, 1, column(MU2.data(), n_iter));
    set_out(nlhs, plhs, 2, column(EE.data(), n_iter));
    set_out(nlhs, plhs, 3, mxCreateDoubleScalar(n_iter));
    set_out(nlhs, plhs, 4, mxCreateDoubleScalar(dtf));
    set_out(nlhs, plhs, 5, mxCreateDoubleScalar(conv));
  } else if (cmd == "itp2") {
    // GPEtaskSpinor/GPEtask2comp.groundstate_itp after the initial state:
    // [MU, MU2, EE, n_iter, dt_final, converged] =
    //     gpe_mex('itp2', h, dt, eps, max_iter, nargout, mixture, N1, N2, M1, M2)
    need(nrhs, 11, c);
    Task& t = task_of(prhs[1]);
    gpe_itp2_args a{};
    a.dt = mxGetScalar(prhs[2]);
    a.eps = mxGetScalar(prhs[3]);
    a.max_iter = as_int(prhs[4]);
    a.nargout = as_int(prhs[5]);
    a.mixture = as_int(prhs[6]);
    a.N1 = mxGetScalar(prhs[7]);
    a.N2 = mxGetScalar(prhs[8]);
    a.M1 = mxGetScalar(prhs[9]);
    a.M2 = mxGetScalar(prhs[10]);
    if (a.max_iter < 1) mexErrMsgIdAndTxt("gpe:args", "itp2: max_iter must be >= 1");
    std::vector<double> MU(a.max_iter), MU2(a.max_iter), EE(a.max_iter);
    int n_iter = 0, conv = 0;
    double dtf = 0;
    check(gpe_itp2(t.h, &a, MU.data(), MU2.data(), EE.data(), &n_iter, &dtf, &conv), t.h, "gpe_itp2");
    const int nh = n_iter > 0 ? n_iter - 1 : 0;   // entries 1..n_iter-1 (groundstate_itp.m:131-142)
    set_out(nlhs, plhs, 0, column(MU.data(), nh));
    set_out(nlhs, plhs, 1, column(MU2.data(), nh));
    set_out(nlhs, plhs, 2, column(EE.data(), nh));
    set_out(nlhs, plhs, 3, mxCreateDoubleScalar(n_iter));
    set_out(nlhs, plhs, 4, mxCreateDoubleScalar(dtf));
    set_out(nlhs, plhs, 5, mxCreateDoubleScalar(conv));
  } else if (cmd == "sgpe_begin") {
    // SGPEtask/solve_split.m:13-47: gpe_mex('sgpe_begin', h, ddt, niter_inner, T, ecut, seed)
    need(nrhs, 7, c);
    Task& t = task_of(prhs[1]);
    gpe_sgpe_args a{};
    a.ddt = mxGetScalar(prhs[2]);
    a.niter_inner = as_int(prhs[3]);
    a.T = mxGetScalar(prhs[4]);
    a.ecut = mxGetScalar(prhs[5]);
    a.seed = (unsigned long long)mxGetScalar(prhs[6]);
    check(gpe_sgpe_begin(t.h, &a), t.h, "gpe_sgpe_begin");
  } else if (cmd == "sgpe_outer") {
    // SGPEtask/solve_split.m:49-80: [time2, ncur] = gpe_mex('sgpe_outer', h, j)
    need(nrhs, 3, c);
    Task& t = task_of(prhs[1]);
    double t2 = 0, n = 0;
    check(gpe_sgpe_outer(t.h, as_int(prhs[2]), &t2, &n), t.h, "gpe_sgpe_outer");
    set_out(nlhs, plhs, 0, mxCreateDoubleScalar(t2));
    set_out(nlhs, plhs, 1, mxCreateDoubleScalar(n));
  } else if (cmd == "sgpe_block") {
    // SGPEtask/solve_split.m:53-76: ncur = gpe_mex('sgpe_block', h, step0)
    need(nrhs, 3, c);
    Task& t = task_of(prhs[1]);
    double n = 0;
    check(gpe_sgpe_block(t.h, (long long)mxGetScalar(prhs[2]), &n), t.h, "gpe_sgpe_block");
    set_out(nlhs, plhs, 0, mxCreateDoubleScalar(n));
  } else if (cmd == "vtd_set") {
    // a Vtd that is a linear combination of fixed fields (GPEtask.m:62-71):
    // gpe_mex('vtd_set', h, G (npts x nbasis), Vtrap (npts x ncomp))
    need(nrhs, 4, c);
    Task& t = task_of(prhs[1]);
    const size_t n = t.npts();
    const size_t nb = mxGetNumberOfElements(prhs[2]) / (n ? n : 1);
    reals(prhs[2], nb * n, "G");
    reals(prhs[3], mxGetNumberOfElements(prhs[3]), "Vtrap");
    check(gpe_vtd_set(t.h, (int)nb, mxGetDoubles(prhs[2]), mxGetDoubles(prhs[3])), t.h, "gpe_vtd_set");
  } else if (cmd == "vtd_coeffs") {
    // per-block getVtotal(time2) on the device: gpe_mex('vtd_coeffs', h, c, target)
    need(nrhs, 4, c);
    Task& t = task_of(prhs[1]);
    check(gpe_vtd_coeffs(t.h, reals(prhs[2], mxGetNumberOfElements(prhs[2]), "c"), as_int(prhs[3])), t.h,
          "gpe_vtd_coeffs");
  } else if (cmd == "detect_core") {
    // lib/util/detect_core.m:1-66 on the resident state:
    // [resp, resm] = gpe_mex('detect_core', h, comp, zslice, x (1 x nx), y (ny x 1), refine)
    need(nrhs, 7, c);
    Task& t = task_of(prhs[1]);
    const double* x = reals(prhs[4], (size_t)t.n[0], "x");
    const double* y = reals(prhs[5], (size_t)t.n[1], "y");
    int cap = 64, np = 0, nm = 0;
    std::vector<double> rp, rm;
    for (;;) {
      rp.assign(2 * (size_t)cap, 0.0);
      rm.assign(2 * (size_t)cap, 0.0);
      check(gpe_detect_core(t.h, as_int(prhs[2]), as_int(prhs[3]), x, y, as_int(prhs[6]), cap, rp.data(), &np,
                            rm.data(), &nm),
            t.h, "gpe_detect_core");
      if (np <= cap && nm <= cap) break;
      cap = np > nm ? np : nm;
    }
    // rows [x, y] (row-major in the library) -> n x 2 MATLAB matrices
    auto rows = [](const std::vector<double>& r, int k) {
      mxArray* a = mxCreateDoubleMatrix(k, 2, mxREAL);
      double* d = mxGetDoubles(a);
      for (int i = 0; i < k; ++i) {
        d[i] = r[2 * (size_t)i];
        d[i + k] = r[2 * (size_t)i + 1];
      }
      return a;
    };
    set_out(nlhs, plhs, 0, rows(rp, np));
    set_out(nlhs, plhs, 1, rows(rm, nm));
  } else if (cmd == "control_set" || cmd == "control_forward" || cmd == "adjoint_grad") {
    // gradient entry point (no reference counterpart):
    //   gpe_mex('control_set', h, V0, G (npts x nbasis))
    //   psi = gpe_mex('control_forward', h, ddt, nsteps, steps_per_interval, coeffs (nbasis x nint))
    //   [loss, grad, recon] = gpe_mex('adjoint_grad', h, ddt, nsteps, steps_per_interval, coeffs, psi_target)
    Task& t = task_of(prhs[1]);
    const size_t n = t.npts();
    if (cmd == "control_set") {
      need(nrhs, 4, c);
      const size_t nb = mxGetNumberOfElements(prhs[3]) / (n ? n : 1);
      check(gpe_control_set(t.h, (int)nb, reals(prhs[2], n, "V0"), reals(prhs[3], nb * n, "G")), t.h,
            "gpe_control_set");
      return;
    }
    need(nrhs, cmd == "adjoint_grad" ? 7 : 6, c);
    const mxArray* cf = prhs[5];
    gpe_control_args a{};
    a.ddt = mxGetScalar(prhs[2]);
    a.nsteps = as_int(prhs[3]);
    a.steps_per_interval = as_int(prhs[4]);
    a.nintervals = (int)mxGetN(cf);
    a.checkpoint_blocks = 0;
    const double* coeffs = reals(cf, mxGetNumberOfElements(cf), "coeffs");   // column-major nbasis x nint
    if (cmd == "control_forward") {
      mxArray* out = grid_complex(t);
      check(gpe_control_forward(t.h, &a, coeffs, cdata(out)), t.h, "gpe_control_forward");
      set_out(nlhs, plhs, 0, out);
    } else {
      std::vector<double> tmp;
      const double* target = complexes(prhs[6], n, tmp, "psi_target");
      double loss = 0, recon = 0;
      mxArray* g = mxCreateDoubleMatrix(mxGetM(cf), mxGetN(cf), mxREAL);
      check(gpe_adjoint_grad(t.h, &a, coeffs, target, &loss, mxGetDoubles(g), &recon), t.h, "gpe_adjoint_grad");
      set_out(nlhs, plhs, 0, mxCreateDoubleScalar(loss));
      set_out(nlhs, plhs, 1, g);
      set_out(nlhs, plhs, 2, mxCreateDoubleScalar(recon));
    }
  } else if (cmd == "rabi_forward" || cmd == "rabi_grad") {
    // two-component gradient (no reference counterpart), Omega(x) = omega0 exp(i k x):
    //   [loss, psi1, psi2] = gpe_mex('rabi_forward', h, ddt, nsteps, omega0, k)
    //   [loss, grad] = gpe_mex('rabi_grad', h, ddt, nsteps, omega0, k, tape_gib)
    need(nrhs, cmd == "rabi_grad" ? 7 : 6, c);
    Task& t = task_of(prhs[1]);
    gpe_rabi_args a{};
    a.ddt = mxGetScalar(prhs[2]);
    a.nsteps = as_int(prhs[3]);
    a.checkpoint_blocks = 0;
    a.omega0 = mxGetScalar(prhs[4]);
    a.k = mxGetScalar(prhs[5]);
    a.tape_gib = cmd == "rabi_grad" ? mxGetScalar(prhs[6]) : 0.0;
    double loss = 0;
    if (cmd == "rabi_forward") {
      mxArray* both = grid_complex(t, 2);   // component 1 then component 2
      check(gpe_rabi_forward(t.h, &a, &loss, cdata(both)), t.h, "gpe_rabi_forward");
      const size_t n = t.npts();
      mxArray* p1 = grid_complex(t);
      mxArray* p2 = grid_complex(t);
      std::memcpy(cdata(p1), cdata(both), 2 * n * sizeof(double));
      std::memcpy(cdata(p2), cdata(both) + 2 * n, 2 * n * sizeof(double));
      mxDestroyArray(both);
      set_out(nlhs, plhs, 0, mxCreateDoubleScalar(loss));
      set_out(nlhs, plhs, 1, p1);
      set_out(nlhs, plhs, 2, p2);
    } else {
      double g[2] = {0, 0};
      check(gpe_rabi_grad(t.h, &a, &loss, g), t.h, "gpe_rabi_grad");
      set_out(nlhs, plhs, 0, mxCreateDoubleScalar(loss));
      set_out(nlhs, plhs, 1, column(g, 2));
    }
  } else if (cmd == "batch_create") {
    // ensemble of 2D tasks (no reference counterpart):
    // b = gpe_mex('batch_create', n (ntraj x 2: [nx ny]), xmax (ntraj x 2: [xmax ymax]))
    need(nrhs, 3, c);
    const size_t nt = mxGetM(prhs[1]);
    if (mxGetN(prhs[1]) != 2 || mxGetM(prhs[2]) != nt || mxGetN(prhs[2]) != 2)
      mexErrMsgIdAndTxt("gpe:args", "batch_create: ntraj x 2 sizes and extents expected");
    const double* ns = reals(prhs[1], 2 * nt, "n");
    const double* xs = reals(prhs[2], 2 * nt, "xmax");
    std::vector<gpe_grid_desc> d(nt);
    Batch B;
    for (size_t i = 0; i < nt; ++i) {
      d[i].ndims = 2;
      d[i].n[0] = (int)ns[i];
      d[i].n[1] = (int)ns[i + nt];
      d[i].xmax[0] = xs[i];
      d[i].xmax[1] = xs[i + nt];
      B.n.push_back({d[i].n[0], d[i].n[1]});
    }
    checkb(gpe_batch_create((int)nt, d.data(), &B.b), nullptr, "gpe_batch_create");
    const uint64_t id = next_id++;
    batches[id] = B;
    set_out(nlhs, plhs, 0, new_id(id));
  } else if (cmd == "batch_destroy") {
    need(nrhs, 2, c);
    auto it = batches.find(id_of(prhs[1]));
    if (it != batches.end()) {
      gpe_batch_destroy(it->second.b);
      batches.erase(it);
    }
  } else if (cmd == "batch_tune") {
    need(nrhs, 2, c);
    Batch& B = batch_of(prhs[1]);
    int npairs = 0;
    std::vector<double> us(64);
    checkb(gpe_batch_tune(B.b, 0, us.data(), (int)us.size(), &npairs), B.b, "gpe_batch_tune");
    set_out(nlhs, plhs, 0, column(us.data(), 2 * npairs));
  } else if (cmd == "batch_set") {
    // gpe_mex('batch_set', b, i (1-based), g, gamma, decay_rate, Ntotal, mu_init, n_recalc, V, psi)
    need(nrhs, 11, c);
    Batch& B = batch_of(prhs[1]);
    const int i = as_int(prhs[2]) - 1;
    if (i < 0 || i >= (int)B.n.size()) mexErrMsgIdAndTxt("gpe:args", "batch_set: trajectory index out of range");
    const size_t n = (size_t)B.n[i][0] * B.n[i][1];
    gpe_params p{};
    p.ncomp = 1;
    p.g[0] = mxGetScalar(prhs[3]);
    p.gamma = mxGetScalar(prhs[4]);
    p.decay_rate = mxGetScalar(prhs[5]);
    p.Ntotal = mxGetScalar(prhs[6]);
    p.mu_init = mxGetScalar(prhs[7]);
    p.n_recalc = as_int(prhs[8]);
    std::vector<double> tmp;
    checkb(gpe_batch_set(B.b, i, &p, reals(prhs[9], n, "V"), complexes(prhs[10], n, tmp, "psi")), B.b,
           "gpe_batch_set");
  } else if (cmd == "batch_get_state") {
    // psi = gpe_mex('batch_get_state', b, i (1-based))
    need(nrhs, 3, c);
    Batch& B = batch_of(prhs[1]);
    const int i = as_int(prhs[2]) - 1;
    if (i < 0 || i >= (int)B.n.size()) mexErrMsgIdAndTxt("gpe:args", "batch_get_state: index out of range");
    mxArray* out = mxCreateNumericMatrix((mwSize)B.n[i][1], (mwSize)B.n[i][0], mxDOUBLE_CLASS, mxCOMPLEX);
    checkb(gpe_batch_get_state(B.b, i, cdata(out)), B.b, "gpe_batch_get_state");
    set_out(nlhs, plhs, 0, out);
  } else if (cmd == "batch_split_begin") {
    // muc = gpe_mex('batch_split_begin', b, ddt, niter_inner, use_graph)
    need(nrhs, 5, c);
    Batch& B = batch_of(prhs[1]);
    std::vector<double> muc(B.n.size());
    checkb(gpe_batch_split_begin(B.b, mxGetScalar(prhs[2]), as_int(prhs[3]), as_int(prhs[4]), muc.data()), B.b,
           "gpe_batch_split_begin");
    set_out(nlhs, plhs, 0, column(muc.data(), (int)muc.size()));
  } else if (cmd == "batch_split_outer") {
    // [time2, muc, ncur] = gpe_mex('batch_split_outer', b, j)
    need(nrhs, 3, c);
    Batch& B = batch_of(prhs[1]);
    std::vector<double> muc(B.n.size()), ncur(B.n.size());
    double t2 = 0;
    checkb(gpe_batch_split_outer(B.b, as_int(prhs[2]), &t2, muc.data(), ncur.data()), B.b,
           "gpe_batch_split_outer");
    set_out(nlhs, plhs, 0, mxCreateDoubleScalar(t2));
    set_out(nlhs, plhs, 1, column(muc.data(), (int)muc.size()));
    set_out(nlhs, plhs, 2, column(ncur.data(), (int)ncur.size()));
  } else if (cmd == "device_info") {
    char info[256];
    check(gpe_device_info(info, (int)sizeof info), nullptr, "gpe_device_info");
    set_out(nlhs, plhs, 0, mxCreateString(info));
  } else {
    mexErrMsgIdAndTxt("gpe:cmd", "gpe_mex: unknown command '%s'", c);
  }
}
