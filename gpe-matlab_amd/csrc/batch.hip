// Batched ragged ensemble driver: many independent 2D GPEtask.solve_split runs
// (different grid sizes, potentials, g, gamma) advanced together.
//
// Every sweep of a block is one launch per distinct pencil length across the
// trajectories of a chain (work items = (trajectory, pencil tile)), so a block of
// n_rec steps costs 2 n_rec + 5 launches per length group instead of per
// trajectory; chains (trajectories of different grid shapes) run on up to four
// streams at once.
// With decay_rate = 0 everywhere, one whole outer iteration (niter_inner steps)
// is captured once into a hipGraph and replayed.  Per-trajectory scalars (mu,
// renormalisation, ncur) live in a DevState array and are reduced on the device.
#include <algorithm>
#include <cmath>
#include <map>

#include "batch.h"
#include "finalize_dev.h"
#include "handle.h"

using namespace gpe_detail;

namespace gpe {

__global__ void __launch_bounds__(256) batch_finalize_kernel(const TrajDev* traj, const int* list, int mode,
                                                             double time2) {
  __shared__ double red[4];
  const TrajDev& tj = traj[list[blockIdx.x]];
  FinParams f{};
  f.partials = tj.partials;
  f.maxwg = tj.maxwg;
  f.cnt[SL_A] = f.cnt[SL_B] = f.cnt[SL_C] = f.cnt[SL_KX] = tj.tiles_x;
  f.cnt[SL_KY] = tj.tiles_y;
  f.st = tj.st;
  f.mode = mode;
  f.renorm = tj.renorm;
  f.w = tj.w;
  f.g = tj.g;
  f.NN0 = tj.NN0;
  f.tau = tj.tau;
  f.time2 = time2;
  finalize_core(f, red);
}

hipError_t launch_batch_finalize(const TrajDev* traj, const int* list, int ntraj, int mode, double time2,
                                 hipStream_t s) {
  hipLaunchKernelGGL(batch_finalize_kernel, dim3((unsigned)ntraj), dim3(256), 0, s, traj, list, mode, time2);
  return hipGetLastError();
}

#define GPE_DECL(N) hipError_t launch_batch_sweep_##N(bool, int, int, const BatchLaunch&, hipStream_t);
GPE_DECL(8) GPE_DECL(12) GPE_DECL(16) GPE_DECL(24) GPE_DECL(32) GPE_DECL(48) GPE_DECL(64) GPE_DECL(96)
GPE_DECL(128) GPE_DECL(192) GPE_DECL(256) GPE_DECL(384) GPE_DECL(512)
#undef GPE_DECL

hipError_t launch_batch_sweep(int n, bool xaxis, int seq, int var, const BatchLaunch& b, hipStream_t s) {
  switch (n) {
#define GPE_C(N) case N: return launch_batch_sweep_##N(xaxis, seq, var, b, s);
    GPE_C(8) GPE_C(12) GPE_C(16) GPE_C(24) GPE_C(32) GPE_C(48) GPE_C(64) GPE_C(96) GPE_C(128) GPE_C(192)
    GPE_C(256) GPE_C(384) GPE_C(512)
#undef GPE_C
    default:
      return hipErrorInvalidValue;
  }
}

}  // namespace gpe

struct gpe_batch {
  struct Traj {
    int nx = 0, ny = 0;
    double dx = 0, dy = 0, w = 0;
    std::vector<double> kx, ky;
    gpe_params prm{};
    bool set = false;
    size_t off = 0, koff_x = 0, koff_y = 0;
    int tiles_x = 0, tiles_y = 0;
  };
  struct Group {
    int n = 0;
    int var = 0;       // plan variant of the group's sweeps
    size_t first = 0;  // offset into the device item array
    int count = 0;
  };
  // Trajectories of different shapes never interact, so they advance on up to four
  // streams at once: the launches of one shape are small and would leave the GPU
  // partly idle between them.  Each chain owns its trajectories' work items.
  struct Chain {
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
    std::vector<Group> gx, gy;
    std::vector<int> traj;
    int* d_list = nullptr;   // traj on the device (finalize)
  };
  int device = 0;
  hipStream_t stream = nullptr;
  std::vector<Traj> tr;
  size_t total = 0, ktotal_x = 0, ktotal_y = 0;
  double2* psi = nullptr;
  double* V = nullptr;
  double2* kcol = nullptr;
  double2* kax = nullptr;
  double* keax_x = nullptr;
  double* keax_y = nullptr;
  std::map<int, double2*> tw;
  DevState* st = nullptr;
  double* partials = nullptr;
  int maxwg = 0;
  gpe::TrajDev* d_traj = nullptr;
  std::vector<gpe::TrajDev> h_traj;
  gpe::WorkItem* d_items = nullptr;
  size_t items_cap = 0;
  std::map<int, int> var_x, var_y;   // plan variant per pencil length of the x / y sweeps
  std::vector<Chain> ch;
  hipEvent_t fork = nullptr;
  double ddt = 0;
  int niter_inner = 0, n_rec = 1;
  bool begun = false;
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  std::string err;
};

namespace {

int bfail(gpe_batch* b, int code, const char* fmt, ...) {
  if (b) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    b->err = buf;
  }
  return code;
}

#define BCHK(b, x)                                                                            \
  do {                                                                                        \
    hipError_t e_ = (x);                                                                      \
    if (e_ != hipSuccess) return bfail((b), GPE_EHIP, "%s: %s", #x, hipGetErrorString(e_));   \
  } while (0)

int bsweep(gpe_batch* b, const gpe_batch::Chain& c, bool xaxis, int seq) {
  for (const auto& g : (xaxis ? c.gx : c.gy)) {
    gpe::BatchLaunch L{b->d_traj, b->d_items + g.first, g.count};
    hipError_t e = gpe::launch_batch_sweep(g.n, xaxis, seq, g.var, L, c.stream);
    if (e != hipSuccess) return bfail(b, GPE_EHIP, "batch sweep (n %d, seq %d): %s", g.n, seq, hipGetErrorString(e));
  }
  return GPE_OK;
}

int bmeasure(gpe_batch* b, const gpe_batch::Chain& c, int mode, double time2) {
  int rc;
  if ((rc = bsweep(b, c, true, gpe::S_KEX))) return rc;
  if ((rc = bsweep(b, c, false, gpe::S_KE))) return rc;
  BCHK(b, gpe::launch_batch_finalize(b->d_traj, c.d_list, (int)c.traj.size(), mode, time2, c.stream));
  return GPE_OK;
}

// solve_split.m:53-78 for every trajectory of one chain
int bblock(gpe_batch* b, const gpe_batch::Chain& c, double time2) {
  int rc;
  for (int i = 0; i < b->n_rec; ++i) {
    if ((rc = bsweep(b, c, true, i == 0 ? gpe::S_OPEN : gpe::S_MID))) return rc;
    if ((rc = bsweep(b, c, false, gpe::S_KIN))) return rc;
  }
  if ((rc = bsweep(b, c, true, gpe::S_CLOSE))) return rc;
  return bmeasure(b, c, gpe::FIN_BLOCK, time2);
}

// Runs fn(chain) on every chain's stream between a fork from and a join into b->stream
// (also inside a graph capture of b->stream, where it records parallel branches).
template <class F>
int on_chains(gpe_batch* b, F fn) {
  BCHK(b, hipEventRecord(b->fork, b->stream));
  for (auto& c : b->ch) {
    BCHK(b, hipStreamWaitEvent(c.stream, b->fork, 0));
    int rc;
    if ((rc = fn(c))) return rc;
    BCHK(b, hipEventRecord(c.done, c.stream));
  }
  for (auto& c : b->ch) BCHK(b, hipStreamWaitEvent(b->stream, c.done, 0));
  return GPE_OK;
}

int bouter(gpe_batch* b, int j) {
  const double dt_outer = b->ddt * b->niter_inner;
  const double time = (double)(j - 1) * dt_outer;
  return on_chains(b, [&](const gpe_batch::Chain& c) {
    int rc;
    for (int jj = 1; jj <= b->niter_inner / b->n_rec; ++jj) {
      const double time2 = time + (double)(jj - 1) * b->ddt * (double)b->n_rec;
      if ((rc = bblock(b, c, time2))) return rc;
    }
    return GPE_OK;
  });
}

// Work-group tiles of every trajectory under the current plan variants, the chains'
// length groups and the device item array.
int build_items(gpe_batch* b) {
  for (auto& t : b->tr) {
    const int vx = b->var_x[t.nx], vy = b->var_y[t.ny];
    // x sweeps are phase sweeps (open / mid / close / kex)
    const int cx = gpe::sweep_cols_per_block_phase(t.nx, vx), cy = gpe::sweep_cols_per_block(t.ny, vy);
    t.tiles_x = (t.ny + cx - 1) / cx;
    t.tiles_y = (t.nx + cy - 1) / cy;
  }
  std::vector<gpe::WorkItem> items;
  for (auto& c : b->ch) {
    c.gx.clear();
    c.gy.clear();
    for (int axis = 0; axis < 2; ++axis) {
      std::map<int, std::vector<gpe::WorkItem>> by_n;
      for (int i : c.traj) {
        const auto& t = b->tr[i];
        const int n = axis == 0 ? t.nx : t.ny;
        const int tiles = axis == 0 ? t.tiles_x : t.tiles_y;
        for (int k = 0; k < tiles; ++k) by_n[n].push_back({i, k});
      }
      for (auto& kv : by_n) {
        gpe_batch::Group g;
        g.n = kv.first;
        g.var = axis == 0 ? b->var_x[g.n] : b->var_y[g.n];
        g.first = items.size();
        g.count = (int)kv.second.size();
        items.insert(items.end(), kv.second.begin(), kv.second.end());
        (axis == 0 ? c.gx : c.gy).push_back(g);
      }
    }
  }
  if (items.size() > b->items_cap) {
    (void)hipFree(b->d_items);
    b->d_items = nullptr;
    b->items_cap = 0;
    if (hipMalloc(&b->d_items, items.size() * sizeof(gpe::WorkItem)) != hipSuccess)
      return bfail(b, GPE_ENOMEM, "work items");
    b->items_cap = items.size();
  }
  BCHK(b, hipMemcpy(b->d_items, items.data(), items.size() * sizeof(gpe::WorkItem), hipMemcpyHostToDevice));
  return GPE_OK;
}

// Device view of trajectory i (pointers, sizes, tiles); the step parameters (g, dt,
// renormalisation) come from its gpe_params when set.
gpe::TrajDev traj_dev(gpe_batch* b, int i, std::complex<double> dt) {
  const auto& t = b->tr[i];
  gpe::TrajDev d{};
  d.psi = b->psi + t.off;
  d.V = b->V + t.off;
  d.kcol = b->kcol + t.koff_x;
  d.kax = b->kax + t.koff_y;
  d.keax_x = b->keax_x + t.koff_x;
  d.keax_y = b->keax_y + t.koff_y;
  d.tw_x = b->tw[t.nx];
  d.tw_y = b->tw[t.ny];
  d.st = b->st + i;
  d.partials = b->partials + (size_t)i * gpe::SL_NSLOT * b->maxwg;
  d.nx = (unsigned)t.nx;
  d.ny = (unsigned)t.ny;
  d.maxwg = b->maxwg;
  d.tiles_x = t.tiles_x;
  d.tiles_y = t.tiles_y;
  d.renorm = (t.prm.gamma > 0 && t.prm.mu_init == 0) ? 1 : 0;
  d.g = t.prm.g[0];
  d.dr = dt.real();
  d.di = dt.imag();
  d.w = t.w;
  d.NN0 = t.prm.Ntotal;
  d.tau = t.prm.decay_rate;
  return d;
}

void drop_graph(gpe_batch* b) {
  if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->gexec = nullptr;
  b->graph = nullptr;
}

}  // namespace

extern "C" {

const char* gpe_batch_last_error(const gpe_batch* b) { return b ? b->err.c_str() : "null batch"; }

int gpe_batch_destroy(gpe_batch* b) {
  if (!b) return GPE_OK;
  (void)hipSetDevice(b->device);
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  drop_graph(b);
  (void)hipFree(b->psi);
  (void)hipFree(b->V);
  (void)hipFree(b->kcol);
  (void)hipFree(b->kax);
  (void)hipFree(b->keax_x);
  (void)hipFree(b->keax_y);
  for (auto& kv : b->tw) (void)hipFree(kv.second);
  (void)hipFree(b->st);
  (void)hipFree(b->partials);
  (void)hipFree(b->d_traj);
  (void)hipFree(b->d_items);
  for (auto& c : b->ch) {
    if (c.stream) {
      (void)hipStreamSynchronize(c.stream);
      (void)hipStreamDestroy(c.stream);
    }
    if (c.done) (void)hipEventDestroy(c.done);
    (void)hipFree(c.d_list);
  }
  if (b->fork) (void)hipEventDestroy(b->fork);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
  return GPE_OK;
}

int gpe_batch_create(int ntraj, const gpe_grid_desc* grids, gpe_batch** out) {
  if (!out || !grids || ntraj < 1) return GPE_EINVAL;
  *out = nullptr;
  for (int i = 0; i < ntraj; ++i) {
    const gpe_grid_desc& d = grids[i];
    if (d.ndims != 2 || !gpe::sweep_length_supported(d.n[0]) || !gpe::sweep_length_supported(d.n[1]) ||
        !(d.xmax[0] > 0) || !(d.xmax[1] > 0))
      return GPE_EINVAL;
  }
  gpe_batch* b = new gpe_batch();
  const char* dev = getenv("GPE_DEVICE");
  b->device = dev ? atoi(dev) : 0;
  if (hipSetDevice(b->device) != hipSuccess) {
    delete b;
    return GPE_EHIP;
  }
  b->tr.resize(ntraj);
  int maxwg = 1;
  for (int i = 0; i < ntraj; ++i) {
    auto& t = b->tr[i];
    t.nx = grids[i].n[0];
    t.ny = grids[i].n[1];
    std::vector<double> x = linspace(-grids[i].xmax[0], grids[i].xmax[0], t.nx);
    std::vector<double> y = linspace(-grids[i].xmax[1], grids[i].xmax[1], t.ny);
    t.dx = x[1] - x[0];
    t.dy = y[1] - y[0];
    t.w = t.dx * t.dy;
    t.kx = kvec(x, t.dx);
    t.ky = kvec(y, t.dy);
    t.off = b->total;
    b->total += (size_t)t.nx * t.ny;
    t.koff_x = b->ktotal_x;
    b->ktotal_x += t.nx;
    t.koff_y = b->ktotal_y;
    b->ktotal_y += t.ny;
    // partials fit every plan variant
    const int cx = gpe::sweep_cols_min(t.nx), cy = gpe::sweep_cols_min(t.ny);
    maxwg = std::max(maxwg, std::max((t.ny + cx - 1) / cx, (t.nx + cy - 1) / cy));
    for (int a = 0; a < 2; ++a) {
      const int n = a == 0 ? t.nx : t.ny;
      const int w = gpe::wisdom_get(2, a, n);
      (a == 0 ? b->var_x : b->var_y)[n] = w > 0 && w < gpe::plan_variants(n) ? w : 0;
    }
  }
  b->maxwg = maxwg;
  // chains: one per grid shape, the smallest shapes (by points) merged until at most
  // four remain; in a chain, work items grouped by swept length
  std::map<std::pair<int, int>, std::vector<int>> by_shape;
  for (int i = 0; i < ntraj; ++i) by_shape[{b->tr[i].nx, b->tr[i].ny}].push_back(i);
  std::vector<std::pair<double, std::vector<int>>> bins;
  double total_pts = 0;
  for (auto& kv : by_shape) {
    bins.push_back({(double)kv.first.first * kv.first.second * kv.second.size(), kv.second});
    total_pts += bins.back().first;
  }
  // a shape holding well over a quarter of the points is cut into chains of about a
  // quarter each, so no single chain is the critical path
  {
    std::vector<std::pair<double, std::vector<int>>> cut;
    for (auto& bn : bins) {
      const int cnt = (int)bn.second.size();
      const int k = std::max(1, std::min(cnt, (int)std::floor(bn.first / (total_pts / 4.0) + 0.5)));
      const double per = bn.first / cnt;
      for (int q = 0; q < k; ++q) {
        const int lo = (int)((long long)cnt * q / k), hi = (int)((long long)cnt * (q + 1) / k);
        cut.push_back({per * (hi - lo), std::vector<int>(bn.second.begin() + lo, bn.second.begin() + hi)});
      }
    }
    bins.swap(cut);
  }
  std::sort(bins.begin(), bins.end(), [](const auto& a, const auto& c) { return a.first > c.first; });
  while (bins.size() > 4) {
    auto last = bins.back();
    bins.pop_back();
    auto& into = bins.back();   // the smallest remaining chain takes the smallest shape
    into.first += last.first;
    into.second.insert(into.second.end(), last.second.begin(), last.second.end());
    std::sort(into.second.begin(), into.second.end());
    std::sort(bins.begin(), bins.end(), [](const auto& a, const auto& c) { return a.first > c.first; });
  }
  b->ch.resize(bins.size());
  for (size_t ci = 0; ci < bins.size(); ++ci) b->ch[ci].traj = bins[ci].second;
  auto bail = [&](int code) {
    gpe_batch_destroy(b);
    return code;
  };
  if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) return bail(GPE_EHIP);
  if (hipEventCreateWithFlags(&b->fork, hipEventDisableTiming) != hipSuccess) return bail(GPE_EHIP);
  for (auto& c : b->ch) {
    if (hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking) != hipSuccess) return bail(GPE_EHIP);
    if (hipEventCreateWithFlags(&c.done, hipEventDisableTiming) != hipSuccess) return bail(GPE_EHIP);
    if (hipMalloc(&c.d_list, c.traj.size() * sizeof(int)) != hipSuccess) return bail(GPE_ENOMEM);
    if (hipMemcpy(c.d_list, c.traj.data(), c.traj.size() * sizeof(int), hipMemcpyHostToDevice) != hipSuccess)
      return bail(GPE_EHIP);
  }
  if (hipMalloc(&b->psi, b->total * sizeof(double2)) != hipSuccess) return bail(GPE_ENOMEM);
  if (hipMalloc(&b->V, b->total * sizeof(double)) != hipSuccess) return bail(GPE_ENOMEM);
  if (hipMalloc(&b->kcol, b->ktotal_x * sizeof(double2)) != hipSuccess) return bail(GPE_ENOMEM);
  if (hipMalloc(&b->kax, b->ktotal_y * sizeof(double2)) != hipSuccess) return bail(GPE_ENOMEM);
  if (hipMalloc(&b->keax_x, b->ktotal_x * sizeof(double)) != hipSuccess) return bail(GPE_ENOMEM);
  if (hipMalloc(&b->keax_y, b->ktotal_y * sizeof(double)) != hipSuccess) return bail(GPE_ENOMEM);
  if (hipMalloc(&b->st, ntraj * sizeof(DevState)) != hipSuccess) return bail(GPE_ENOMEM);
  if (hipMalloc(&b->partials, (size_t)ntraj * gpe::SL_NSLOT * maxwg * sizeof(double)) != hipSuccess)
    return bail(GPE_ENOMEM);
  if (hipMalloc(&b->d_traj, ntraj * sizeof(gpe::TrajDev)) != hipSuccess) return bail(GPE_ENOMEM);
  if (build_items(b) != GPE_OK) return bail(GPE_ENOMEM);
  std::vector<double> kex(b->ktotal_x), key(b->ktotal_y);
  for (auto& t : b->tr) {
    for (int n : {t.nx, t.ny})
      if (!b->tw.count(n)) {
        double2* p = nullptr;
        std::vector<double2> tw = twiddles(n);
        if (hipMalloc(&p, n * sizeof(double2)) != hipSuccess) return bail(GPE_ENOMEM);
        b->tw[n] = p;
        if (hipMemcpy(p, tw.data(), n * sizeof(double2), hipMemcpyHostToDevice) != hipSuccess) return bail(GPE_EHIP);
      }
    for (int i = 0; i < t.nx; ++i) kex[t.koff_x + i] = t.kx[i] * t.kx[i] / 2.0 / (double)t.nx;
    for (int i = 0; i < t.ny; ++i) key[t.koff_y + i] = t.ky[i] * t.ky[i] / 2.0 / (double)t.ny;
  }
  if (hipMemcpy(b->keax_x, kex.data(), kex.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(b->keax_y, key.data(), key.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
    return bail(GPE_EHIP);
  std::vector<DevState> s0(ntraj);
  for (auto& s : s0) {
    s = DevState{};
    s.scale = 1.0;
    s.mu_itp = 1.0;
  }
  if (hipMemcpy(b->st, s0.data(), ntraj * sizeof(DevState), hipMemcpyHostToDevice) != hipSuccess) return bail(GPE_EHIP);
  b->h_traj.resize(ntraj);
  *out = b;
  return GPE_OK;
}

int gpe_batch_set(gpe_batch* b, int i, const gpe_params* p, const double* V, const double* psi) {
  if (!b || !p || i < 0 || i >= (int)b->tr.size()) return b ? bfail(b, GPE_EINVAL, "bad trajectory") : GPE_EINVAL;
  if (p->ncomp != 1 || p->n_recalc < 1) return bfail(b, GPE_EINVAL, "one component, n_recalc >= 1");
  BCHK(b, hipSetDevice(b->device));
  auto& t = b->tr[i];
  const size_t n = (size_t)t.nx * t.ny;
  if (V) BCHK(b, hipMemcpyAsync(b->V + t.off, V, n * sizeof(double), hipMemcpyHostToDevice, b->stream));
  if (psi) BCHK(b, hipMemcpyAsync(b->psi + t.off, psi, n * sizeof(double2), hipMemcpyHostToDevice, b->stream));
  BCHK(b, hipStreamSynchronize(b->stream));
  t.prm = *p;
  t.set = true;
  b->begun = false;
  drop_graph(b);
  if (psi) {
    DevState s;
    BCHK(b, hipMemcpy(&s, b->st + i, sizeof s, hipMemcpyDeviceToHost));
    s.scale = 1.0;
    BCHK(b, hipMemcpy(b->st + i, &s, sizeof s, hipMemcpyHostToDevice));
  }
  return GPE_OK;
}

int gpe_batch_get_state(gpe_batch* b, int i, double* psi) {
  if (!b || !psi || i < 0 || i >= (int)b->tr.size()) return b ? bfail(b, GPE_EINVAL, "bad trajectory") : GPE_EINVAL;
  BCHK(b, hipSetDevice(b->device));
  auto& t = b->tr[i];
  const size_t n = (size_t)t.nx * t.ny;
  DevState s;
  BCHK(b, hipMemcpyAsync(&s, b->st + i, sizeof s, hipMemcpyDeviceToHost, b->stream));
  BCHK(b, hipMemcpyAsync(psi, b->psi + t.off, n * sizeof(double2), hipMemcpyDeviceToHost, b->stream));
  BCHK(b, hipStreamSynchronize(b->stream));
  if (s.scale != 1.0)
    for (size_t k = 0; k < 2 * n; ++k) psi[k] = psi[k] * s.scale;
  return GPE_OK;
}

int gpe_batch_split_begin(gpe_batch* b, double ddt, int niter_inner, int use_graph, double* muc_out) {
  if (!b || niter_inner < 1) return b ? bfail(b, GPE_EINVAL, "niter_inner must be >= 1") : GPE_EINVAL;
  BCHK(b, hipSetDevice(b->device));
  const int ntraj = (int)b->tr.size();
  const int nrec0 = b->tr[0].prm.n_recalc;
  bool decay = false;
  for (auto& t : b->tr) {
    if (!t.set) return bfail(b, GPE_ESTATE, "every trajectory needs gpe_batch_set");
    if (t.prm.n_recalc != nrec0) return bfail(b, GPE_EINVAL, "all trajectories need the same n_recalc");
    if (!(t.prm.Ntotal > 0)) return bfail(b, GPE_EINVAL, "batched trajectories need Ntotal > 0");
    decay = decay || t.prm.decay_rate > 0;
  }
  drop_graph(b);
  // materialise pending renormalisations of a previous run
  std::vector<DevState> s(ntraj);
  BCHK(b, hipMemcpy(s.data(), b->st, ntraj * sizeof(DevState), hipMemcpyDeviceToHost));
  for (int i = 0; i < ntraj; ++i)
    if (s[i].scale != 1.0) {
      auto& t = b->tr[i];
      const size_t n = (size_t)t.nx * t.ny;
      std::vector<double2> h(n);
      BCHK(b, hipMemcpy(h.data(), b->psi + t.off, n * sizeof(double2), hipMemcpyDeviceToHost));
      for (auto& v : h) v = make_double2(v.x * s[i].scale, v.y * s[i].scale);
      BCHK(b, hipMemcpy(b->psi + t.off, h.data(), n * sizeof(double2), hipMemcpyHostToDevice));
      s[i].scale = 1.0;
    }
  b->n_rec = niter_inner % nrec0 ? 1 : nrec0;                   // solve_split.m:24-27
  b->niter_inner = niter_inner;
  b->ddt = ddt;
  std::vector<double2> kc(b->ktotal_x), ka(b->ktotal_y);
  for (int i = 0; i < ntraj; ++i) {
    auto& t = b->tr[i];
    const std::complex<double> dt = cdiv({0.0, ddt}, {1.0, t.prm.gamma});   // solve_split.m:31
    for (int k = 0; k < t.nx; ++k) {
      const std::complex<double> e = std::exp(-(t.kx[k] * t.kx[k] / 2) * dt);
      kc[t.koff_x + k] = make_double2(e.real(), e.imag());
    }
    const double inv = 1.0 / ((double)t.nx * t.ny);
    for (int k = 0; k < t.ny; ++k) {
      const std::complex<double> e = std::exp(-(t.ky[k] * t.ky[k] / 2) * dt) * inv;
      ka[t.koff_y + k] = make_double2(e.real(), e.imag());
    }
    b->h_traj[i] = traj_dev(b, i, dt);
  }
  BCHK(b, hipMemcpy(b->kcol, kc.data(), kc.size() * sizeof(double2), hipMemcpyHostToDevice));
  BCHK(b, hipMemcpy(b->kax, ka.data(), ka.size() * sizeof(double2), hipMemcpyHostToDevice));
  BCHK(b, hipMemcpy(b->d_traj, b->h_traj.data(), ntraj * sizeof(gpe::TrajDev), hipMemcpyHostToDevice));
  BCHK(b, hipMemcpy(b->st, s.data(), ntraj * sizeof(DevState), hipMemcpyHostToDevice));
  // solve_split.m:40-46: initial muc of every trajectory
  int rc;
  if ((rc = on_chains(b, [&](const gpe_batch::Chain& c) { return bmeasure(b, c, gpe::FIN_SUMS, 0.0); }))) return rc;
  BCHK(b, hipStreamSynchronize(b->stream));
  BCHK(b, hipMemcpy(s.data(), b->st, ntraj * sizeof(DevState), hipMemcpyDeviceToHost));
  for (int i = 0; i < ntraj; ++i) {
    const auto& t = b->tr[i];
    // same association as gpe_split_begin: ((sum V n w + kinetic w) + g sum n^2 w) / NN0
    const double m1 = s[i].sB * t.w, m2 = s[i].sC * t.w, m3 = s[i].sK * t.w;
    const double muc = ((m1 + m3) + t.prm.g[0] * m2) / t.prm.Ntotal;
    s[i].mu = t.prm.mu_init > 0 ? t.prm.mu_init : muc;
    s[i].scale = 1.0;
    s[i].NNN = t.prm.Ntotal;
    s[i].muc = muc;
    s[i].ncur = t.prm.Ntotal;
    if (muc_out) muc_out[i] = muc;
  }
  BCHK(b, hipMemcpy(b->st, s.data(), ntraj * sizeof(DevState), hipMemcpyHostToDevice));
  if (use_graph && !decay) {
    // capture one outer iteration; block times only matter when decay_rate > 0
    BCHK(b, hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    rc = bouter(b, 1);
    hipGraph_t g = nullptr;
    const hipError_t ec = hipStreamEndCapture(b->stream, &g);
    if (rc) {
      if (g) (void)hipGraphDestroy(g);
      return rc;
    }
    if (ec != hipSuccess) return bfail(b, GPE_EHIP, "graph capture: %s", hipGetErrorString(ec));
    b->graph = g;
    BCHK(b, hipGraphInstantiate(&b->gexec, g, nullptr, nullptr, 0));
  }
  b->begun = true;
  return GPE_OK;
}

int gpe_batch_split_outer(gpe_batch* b, int j, double* time2_out, double* muc, double* ncur) {
  if (!b) return GPE_EINVAL;
  if (!b->begun) return bfail(b, GPE_ESTATE, "gpe_batch_split_begin has not been called");
  BCHK(b, hipSetDevice(b->device));
  int rc;
  if (b->gexec) {
    BCHK(b, hipGraphLaunch(b->gexec, b->stream));
  } else if ((rc = bouter(b, j))) {
    return rc;
  }
  const int ntraj = (int)b->tr.size();
  std::vector<DevState> s(ntraj);
  BCHK(b, hipMemcpyAsync(s.data(), b->st, ntraj * sizeof(DevState), hipMemcpyDeviceToHost, b->stream));
  BCHK(b, hipStreamSynchronize(b->stream));
  const double time = (double)(j - 1) * (b->ddt * b->niter_inner);
  if (time2_out) *time2_out = time + (double)(b->niter_inner / b->n_rec - 1) * b->ddt * (double)b->n_rec;
  for (int i = 0; i < ntraj; ++i) {
    if (muc) muc[i] = s[i].muc;
    if (ncur) ncur[i] = s[i].ncur;
  }
  // the block end divides by each trajectory's norm (solve_split.m:67-73)
  for (int i = 0; i < ntraj; ++i)
    if (!std::isfinite(s[i].ncur) || !std::isfinite(s[i].muc) || !std::isfinite(s[i].scale) || !std::isfinite(s[i].mu))
      return bfail(b, GPE_ENONFINITE, "trajectory %d: block reduction is not finite (ncur %g, mu %g)", i, s[i].ncur,
                   s[i].muc);
  return GPE_OK;
}

// The batch counterpart of gpe_tune: the (axis, length) pairs of the ensemble are
// tuned one after another (coordinate descent), each by timing the whole ensemble's
// steady-state step (x sweeps S_MID, then y sweeps S_KIN, one launch per length) with
// that pair at each variant and the other pairs at their current choice.
int gpe_batch_tune(gpe_batch* b, int remeasure, double* times_us, int cap, int* npairs) {
  if (!b) return GPE_EINVAL;
  BCHK(b, hipSetDevice(b->device));
  std::vector<std::pair<int, int>> pairs;   // (axis, n), ascending
  for (int a = 0; a < 2; ++a) {
    std::map<int, int> seen;
    for (auto& t : b->tr) seen[a == 0 ? t.nx : t.ny] = 1;
    for (auto& kv : seen) pairs.push_back({a, kv.first});
  }
  if (npairs) *npairs = (int)pairs.size();
  if (times_us)
    for (int i = 0; i < std::min(cap, 2 * (int)pairs.size()); ++i) times_us[i] = std::nan("");
  // current choice of every pair: recorded plan, else variant 0
  std::map<std::pair<int, int>, int> cur;
  std::vector<size_t> todo;
  for (size_t pi = 0; pi < pairs.size(); ++pi) {
    const int w = gpe::wisdom_get(2, pairs[pi].first, pairs[pi].second);
    cur[pairs[pi]] = w > 0 && w < gpe::plan_variants(pairs[pi].second) ? w : 0;
    if (gpe::plan_variants(pairs[pi].second) > 1 && (remeasure || w < 0)) todo.push_back(pi);
  }
  if (!todo.empty()) {
    const int ntraj = (int)b->tr.size();
    BCHK(b, hipStreamSynchronize(b->stream));
    for (int i = 0; i < ntraj; ++i) {
      const auto& t = b->tr[i];
      const std::complex<double> dt = b->ddt != 0 ? cdiv({0.0, b->ddt}, {1.0, t.prm.gamma}) : std::complex<double>(0, 0);
      b->h_traj[i] = traj_dev(b, i, dt);
    }
    BCHK(b, hipMemcpy(b->d_traj, b->h_traj.data(), ntraj * sizeof(gpe::TrajDev), hipMemcpyHostToDevice));
    double2* save = nullptr;
    gpe::WorkItem* items = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = GPE_OK;
    auto chk = [&](hipError_t e, const char* what) {
      if (e != hipSuccess && rc == GPE_OK) rc = bfail(b, GPE_EHIP, "batch tuning (%s): %s", what, hipGetErrorString(e));
      return rc == GPE_OK;
    };
    size_t maxitems = 0;
    for (auto& t : b->tr) {
      const int cx = gpe::sweep_cols_min(t.nx), cy = gpe::sweep_cols_min(t.ny);
      maxitems += (size_t)((t.ny + cx - 1) / cx + (t.nx + cy - 1) / cy);
    }
    // one ensemble step under the variants in `cur`: every x group, then every y group
    struct TG { int axis, n, var; size_t first; int count; };
    auto step_us = [&](double* us) {
      std::vector<gpe::WorkItem> it;
      std::vector<TG> groups;
      for (auto& pr : pairs) {
        const int axis = pr.first, n = pr.second, v = cur[pr];
        TG g{axis, n, v, it.size(), 0};
        for (int i = 0; i < ntraj; ++i) {
          const auto& t = b->tr[i];
          if ((axis == 0 ? t.nx : t.ny) != n) continue;
          const int c = axis == 0 ? gpe::sweep_cols_per_block_phase(n, v) : gpe::sweep_cols_per_block(n, v);
          const int tiles = ((axis == 0 ? t.ny : t.nx) + c - 1) / c;
          for (int k = 0; k < tiles; ++k) it.push_back({i, k});
        }
        g.count = (int)(it.size() - g.first);
        groups.push_back(g);
      }
      if (!chk(hipMemcpy(items, it.data(), it.size() * sizeof(gpe::WorkItem), hipMemcpyHostToDevice), "items"))
        return;
      const int warm = 2, reps = 4;
      for (int r = 0; r < warm + reps && rc == GPE_OK; ++r) {
        if (r == warm && !chk(hipEventRecord(e0, b->stream), "event")) return;
        for (auto& g : groups) {
          const gpe::BatchLaunch L{b->d_traj, items + g.first, g.count};
          if (!chk(gpe::launch_batch_sweep(g.n, g.axis == 0, g.axis == 0 ? gpe::S_MID : gpe::S_KIN, g.var, L,
                                           b->stream), "sweep"))
            return;
        }
      }
      float ms = 0.f;
      if (chk(hipEventRecord(e1, b->stream), "event") && chk(hipEventSynchronize(e1), "sync") &&
          chk(hipEventElapsedTime(&ms, e0, e1), "elapsed"))
        *us = 1e3 * (double)ms / reps;
    };
    // the sweeps run in place: keep every trajectory's state and put it back afterwards
    if (chk(hipMalloc(&save, b->total * sizeof(double2)), "save") &&
        chk(hipMemcpyAsync(save, b->psi, b->total * sizeof(double2), hipMemcpyDeviceToDevice, b->stream), "save") &&
        // timed on a synthetic field of unit modulus and, before gpe_batch_split_begin
        // has built them, unit-modulus kinetic tables (all-zero buffers run faster)
        chk(gpe::launch_tune_fill(b->psi, b->total, nullptr, 0, b->stream), "fill") &&
        (b->begun || (chk(gpe::launch_tune_fill(b->kcol, b->ktotal_x, nullptr, 0, b->stream), "fill") &&
                      chk(gpe::launch_tune_fill(b->kax, b->ktotal_y, nullptr, 0, b->stream), "fill"))) &&
        chk(hipMalloc(&items, maxitems * sizeof(gpe::WorkItem)), "items") && chk(hipEventCreate(&e0), "event") &&
        chk(hipEventCreate(&e1), "event")) {
      for (size_t pi : todo) {
        if (rc) break;
        // rounds alternate the variants; each keeps its best round
        double t_us[2] = {1e300, 1e300};
        for (int rv = 0; rv < 6 && rc == GPE_OK; ++rv) {
          const int v = rv & 1;
          cur[pairs[pi]] = v;
          double us = 1e300;
          step_us(&us);
          t_us[v] = std::min(t_us[v], us);
        }
        if (rc) break;
        if (times_us && 2 * (int)pi + 1 < cap) {
          times_us[2 * pi] = t_us[0];
          times_us[2 * pi + 1] = t_us[1];
        }
        // variant 0 is the default: the other plan must be 1.5 % faster to replace it (an
        // ensemble step is milliseconds; its best rounds agree to a few tenths of a percent)
        const int best = t_us[1] < 0.985 * t_us[0] ? 1 : 0;
        cur[pairs[pi]] = best;
        if (remeasure) gpe::wisdom_put(2, pairs[pi].first, pairs[pi].second, best);
        else gpe::wisdom_put_if_absent(2, pairs[pi].first, pairs[pi].second, best);
      }
    }
    if (save) {
      chk(hipMemcpyAsync(b->psi, save, b->total * sizeof(double2), hipMemcpyDeviceToDevice, b->stream), "restore");
      chk(hipStreamSynchronize(b->stream), "restore");
    }
    (void)hipFree(save);
    (void)hipFree(items);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (rc) return rc;
  }
  // adopt the recorded plans: new tiles and work items, no captured graph
  for (auto& pr : pairs) {
    const int w = gpe::wisdom_get(2, pr.first, pr.second);
    (pr.first == 0 ? b->var_x : b->var_y)[pr.second] = w > 0 && w < gpe::plan_variants(pr.second) ? w : 0;
  }
  drop_graph(b);
  b->begun = false;
  return build_items(b);
}

}  // extern "C"
