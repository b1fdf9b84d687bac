// Plan tuner and plan wisdom: the counterpart of the reference's FFTW planning.
//
// GPEtask.m:28 (fftw_set_wisdom = true) and GPEtask.m:55-58 (fftw('dwisdom',[]);
// fftw('planner','measure'); fftn(grid.mesh.x)) make MATLAB time candidate FFT plans
// for the grid once, when a task is built, and keep the fastest in FFTW's wisdom.
// Here every length with two plans (64, 128, 256, 384, 512; fft_dev.h) may run
// either: the tuner times the steady-state solve_split step (x: the phase sweep S_MID;
// 3D: y forward, z kinetic, y inverse; 2D: y kinetic) for every combination of the
// axes' variants on the handle's own buffers, in context, because a sweep timed alone
// runs on data the previous launch of the same sweep left in the caches and ranks the
// plans differently from the step.  The fastest combination is recorded in a
// process-wide table keyed by (dimensions, axis, length).  Handles and ensembles
// created later read that table, so all runs of one process on one grid shape use the
// same plans and round identically.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <string>
#include <vector>

#include "handle.h"

using namespace gpe_detail;

namespace {

// Relative gain a non-default plan combination needs to be recorded.
constexpr double kMargin = 0.015;
// Each timing covers at least this long (repetitions from a pilot timing): four steps of
// a 256^2 grid are 60 us, too short to resolve the few percent between plans.
constexpr double kMinTimedUs = 2000.0;

// Average time in microseconds of one steady-state step under the handle's current
// variants, over `reps` steps after two untimed ones.
int time_step(gpe_handle* h, int reps, double* us) {
  hipEvent_t e0 = nullptr, e1 = nullptr;
  HIPCHK(h, hipEventCreate(&e0));
  hipError_t e = hipEventCreate(&e1);
  if (e != hipSuccess) {
    (void)hipEventDestroy(e0);
    return fail(h, GPE_EHIP, "hipEventCreate: %s", hipGetErrorString(e));
  }
  const int warm = 2;
  auto sw = [&](int axis, int seq) {
    if (e != hipSuccess) return;
    SweepParams p = base_params(h, axis, h->psi[0], h->V[0]);
    e = gpe::launch_sweep(h->ax[axis].n, seq, axis != 1, p, h->stream, nullptr, h->var[axis]);
  };
  for (int i = 0; i < warm + reps && e == hipSuccess; ++i) {
    if (i == warm) e = hipEventRecord(e0, h->stream);
    sw(0, gpe::S_MID);
    if (h->ndims == 3) {
      sw(1, gpe::S_FWD);
      sw(2, gpe::S_KIN);
      sw(1, gpe::S_INV);
    } else {
      sw(1, gpe::S_KIN);
    }
  }
  if (e == hipSuccess) e = hipEventRecord(e1, h->stream);
  if (e == hipSuccess) e = hipEventSynchronize(e1);
  float ms = 0.f;
  if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (e != hipSuccess) return fail(h, GPE_EHIP, "tuning step: %s", hipGetErrorString(e));
  *us = 1e3 * (double)ms / reps;
  return GPE_OK;
}

}  // namespace

extern "C" {

int gpe_tune(gpe_handle* h, int remeasure, double* times_us) {
  if (!h) return GPE_EINVAL;
  HIPCHK(h, hipSetDevice(h->device));
  const double nan = std::numeric_limits<double>::quiet_NaN();
  if (times_us)
    for (int i = 0; i < 6; ++i) times_us[i] = nan;
  // axes to measure; the others keep the recorded (or default) plan
  int base[3] = {0, 0, 0};
  std::vector<int> axes;
  for (int a = 0; a < h->ndims; ++a) {
    const int n = h->ax[a].n, w = gpe::wisdom_get(h->ndims, a, n);
    base[a] = w > 0 && w < gpe::plan_variants(n) ? w : 0;
    if (gpe::plan_variants(n) > 1 && (remeasure || w < 0)) axes.push_back(a);
  }
  if (!axes.empty()) {
    // the sweeps run in place on component 1: keep the state and put it back afterwards
    HIPCHK(h, hipStreamSynchronize(h->stream));
    double2* save = nullptr;
    HIPCHK(h, hipMalloc(&save, h->npts * sizeof(double2)));
    int rc = GPE_OK;
    hipError_t e = hipMemcpyAsync(save, h->psi[0], h->npts * sizeof(double2), hipMemcpyDeviceToDevice, h->stream);
    // The step is timed on a synthetic field (unit modulus, hashed phases), on the
    // handle's potential (a hashed one when none is set) and, outside a run, on kinetic
    // tables of unit modulus; every solver call rebuilds its tables.  Fresh buffers hold
    // zeros, on which the step ran about 10 % faster than on real data.
    if (e == hipSuccess)
      e = gpe::launch_tune_fill(h->psi[0], h->npts, h->have_V[0] ? nullptr : h->V[0], h->npts, h->stream);
    if (e == hipSuccess && !h->begun && !h->sg_begun) {
      const Axis& m = h->ax[mid_axis(h)];
      e = gpe::launch_tune_fill(h->kcol, m.ncols, nullptr, 0, h->stream);
      if (e == hipSuccess) e = gpe::launch_tune_fill(h->kax, m.n, nullptr, 0, h->stream);
    }
    if (e != hipSuccess) rc = fail(h, GPE_EHIP, "tuning: %s", hipGetErrorString(e));
    const int keep[3] = {h->var[0], h->var[1], h->var[2]};
    const int ncomb = 1 << axes.size();
    int reps = 4;
    if (rc == GPE_OK) {
      for (int a = 0; a < 3; ++a) h->var[a] = base[a];
      double pilot = 0;
      rc = time_step(h, 4, &pilot);
      if (rc == GPE_OK && pilot > 0) reps = (int)std::fmin(256.0, std::fmax(4.0, std::ceil(kMinTimedUs / pilot)));
    }
    std::vector<double> best_t(ncomb, 1e300);
    // rounds run every combination in turn, so clock ramps and other traffic fall on
    // all of them; each combination keeps its best round
    for (int r = 0; r < 3 && rc == GPE_OK; ++r)
      for (int m = 0; m < ncomb && rc == GPE_OK; ++m) {
        for (int a = 0; a < 3; ++a) h->var[a] = base[a];
        for (size_t i = 0; i < axes.size(); ++i) h->var[axes[i]] = (m >> i) & 1;
        double us = 0;
        rc = time_step(h, reps, &us);
        if (us < best_t[m]) best_t[m] = us;
      }
    for (int a = 0; a < 3; ++a) h->var[a] = keep[a];
    e = hipMemcpyAsync(h->psi[0], save, h->npts * sizeof(double2), hipMemcpyDeviceToDevice, h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    if (e != hipSuccess && rc == GPE_OK) rc = fail(h, GPE_EHIP, "tuning: %s", hipGetErrorString(e));
    (void)hipFree(save);
    if (rc) return rc;
    // variant 0 is the default; another combination must beat the one that keeps
    // every measured axis at variant 0 by kMargin to replace it (timings of 2 ms differ
    // by a few tenths of a percent between rounds)
    int mbest = 0;
    for (int m = 1; m < ncomb; ++m)
      if (best_t[m] < best_t[mbest]) mbest = m;
    if (mbest != 0 && !(best_t[mbest] < (1.0 - kMargin) * best_t[0])) mbest = 0;
    for (size_t i = 0; i < axes.size(); ++i) {
      const int a = axes[i], n = h->ax[a].n, v = (mbest >> i) & 1;
      if (times_us)
        for (int m = 0; m < ncomb; ++m) {
          double& slot = times_us[2 * a + ((m >> i) & 1)];
          if (!(slot <= best_t[m])) slot = best_t[m];
        }
      if (remeasure) gpe::wisdom_put(h->ndims, a, n, v);
      else gpe::wisdom_put_if_absent(h->ndims, a, n, v);
    }
  }
  // adopt the recorded plans; a captured outer iteration holds the old ones
  bool changed = false;
  for (int a = 0; a < h->ndims; ++a) {
    const int w = gpe::wisdom_get(h->ndims, a, h->ax[a].n);
    const int v = w > 0 && w < gpe::plan_variants(h->ax[a].n) ? w : 0;
    changed = changed || v != h->var[a];
    h->var[a] = v;
  }
  if (changed) drop_outer_graph(h);
  return GPE_OK;
}

int gpe_get_plan(gpe_handle* h, int axis, int* variant) {
  if (!h || !variant || axis < 0 || axis >= h->ndims) return h ? fail(h, GPE_EINVAL, "bad axis") : GPE_EINVAL;
  *variant = h->var[axis];
  return GPE_OK;
}

int gpe_wisdom_get(int ndims, int axis, int n) { return gpe::wisdom_get(ndims, axis, n); }

int gpe_wisdom_set(int ndims, int axis, int n, int variant) {
  if ((ndims != 2 && ndims != 3) || axis < 0 || axis >= ndims || variant < 0 ||
      variant >= gpe::plan_variants(n))
    return GPE_EINVAL;
  gpe::wisdom_put(ndims, axis, n, variant);
  return GPE_OK;
}

int gpe_wisdom_forget(void) {
  gpe::wisdom_clear();
  return GPE_OK;
}

// Text form "ndims axis n variant" per line (FFTW's wisdom export/import).
int gpe_wisdom_export(char* buf, int len, int* needed) {
  std::string s;
  for (const auto& e : gpe::wisdom_entries())
    s += std::to_string(e[0]) + " " + std::to_string(e[1]) + " " + std::to_string(e[2]) + " " + std::to_string(e[3]) + "\n";
  if (needed) *needed = (int)s.size() + 1;
  if (!buf || len < (int)s.size() + 1) return buf ? GPE_EINVAL : GPE_OK;
  std::memcpy(buf, s.c_str(), s.size() + 1);
  return GPE_OK;
}

int gpe_wisdom_import(const char* text) {
  if (!text) return GPE_EINVAL;
  std::vector<std::array<int, 4>> rows;
  const char* p = text;
  while (*p) {
    int nd, ax, n, v, used = 0;
    while (*p == ' ' || *p == '\n' || *p == '\r' || *p == '\t') ++p;
    if (!*p) break;
    if (std::sscanf(p, "%d %d %d %d%n", &nd, &ax, &n, &v, &used) != 4 || used <= 0) return GPE_EINVAL;
    if ((nd != 2 && nd != 3) || ax < 0 || ax >= nd || v < 0 || v >= gpe::plan_variants(n)) return GPE_EINVAL;
    rows.push_back({nd, ax, n, v});
    p += used;
  }
  for (const auto& r : rows) gpe::wisdom_put(r[0], r[1], r[2], r[3]);   // all or nothing
  return GPE_OK;
}

}  // extern "C"
