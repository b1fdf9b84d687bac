// Batched ragged ensemble: device interface.
#pragma once
#include "state.h"
#include "sweep.h"

namespace gpe {

// One 2D trajectory of the ensemble (column-major [ny, nx] like a single grid).
struct TrajDev {
  double2* psi;
  const double* V;
  const double2* kcol;       // kinetic factor per x index (mid-spectrum axis y)
  const double2* kax;        // kinetic factor per y index, incl. 1/(nx ny)
  const double* keax_x;
  const double* keax_y;
  const double2* tw_x;
  const double2* tw_y;
  DevState* st;
  double* partials;          // SL_NSLOT * maxwg doubles
  unsigned nx, ny;
  int maxwg;
  int tiles_x, tiles_y;      // work-groups of an x / y sweep (reduction counts)
  int renorm;
  double g, dr, di;
  double w, NN0, tau;
};

struct WorkItem {
  int traj;
  int tile;
};

struct BatchLaunch {
  const TrajDev* traj;
  const WorkItem* items;
  int nitems;
};

// xaxis: x sweep (strided, phase sequences) or y sweep (contiguous, kinetic / KE)
// over items whose swept axis has length n
// with tiles of plan variant var
hipError_t launch_batch_sweep(int n, bool xaxis, int seq, int var, const BatchLaunch& b, hipStream_t s);
// one work-group per trajectory list[0 .. ntraj): FIN_BLOCK or FIN_SUMS on its partials
hipError_t launch_batch_finalize(const TrajDev* traj, const int* list, int ntraj, int mode, double time2,
                                 hipStream_t s);

}  // namespace gpe
