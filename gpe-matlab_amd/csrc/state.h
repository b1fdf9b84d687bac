// Block-boundary kernels: second-stage reductions and device-side scalars.
#pragma once
#include <array>
#include <vector>

#include "sweep.h"

namespace gpe {

enum FinMode : int {
  FIN_SUMS = 0, FIN_BLOCK = 1, FIN_ITP = 2, FIN_MAX = 3, FIN_SPINOR = 4,
  FIN_ITP2 = 5,    // two-component ITP renormalisation and MU histories
  FIN_ITP2M = 6    // two-component ITP applyham / energy histories
};

struct FinParams {
  const double* partials;
  int maxwg;
  int cnt[SL_NSLOT];   // number of partials in each slot (0 = unused)
  DevState* st;
  int mode;
  int renorm;          // FIN_BLOCK: gamma > 0 && mu_init == 0
  double w;            // grid weight dx*dy(*dz)
  double g;
  double NN0, tau, time2;
  double dt;           // FIN_ITP / FIN_MAX: real time step
  double Ntotal, mu_init;
  double* MU;          // FIN_ITP: device history
  int iter;
  double g11, g12, g21, g22;   // FIN_SPINOR: interaction matrix
  // FIN_ITP2 / FIN_ITP2M
  int mixture;         // per-component norms N1, N2 (else one total norm n1t)
  double n1t, n2t;     // target norms
  double* MU2;         // device histories, entry `iter`
  double* EE;
  int which;           // FIN_ITP2M: 1 = write MU2, 2 = write EE
};

hipError_t launch_finalize(const FinParams& f, hipStream_t s);
hipError_t launch_itp_check(const double2* psi, double2* old, const DevState* st, long long n, double* partials,
                            int maxwg, int nblocks, hipStream_t s);
hipError_t launch_scale(double2* psi, DevState* st, int which, long long n, int nblocks, hipStream_t s,
                        bool reset = true);
// pencils per work-group of the non-phase sweeps of length n under plan variant var
int sweep_cols_per_block(int n, int var = 0);
// the same for the phase sweeps (open / mid / close / kex / ITP on the x axis)
int sweep_cols_per_block_phase(int n, int var = 0);
// fewest pencils per work-group over every variant and sequence class (sizes partials)
int sweep_cols_min(int n);
// number of plan variants of length n (0: unsupported length)
int plan_variants(int n);
// process-wide plan wisdom keyed by (dimensions, axis, length); -1 = none recorded
int wisdom_get(int ndims, int axis, int n);
void wisdom_put(int ndims, int axis, int n, int var);
bool wisdom_put_if_absent(int ndims, int axis, int n, int var);
void wisdom_clear();
// every recorded (ndims, axis, n, variant)
std::vector<std::array<int, 4>> wisdom_entries();
hipError_t launch_phase_selftest(const double* er, const double* ei, double2* out, int n, hipStream_t s);
// Synthetic inputs for the plan tuners: z[0..nz) of unit modulus with a hashed phase
// and, when v is not null, v[0..nv) hashed in [0, 10).
hipError_t launch_tune_fill(double2* z, long long nz, double* v, long long nv, hipStream_t s);
// Array order conversion between C order [ny][nmid][nlast] (last index fastest) and
// column-major [ny, nmid, nlast] (first index fastest).  to_colmajor: in is C order,
// out column-major; otherwise the reverse, with every value multiplied by `scale`
// (skipped when scale == 1).  elem: 16 (complex128) or 8 (float64) bytes.
hipError_t launch_reorder(const void* in, void* out, int elem, unsigned ny, unsigned nmid, unsigned nlast,
                          bool to_colmajor, double scale, hipStream_t s);

}  // namespace gpe
