// Adjoint split-step: device interface.
#pragma once
#include "sweep.h"

namespace gpe {

enum AdjSeq : int { A_OPEN = 0, A_MID = 1, A_CLOSE = 2 };

struct AdjParams {
  double2* psi;          // state, reconstructed backwards
  double2* lam;          // adjoint 2 dL/d conj(psi)
  const double* V;       // potential of the current interval
  double* gV;            // accumulated dL/dV of the current interval
  const double2* tw;
  DevState* st;
  unsigned ncols, A, sa, sb, stride;
  double g;
  double h;              // real time step ddt (gamma = 0: dt = i ddt)
};

hipError_t launch_adj_x(int n, int seq, int var, const AdjParams& p, hipStream_t s, int* nblocks);

}  // namespace gpe
