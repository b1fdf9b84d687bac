// Host/device interface of the pencil sweep kernels.
//
// A "sweep" reads every pencil of one axis once, runs a fused sequence
// (inverse FFT, pointwise phase, forward FFT, kinetic multiply, reductions) in
// registers and LDS, and writes the pencil back once.  The grid is stored
// column-major as MATLAB's [ny, nx, nz] (y contiguous), interleaved complex128.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace gpe {

// Sequences a sweep can run on each pencil.
enum Seq : int {
  S_FWD = 0,    // forward FFT
  S_INV = 1,    // inverse FFT (no 1/N)
  S_KIN = 2,    // forward, x kcol[col]*kax[i], inverse          (mid-spectrum axis)
  S_OPEN = 3,   // x scale, block-open half phase, forward        (solve_split.m:55-56)
  S_MID = 4,    // inverse, full phase, forward                   (solve_split.m:60-61)
  S_CLOSE = 5,  // inverse, full phase, block-close half phase    (solve_split.m:60-64)
  S_KEX = 6,    // reduce |psi|^2, V|psi|^2, |psi|^4, then forward and reduce k^2/2 |.|^2
  S_KE = 7,     // forward and reduce k^2/2 |.|^2
  S_ITPA = 8,   // inverse, x stored real operator N, reduce |psi|^2 (groundstate_itp.m:41-44)
  S_ITPB = 9,   // x mu, N = exp(-dt/2 (g mu^2 n + V)), store N, x N, forward (groundstate_itp.m:76,40)
  S_NSEQ = 10
};

// Reduction slots of the partial-sum buffer.
enum Slot : int { SL_A = 0, SL_B = 1, SL_C = 2, SL_KX = 3, SL_KY = 4, SL_KZ = 5, SL_MAX = 6, SL_NSLOT = 8 };

// Scalars that evolve on the device between kernels (no host round trip).
struct DevState {
  double mu;      // chemical potential in VV = V - mu
  double scale;   // pending real scale of the stored state (renormalisation)
  double muc;     // last reported chemical potential
  double ncur;    // last block norm
  double NNN;     // current target norm
  double sA, sB, sC, sK;  // last finalized raw sums
  double delta;   // ITP max-norm change
  double mu_itp;  // ITP renormalisation factor of the last iteration
  double pad[4];
};

struct SweepParams {
  double2* psi;              // pencil data (in place)
  const double* V;           // potential (same layout), may be null when unused
  const double2* tw;         // forward twiddles e^{-2 pi i j/n}, j < n
  const double2* kcol;       // kinetic factor per column (S_KIN)
  const double2* kax;        // kinetic factor per axis index incl. 1/Npts (S_KIN)
  const double* keax;        // k_i^2/2/n per axis index (S_KEX, S_KE)
  double* nop;               // stored real ITP operator (S_ITPA reads, S_ITPB writes)
  double* partials;          // reduction partials [slot * maxwg + block]
  DevState* st;              // device scalars
  unsigned ncols;            // number of pencils
  unsigned A, sa, sb;        // pencil base = (col % A) * sa + (col / A) * sb
  unsigned stride;           // element stride along the pencil (grid <= 2^27 points)
  int maxwg;                 // slot pitch of partials
  int slot_ke;               // slot receiving k^2 sums
  double g;                  // interaction constant
  double dr, di;             // complex time step dt = dr + i di (ITP: dr = dt, di = 0)
};

// Launch one sweep; returns hipErrorInvalidValue for an unsupported length.
hipError_t launch_sweep(int n, int seq, bool colfast, const SweepParams& p, hipStream_t s, int* nblocks);
bool sweep_length_supported(int n);

}  // namespace gpe
