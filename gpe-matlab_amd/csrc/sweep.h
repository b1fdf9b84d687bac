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
  S_CLOSEM = 10,  // S_CLOSE, then the S_KEX reductions of the stored state in the same pass
  // rotating frame (scalar omega, set_kinop.m:5-7, ssft_kin_step.m:22-25): the x half
  // steps Kx = forward, x rtab[row(col)][kx], inverse, all in real space between sweeps
  S_RKIN = 11,    // forward, x rtab[row][i], inverse (2D y step, table over (x, ky))
  S_ROPEN = 12,   // scale, block-open half phase, Kx
  S_RMID = 13,    // Kx, full phase, Kx
  S_RCLOSEM = 14, // Kx, full and closing half phase; block-end sums incl. the x kinetic term
  S_RITPA = 15,   // Kx, x stored operator N, reduce |psi|^2
  S_RITPB = 16,   // x mu, N = exp(-dt/2 (g mu^2 n + V)), store N, x N, Kx
  // stochastic GPE (SGPEtask/solve_split.m:55-72): the noise of the next step is added
  // in real space before the forward x transform
  S_SOPEN = 17,   // block-open half phase, + noise, forward
  S_SMID = 18,    // inverse, full phase, + noise, forward
  S_KINM = 19,    // S_KIN with the energy cut-off mask kk < ecut
  // 3D steps with a 256-point y axis split as 16 x 16 (y = a + 16 b): the b factor runs
  // in the x sweeps (tiles of one a x 16 b x nx), the a factor in the z sweep, so a step
  // is two sweeps; between them the state is x-spectral with the y factor over b applied
  S_YKIN = 20,    // y factor over a (16 adjacent pencils), forward, x kcol[col]*kax[i], inverse, y factor back
  S_YOPEN = 21,   // scale, block-open half phase, forward x, y factor over b
  S_YMID = 22,    // y factor over b back, inverse x, full phase, forward x, y factor over b
  S_YCLOSEM = 23, // y factor over b back, inverse x, full and closing half phase; block-end sums
  // the stochastic GPE on the same split: noise after the phase, and the cut-off mask
  S_YSOPEN = 24,  // S_YOPEN, + noise before the forward x transform
  S_YSMID = 25,   // S_YMID, + noise before the forward x transform
  S_YKINM = 26,   // S_YKIN with the energy cut-off mask kk < ecut
  // the rotating frame on the same split: x stays real between the sweeps, the x half
  // steps Kx run around the phase in the x sweep, the z sweep takes the (ky, x) table
  S_YROPEN = 27,   // scale, block-open half phase, Kx, y factor over b
  S_YRMID = 28,    // y factor over b back, Kx, full phase, Kx, y factor over b
  S_YRCLOSEM = 29, // y factor over b back, Kx, full and closing half phase; block-end sums
  // y-split steps: S_KE along y of a state in block order (block ends inside an outer
  // iteration)
  S_YKE = 30,
  // y-split block ends inside an outer iteration hand over in the sweeps' own mixed form
  // (x spectral, y factor over b applied):
  S_YCLOSEB = 31,  // S_YMID with the closing phase and the block-end sums of |psi|^2, V|psi|^2, |psi|^4
  S_YOPENB = 32,   // S_YMID with the pending scale and the block-open half phase instead of the full phase
  S_YKEZ = 33,     // forward z, y factor over a: the whole kinetic sum from the full spectrum
  // S_KIN that also sums kk |.|^2 over the spectrum it multiplied (3D spinor blocks end
  // with a kinetic step, so that spectrum is the block-end state's)
  S_KINE = 34,
  S_NSEQ = 35
};

// Reduction slots of the partial-sum buffer.  Spinor runs use C for |psi1|^4,
// C12 for |psi1|^2|psi2|^2, C2 for |psi2|^4, X for Re conj(psi1) Omega psi2 and
// KX2..KZ2 for the kinetic sums of the second component.
enum Slot : int {
  SL_A = 0, SL_B = 1, SL_C = 2, SL_KX = 3, SL_KY = 4, SL_KZ = 5, SL_MAX = 6,
  SL_C12 = 7, SL_C2 = 8, SL_X = 9, SL_KX2 = 10, SL_KY2 = 11, SL_KZ2 = 12, SL_NSLOT = 16
};

// Scalars that evolve on the device between kernels (no host round trip).
// Work-groups are dealt to the 8 XCDs round robin (group b runs on XCD b % 8).  When a
// tile of pencils covers only half of a 128-byte L2 line of an 8-byte array (8 pencils
// in a row of V), neighbouring tiles 2j and 2j+1 go to groups on the same XCD, so its L2
// fetches each line once; groups beyond the last multiple of 16 keep their index.
__device__ __forceinline__ unsigned xcd_pair_tile(unsigned b, unsigned nb) {
  const unsigned full = nb & ~15u;
  if (b >= full) return b;
  const unsigned x = b & 7u, s = b >> 3;
  return ((s >> 1) << 4) + (x << 1) + (s & 1u);
}

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
  double sc[2];   // two-component ITP: pending renormalisation factor of each component
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
  int apply_scale;           // S_FWD: multiply by the pending scale st->scale first
  // rotating frame
  const double2* rtab;       // S_R*: kinetic table [row][i], row = col % trow (incl. 1/N of the axis)
  unsigned trow;             // row modulus of rtab and rcoord
  const double* k1ax;        // S_KEX/S_KE/S_RCLOSEM: k_i / n of the swept axis
  const double* rcoord;      // coordinate per row (y for x pencils, x for y pencils)
  double rot_w;              // kinetic-sum cross term weight: sum (k^2/2 + rot_w coord k) |F psi|^2 / n
  // stochastic GPE
  double noise_var;          // sqrt(T gamma ddt / weight)
  unsigned long long noise_step, noise_seed;   // Philox counter word 1-2 and key
  const int* kmax;           // S_KINM: largest |k| index of the mid axis kept in each column (-1: none)
  // y-split 3D steps (S_Y*)
  const double2* twy;        // y twiddles e^{-2 pi i j/256}
  unsigned gnx, gny, gnz;    // grid sizes (x sweeps of S_YOPEN/S_YMID/S_YCLOSEM address the grid directly)
  // S_FWD / S_INV / S_KIN of sweep_kernel: a second array of the same shape, swept by
  // the second half of the grid with the same tables (null: one array)
  double2* psi2;
  // y-split 3D steps: the state in block order, from a block's opening x sweep (which
  // reads psi) to its closing one (which writes psi); the z sweeps run on it in place
  double2* psiy;
  // y-split blocks of one outer iteration hand the state over in block order: the
  // opening x sweep reads psiy instead of psi when yblk_in, the closing one writes
  // psiy instead of psi when yblk_out
  int yblk_in, yblk_out;
  // S_YKEZ: (kx^2 + ky^2)/2/npts per column in the z sweep's split order (as kcol),
  // then kz^2/2/npts per kz index.  S_KINE: its column weights (keax: those of the
  // swept axis)
  const double* kecol;
  int slot_ke2;              // S_KINE with psi2: slot of the second array's sums
};

// Two-component x-sweep parameters (GPEtaskSpinor).
enum SpSeq : int { SP_PR = 0, SP_KEX = 1, SP_ADJ = 2, SP_ITPA = 3, SP_ITPB = 4, SP_ITPC = 5, SP_KEXS = 6 };

struct SpinorParams {
  double2* psi1;
  double2* psi2;
  const double* V1;
  const double* V2;
  const double2* cpl;        // space-dependent coupling Omega(x), or null for a scalar
  double cpl_max2;           // max |Omega(x)|^2 over the array (selects the series Rabi factors)
  const double2* tw;
  double* partials;
  DevState* st;
  unsigned ncols, A, sa, sb, stride;
  int maxwg;
  double g11, g12, g21, g22;
  double dr, di;             // complex dt
  double2 cpl_s;             // scalar coupling (cpl == null)
  double2 cosom, sinomm, sinomp;  // scalar Rabi coefficients (solve_split.m:61-65), cpl == null
  double nlincpl;            // nonlinear coupling (solve_split.m:84-89)
  const double* keax;        // k^2/2/n along x (SP_KEX)
  // Rabi adjoint (gamma = 0, Omega(x) = om0 exp(i kw x)):
  double2* ust1;             // SP_PR: if set, store the real-space state before the step there;
  double2* ust2;             // SP_ADJ: read that stored state back
  const double* xc;          // SP_ADJ: x coordinate per x index
  double om0, kw;            // SP_ADJ: coupling amplitude and wave number
  double* gacc;              // SP_ADJ: per-work-group accumulators of dL/d om0 [0, maxwg) and dL/d kw
  // two-component imaginary time (dt = dr real):
  double* pbuf;              // SP_ITPA writes, SP_ITPB reads exp(-(V_k + g_k1 n1 + g_k2 n2) dt), 2 x npts
  unsigned npts;             // grid points (offset of component 2 in pbuf)
  int apply_sc;              // SP_ITPA, SP_KEX: multiply the loaded state by st->sc[k] first
  int mixture;               // SP_ITPB: no Rabi rotation (incoherent mixture)
};

// Launch one sweep; returns hipErrorInvalidValue for an unsupported length.
// var selects the plan variant of the tunable sequences (0 = default plan).
hipError_t launch_sweep(int n, int seq, bool colfast, const SweepParams& p, hipStream_t s, int* nblocks,
                        int var = 0);
bool sweep_length_supported(int n);
hipError_t launch_spinor_sweep(int n, int seq, const SpinorParams& p, hipStream_t s, int* nblocks);
// x sweeps of the y-split 3D step (S_YOPEN, S_YMID, S_YCLOSEM, S_YSOPEN, S_YSMID, S_YROPEN,
// S_YRMID, S_YRCLOSEM) for x length n; p.V is the potential in row order
// (launch_ysplit_permute_v)
hipError_t launch_ysplit_x(int n, int seq, const SweepParams& p, hipStream_t s, int* nblocks);
bool ysplit_x_supported(int n);
// Vp[((z 16 + a) 16 + b) nx + x] = V[a + 16 b + 256 x + 256 nx z]: the potential in the
// row order of the y-split x sweeps
hipError_t launch_ysplit_permute_v(const double* V, double* Vp, unsigned nx, unsigned nz, hipStream_t s);
hipError_t launch_rowdft_selftest(const double2* in, double2* out_f, double2* out_i, hipStream_t s);

}  // namespace gpe
