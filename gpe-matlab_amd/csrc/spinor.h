// Two-component (GPEtaskSpinor) drivers.
#pragma once
#include "sweep.h"
struct gpe_handle;
namespace gpe { struct FinParams; }

namespace gpe_spinor {
gpe::SpinorParams spinor_params(gpe_handle* h);
int spinor_sweep(gpe_handle* h, int seq, const gpe::SpinorParams& p, int* nb);
// Coefficients of a scalar coupling (solve_split.m:61-65).
void rabi_scalars(gpe_handle* h);
// y/z sweeps of one step on one array with the kinetic table pair (kcol, kax).
int spinor_rest_of_step(gpe_handle* h, double2* psi, double2* kcol, double2* kax);
// One block of n_rec steps.  tape (if set) receives the real-space state of both
// components before each step (2 * n_rec arrays); reduce = false skips the
// block-end chemical potential and renormalisation.  blk_in / blk_out (3D with
// reduce): the block opens from / ends in the stored x spectrum of a neighbouring
// block of the same outer iteration instead of the real-space state.
int spinor_block(gpe_handle* h, double time2, double2* tape, bool reduce, bool blk_in = false, bool blk_out = false);
// SP_KEX (SP_KEXS when spectral) + y/z kinetic sums of both components into the
// partials; fills f's counts.  ke_groups > 0: the kinetic sums are already in SL_KZ /
// SL_KZ2 over that many work-groups per component (S_KINE), and no kinetic sweeps run.
int spinor_reduce(gpe_handle* h, gpe::FinParams* f, bool scaled, bool spectral = false, int ke_groups = 0);
}  // namespace gpe_spinor

namespace gpe {
int spinor_measure(gpe_handle* h, double out[5]);
int spinor_split_begin(gpe_handle* h, double ddt, int niter_inner, double* muc_out);
int spinor_split_block(gpe_handle* h, double time2, double* muc, double* ncur);
int spinor_split_outer(gpe_handle* h, int j, double* time2, double* muc, double* ncur);
}  // namespace gpe
