// Two-component (GPEtaskSpinor) drivers.
#pragma once
struct gpe_handle;

namespace gpe {
int spinor_measure(gpe_handle* h, double out[5]);
int spinor_split_begin(gpe_handle* h, double ddt, int niter_inner, double* muc_out);
int spinor_split_block(gpe_handle* h, double time2, double* muc, double* ncur);
int spinor_split_outer(gpe_handle* h, int j, double* time2, double* muc, double* ncur);
}  // namespace gpe
