// Two-component (GPEtaskSpinor) drivers.
#include "handle.h"
#include "spinor.h"

using namespace gpe_detail;

namespace gpe {
int spinor_measure(gpe_handle* h, double*) { return fail(h, GPE_EINVAL, "spinor path not built"); }
int spinor_split_begin(gpe_handle* h, double, int, double*) { return fail(h, GPE_EINVAL, "spinor path not built"); }
int spinor_split_block(gpe_handle* h, double, double*, double*) { return fail(h, GPE_EINVAL, "spinor path not built"); }
int spinor_split_outer(gpe_handle* h, int, double*, double*, double*) { return fail(h, GPE_EINVAL, "spinor path not built"); }
}  // namespace gpe
