// Pencil sweep kernels for one FFT length.  This file is compiled once per
// supported length with -DGPE_N=<n>; each object exports launch_sweep_<n>.
#include "sweep_body.h"

#ifndef GPE_N
#error "compile with -DGPE_N=<pencil length>"
#endif

namespace gpe {

#define GPE_CAT2(a, b) a##b
#define GPE_CAT(a, b) GPE_CAT2(a, b)

hipError_t GPE_CAT(launch_sweep_, GPE_N)(int seq, bool colfast, const SweepParams& p, hipStream_t s,
                                          int* nblocks) {
  constexpr int N = GPE_N;
  long long nb = 0;
  const dim3 block(256);
  // pencils per work-group follow the sequence's plan (sweep_cols)
#define GPE_L(CF, S)                                                                                   \
  case S: {                                                                                            \
    constexpr int C = sweep_cols<N, S>();                                                              \
    nb = ((long long)p.ncols + C - 1) / C;                                                             \
    if (nb > 0) hipLaunchKernelGGL((sweep_kernel<N, C, CF, S>), dim3((unsigned)nb), block, 0, s, p);   \
    break;                                                                                             \
  }
  if (colfast) {
    switch (seq) {
      GPE_L(true, S_FWD) GPE_L(true, S_INV) GPE_L(true, S_KIN) GPE_L(true, S_OPEN) GPE_L(true, S_MID)
      GPE_L(true, S_CLOSE) GPE_L(true, S_KEX) GPE_L(true, S_KE) GPE_L(true, S_ITPA) GPE_L(true, S_ITPB)
      GPE_L(true, S_CLOSEM) GPE_L(true, S_ROPEN) GPE_L(true, S_RMID) GPE_L(true, S_RCLOSEM)
      GPE_L(true, S_RITPA) GPE_L(true, S_RITPB) GPE_L(true, S_SOPEN) GPE_L(true, S_SMID) GPE_L(true, S_KINM)
      default: return hipErrorInvalidValue;
    }
  } else {
    switch (seq) {
      GPE_L(false, S_FWD) GPE_L(false, S_INV) GPE_L(false, S_KIN) GPE_L(false, S_KE)
      GPE_L(false, S_KEX) GPE_L(false, S_RKIN) GPE_L(false, S_KINM)
      default: return hipErrorInvalidValue;
    }
  }
#undef GPE_L
  if (nblocks) *nblocks = (int)nb;
  return nb > 0 ? hipGetLastError() : hipSuccess;
}

}  // namespace gpe
