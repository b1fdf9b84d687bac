// Pencil sweep kernels for one FFT length.  This file is compiled once per
// supported length with -DGPE_N=<n>; each object exports launch_sweep_<n>.
#include "sweep_body.h"

#ifndef GPE_N
#error "compile with -DGPE_N=<pencil length>"
#endif

namespace gpe {

#define GPE_CAT2(a, b) a##b
#define GPE_CAT(a, b) GPE_CAT2(a, b)

hipError_t GPE_CAT(launch_sweep_, GPE_N)(int seq, bool colfast, int var, const SweepParams& p, hipStream_t s,
                                          int* nblocks) {
  constexpr int N = GPE_N;
  long long nb = 0;
  const dim3 block(256);
  // pencils per work-group follow the sequence's plan (sweep_cols); variant 1 exists
  // for the tunable sequences of lengths with a second plan
  if (!has_variant<N>() || !tunable_seq(seq)) var = 0;
  if (p.psi2 && seq != S_FWD && seq != S_INV && seq != S_KIN && seq != S_KINE) return hipErrorInvalidValue;
#define GPE_LV(CF, S, V)                                                                               \
  {                                                                                                    \
    constexpr int C = sweep_cols<N, S, V>();                                                           \
    nb = ((long long)p.ncols + C - 1) / C * (p.psi2 ? 2 : 1);                                          \
    if (nb > 0) hipLaunchKernelGGL((sweep_kernel<N, C, CF, S, V>), dim3((unsigned)nb), block, 0, s, p); \
  }
#define GPE_L(CF, S) \
  case S: GPE_LV(CF, S, 0) break;
#if GPE_N == 64 || GPE_N == 128 || GPE_N == 256 || GPE_N == 384 || GPE_N == 512
  static_assert(has_variant<N>(), "variant list out of date");
#define GPE_T(CF, S)                                   \
  case S:                                              \
    if (var == 1) { GPE_LV(CF, S, 1) break; }          \
    GPE_LV(CF, S, 0) break;
#else
  static_assert(!has_variant<N>(), "variant list out of date");
#define GPE_T(CF, S) GPE_L(CF, S)
#endif
  if (colfast) {
    switch (seq) {
      GPE_T(true, S_FWD) GPE_T(true, S_INV) GPE_T(true, S_KIN) GPE_T(true, S_OPEN) GPE_T(true, S_MID)
      GPE_T(true, S_CLOSE) GPE_T(true, S_KEX) GPE_T(true, S_KE) GPE_L(true, S_ITPA) GPE_L(true, S_ITPB)
      GPE_T(true, S_CLOSEM) GPE_T(true, S_KINE) GPE_L(true, S_ROPEN) GPE_L(true, S_RMID) GPE_L(true, S_RCLOSEM)
      GPE_L(true, S_RITPA) GPE_L(true, S_RITPB) GPE_L(true, S_SOPEN) GPE_L(true, S_SMID) GPE_L(true, S_KINM)
#if GPE_N == 64 || GPE_N == 128 || GPE_N == 256
      GPE_L(true, S_YKIN) GPE_L(true, S_YKINM) GPE_L(true, S_YKEZ)
#endif
      default: return hipErrorInvalidValue;
    }
  } else {
    switch (seq) {
      GPE_T(false, S_FWD) GPE_T(false, S_INV) GPE_T(false, S_KIN) GPE_T(false, S_KE)
      GPE_T(false, S_KEX) GPE_L(false, S_RKIN) GPE_L(false, S_KINM)
#if GPE_N == 256
      GPE_T(false, S_YKE)
#endif
      default: return hipErrorInvalidValue;
    }
  }
#undef GPE_T
#undef GPE_L
#undef GPE_LV
  if (nblocks) *nblocks = (int)nb;
  return nb > 0 ? hipGetLastError() : hipSuccess;
}

}  // namespace gpe
