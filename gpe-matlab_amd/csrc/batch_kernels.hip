// Batched ragged ensemble sweeps for one pencil length (-DGPE_N=<n>): one launch
// sweeps one axis of every trajectory whose axis has that length.
//
// Work-group b takes work item items[b] = (trajectory, pencil tile); the tile is
// processed by the same sweep body, FFT plan and reduction order as the
// single-grid kernels, so a trajectory of the ensemble reproduces a
// single-trajectory run bit for bit.
#include "batch.h"
#include "sweep_body.h"

#ifndef GPE_N
#error "compile with -DGPE_N=<pencil length>"
#endif

namespace gpe {

template <int N, bool COLFAST, int SEQ, int VAR>
__global__ void __launch_bounds__(256) batch_sweep_kernel(const BatchLaunch b) {
  constexpr int C = sweep_cols<N, SEQ, VAR>();
  __shared__ double lds[Lds<N, C, COLFAST>::SIZE];
  __shared__ double2 twl[SEQ == S_KIN ? 2 * N : N];
  __shared__ double red[4];
  // 8-pencil tiles: consecutive items (neighbouring tiles of one trajectory) on one XCD
  const WorkItem it = b.items[C < 16 ? xcd_pair_tile(blockIdx.x, gridDim.x) : blockIdx.x];
  const TrajDev& tj = b.traj[it.traj];
  SweepParams p{};
  p.psi = tj.psi;
  p.V = tj.V;
  p.kcol = tj.kcol;
  p.kax = tj.kax;
  p.partials = tj.partials;
  p.st = tj.st;
  p.maxwg = tj.maxwg;
  p.g = tj.g;
  p.dr = tj.dr;
  p.di = tj.di;
  if (COLFAST) {  // x pencils: columns iy, stride ny
    p.tw = tj.tw_x;
    p.keax = tj.keax_x;
    p.ncols = tj.ny;
    p.A = tj.ny;
    p.sa = 1;
    p.sb = tj.ny * tj.nx;
    p.stride = tj.ny;
    p.slot_ke = SL_KX;
  } else {        // y pencils: columns ix, contiguous
    p.tw = tj.tw_y;
    p.keax = tj.keax_y;
    p.ncols = tj.nx;
    p.A = 1;
    p.sa = 0;
    p.sb = tj.ny;
    p.stride = 1;
    p.slot_ke = SL_KY;
  }
  sweep_body<N, C, COLFAST, SEQ, VAR>(p, p.psi, (unsigned)it.tile, (unsigned)it.tile, lds, twl, red);
}

#define GPE_CAT2(a, b) a##b
#define GPE_CAT(a, b) GPE_CAT2(a, b)

// items must all belong to trajectories whose swept axis has length GPE_N, with tiles
// of the plan variant `var`
hipError_t GPE_CAT(launch_batch_sweep_, GPE_N)(bool xaxis, int seq, int var, const BatchLaunch& b, hipStream_t s) {
  constexpr int N = GPE_N;
  if (b.nitems <= 0) return hipSuccess;
  const dim3 grid((unsigned)b.nitems), block(256);
#if GPE_N == 64 || GPE_N == 128 || GPE_N == 256 || GPE_N == 384 || GPE_N == 512
#define GPE_L(CF, S)                                                                              \
  case S:                                                                                         \
    if (var == 1) hipLaunchKernelGGL((batch_sweep_kernel<N, CF, S, 1>), grid, block, 0, s, b);    \
    else hipLaunchKernelGGL((batch_sweep_kernel<N, CF, S, 0>), grid, block, 0, s, b);              \
    break;
#else
  if (var != 0) return hipErrorInvalidValue;
#define GPE_L(CF, S) \
  case S: hipLaunchKernelGGL((batch_sweep_kernel<N, CF, S, 0>), grid, block, 0, s, b); break;
#endif
  if (xaxis) {
    switch (seq) {
      GPE_L(true, S_OPEN) GPE_L(true, S_MID) GPE_L(true, S_CLOSE) GPE_L(true, S_KEX)
      default: return hipErrorInvalidValue;
    }
  } else {
    switch (seq) {
      GPE_L(false, S_KIN) GPE_L(false, S_KE)
      default: return hipErrorInvalidValue;
    }
  }
#undef GPE_L
  return hipGetLastError();
}

}  // namespace gpe
