// x sweeps of the y-split 3D step for one x length (-DGPE_N=<nx>): S_YOPEN, S_YMID,
// S_YCLOSEM (solve_split.m:55-56, 60-61, 60-65).
//
// The y axis has 256 points, y = a + 16 b.  A work-group takes the tile of one z
// plane with a in {2 ap, 2 ap + 1} and every b: 32 x-pencils, pencil c <-> (a = 2 ap
// + c/16, b = c % 16), T = nx/E threads each (tid = t*32 + c).  Each memory
// instruction of a wave then reads 32-byte pieces (the two a of one b) of 32 lines,
// and the 8 tiles of a plane run on one XCD at about the same time, so its L2 serves
// the other pieces of each line.  The 16 pencils with one a fill a DPP row, so the
// y factor over b (radix-2 across lanes, rowdft_dev.h) needs no data movement; the
// twiddle W256^{a k1} of the 16 x 16 split is applied here.
//
// The potential is read in tile order (launch_ysplit_permute_v): one contiguous
// 256-byte row per pencil index m and pair of lanes t.
#include <type_traits>

#include "rowdft_dev.h"
#include "sweep_body.h"

#ifndef GPE_N
#error "compile with -DGPE_N=<x length>"
#endif

namespace gpe {

// Sum over a work-group of NW wavefronts; valid on thread 0.  `red` holds NW doubles.
template <int NW>
__device__ __forceinline__ double block_sum_w(double v, double* red) {
  v = wave_sum(v);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) red[w] = v;
  __syncthreads();
  double r = 0.0;
  if (threadIdx.x == 0) {
    r = red[0];
#pragma unroll
    for (int i = 1; i < NW; ++i) r += red[i];
  }
  return r;
}

// PV 0: the plan of the phase sweeps (8.8.4 for 256: 8 elements per thread);
// PV 1: Plan<N> (16.16 for 256: 16 elements per thread, half the threads).
template <int N, int PV>
using YPlan = typename std::conditional<PV == 0, typename SweepPlan<N, S_MID, 0>::type, Plan<N>>::type;

// NA = number of a values per tile (1: 16 pencils, 16-byte pieces; 2: 32 pencils,
// 32-byte pieces).  Tiles per z plane: 16 / NA.
template <int N, int NA, int PV>
constexpr int ythreads() { return 16 * NA * YPlan<N, PV>::T; }

template <int N, int SEQ, int NA, int PV>
__global__ void __launch_bounds__((ythreads<N, NA, PV>())) ysplit_x_kernel(const SweepParams p) {
  using PL = YPlan<N, PV>;
  constexpr int T = PL::T, E = PL::E, C = 16 * NA, NT = C * T, NW = NT / 64, TPP = 16 / NA;
  static_assert(NT % 64 == 0, "whole wavefronts");
  __shared__ double lds[Lds<N, C, true>::SIZE];
  __shared__ double2 twl[N];
  __shared__ double2 twy[256];
  __shared__ double red[NW];
  const int tid = threadIdx.x;
  for (int i = tid; i < N; i += NT) twl[i] = p.tw[i];
  for (int i = tid; i < 256; i += NT) twy[i] = p.twy[i];
  __syncthreads();

  // the TPP tiles of a z plane go to one XCD (group b runs on XCD b % 8); gridDim = TPP nz
  const unsigned bi = blockIdx.x, xcd = bi & 7u, s8 = bi >> 3;
  const unsigned ap = s8 % (unsigned)TPP, z = (s8 / (unsigned)TPP) * 8u + xcd;
  const int c = tid % C, t = tid / C;
  const int r = c & 15;                        // lane in the DPP row = b
  const int a = NA * (int)ap + (c >> 4);
  const unsigned ny = p.gny, nx = p.gnx;
  const unsigned base = (unsigned)(a + 16 * r) + ny * (unsigned)t + ny * nx * z;
  const unsigned step = (unsigned)T * ny;
  const unsigned vbase = (z * (unsigned)TPP + ap) * (unsigned)C * nx + (unsigned)t * (unsigned)C + (unsigned)c;
  const unsigned vstep = (unsigned)T * (unsigned)C;

  double2 v[E];
#pragma unroll
  for (int m = 0; m < E; ++m) v[m] = p.psi[base + (unsigned)m * step];
  double vv[E];
#pragma unroll
  for (int m = 0; m < E; ++m) vv[m] = p.V[vbase + (unsigned)m * vstep];

  RowFactors rf;
  rf.init(r, twy, 16);
  // W256^{a k1}, k1 = bitrev4(b lane) after the forward factor (read where used)
  const int ia = a * bitrev4(r);
  const bool has_re = p.dr != 0.0;
  const double mu = p.st->mu;

  if constexpr (SEQ == S_YOPEN) {
    // solve_split.m:55-56 with the pending renormalisation (solve_split.m:72-73)
    const double sc = p.st->scale;
#pragma unroll
    for (int m = 0; m < E; ++m) {
      const double2 x = cscale(v[m], sc);
      const double aa = -(vv[m] - mu) - p.g * cabs2(x);
      v[m] = mul_cexp(x, aa * p.dr * 0.5, aa * p.di * 0.5, has_re);
    }
  } else {
    // the previous x sweep left the y factor over b applied: undo it
    {
      const double2 wa = twy[ia];
#pragma unroll
      for (int m = 0; m < E; ++m) v[m] = cmulc(v[m], wa);
    }
    row_dft16_inv<E>(v, rf);
    fft_pencil<PL, +1, C, true>(v, t, c, lds, twl);
#pragma unroll
    for (int m = 0; m < E; ++m) {
      const double vm = vv[m] - mu;
      const double aa = -vm - p.g * cabs2(v[m]);                  // solve_split.m:60
      double2 x = mul_cexp(v[m], aa * p.dr, aa * p.di, has_re);    // solve_split.m:61
      if constexpr (SEQ == S_YCLOSEM) {
        const double bb = vm + p.g * cabs2(x);                       // solve_split.m:63
        x = mul_cexp(x, bb * p.dr * 0.5, bb * p.di * 0.5, has_re);   // solve_split.m:64
      }
      v[m] = x;
    }
  }
  if constexpr (SEQ == S_YOPEN || SEQ == S_YMID) {
    fft_pencil<PL, -1, C, true>(v, t, c, lds, twl);
    row_dft16_fwd<E>(v, rf);
    const double2 wa = twy[ia];
#pragma unroll
    for (int m = 0; m < E; ++m) v[m] = cmul(v[m], wa);
  }
#pragma unroll
  for (int m = 0; m < E; ++m) p.psi[base + (unsigned)m * step] = v[m];

  if constexpr (SEQ == S_YCLOSEM) {
    // block-end reductions of the stored real-space state (solve_split.m:65-77)
    double accA = 0.0, accB = 0.0, accC = 0.0, accK = 0.0;
#pragma unroll
    for (int m = 0; m < E; ++m) {
      const double n = cabs2(v[m]);
      accA += n;
      accB += vv[m] * n;
      accC += n * n;
    }
    fft_pencil<PL, -1, C, true>(v, t, c, lds, twl);
#pragma unroll
    for (int m = 0; m < E; ++m) accK += p.keax[t + T * m] * cabs2(v[m]);
    const double sa = block_sum_w<NW>(accA, red);
    const double sb = block_sum_w<NW>(accB, red);
    const double scc = block_sum_w<NW>(accC, red);
    const double sk = block_sum_w<NW>(accK, red);
    if (tid == 0) {
      const unsigned wg = blockIdx.x;
      p.partials[SL_A * p.maxwg + wg] = sa;
      p.partials[SL_B * p.maxwg + wg] = sb;
      p.partials[SL_C * p.maxwg + wg] = scc;
      p.partials[p.slot_ke * p.maxwg + wg] = sk;
    }
  }
}

#define GPE_CAT2(a, b) a##b
#define GPE_CAT(a, b) GPE_CAT2(a, b)

// Templated on the length as well: this file is compiled once per length, and a helper
// named only by NA would be one merged symbol across the objects.
template <int N, int NA, int PV>
static hipError_t launch_na(int seq, const SweepParams& p, hipStream_t s, int* nblocks) {
  const dim3 grid((16u / NA) * p.gnz), block(ythreads<N, NA, PV>());
  switch (seq) {
    case S_YOPEN: hipLaunchKernelGGL((ysplit_x_kernel<N, S_YOPEN, NA, PV>), grid, block, 0, s, p); break;
    case S_YMID: hipLaunchKernelGGL((ysplit_x_kernel<N, S_YMID, NA, PV>), grid, block, 0, s, p); break;
    case S_YCLOSEM: hipLaunchKernelGGL((ysplit_x_kernel<N, S_YCLOSEM, NA, PV>), grid, block, 0, s, p); break;
    default: return hipErrorInvalidValue;
  }
  if (nblocks) *nblocks = (int)grid.x;
  return hipGetLastError();
}

// One value of a per tile (16 pencils) with Plan<N> (16 elements per thread at 256):
// the fastest of the measured tile/plan combinations.
hipError_t GPE_CAT(launch_ysplit_x_, GPE_N)(int seq, const SweepParams& p, hipStream_t s, int* nblocks) {
  constexpr int N = GPE_N;
  if (p.gny != 256u || p.gnx != (unsigned)N || p.gnz % 8u != 0u || p.gnz == 0u) return hipErrorInvalidValue;
  return launch_na<N, 1, 1>(seq, p, s, nblocks);
}

}  // namespace gpe
