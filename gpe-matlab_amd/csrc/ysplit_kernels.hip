// x sweeps of the y-split 3D step for one x length (-DGPE_N=<nx>): S_YOPEN, S_YMID,
// S_YCLOSEM (solve_split.m:55-56, 60-61, 60-65), their block-end forms S_YCLOSEB and
// S_YOPENB (below), the stochastic S_YSOPEN, S_YSMID
// (SGPEtask/solve_split.m:55-59, 69 then 59) and the rotating-frame S_YROPEN, S_YRMID,
// S_YRCLOSEM (the same with the x half steps of ssft_kin_step.m:22-25).
//
// The y axis has 256 points, y = a + 16 b.  A work-group takes the tile of one z
// plane and one value of a with every b: C = 16 x-pencils, pencil c <-> b slot c,
// T = nx/16 threads each, 16 elements per thread.  The x transforms run in that pencil
// layout (tid = t C + c); the tile then moves through LDS to the row layout (thread
// q = x holds the 16 b slots of one (a, x), row_exchange16), where the y factor over b
// is a 16-point DFT in registers and the phase step applies.  Between the sweeps the b
// slot r holds the spectral index k1 = bitrev4(r), times the twiddle W256^{a k1} of the
// 16 x 16 split; the z sweep applies the factor over a.  From the opening x sweep of
// an outer iteration to the closing one of its last block, the state lives in a second
// array (psiy), where the 256 values of every (x, z) column are in the block order
// ysplit_pos(a, r) (rowdft_dev.h) instead of y = a + 16 r.
//
// The potential is read in row order (launch_ysplit_permute_v): for each b, the nx
// values of the tile's rows are contiguous.
//
// Block ends inside an outer iteration: S_YCLOSEB is S_YMID with the closing phase and
// the real-space sums, and stores the mixed form as S_YMID does; the z sweep S_YKEZ
// then sums the kinetic term over the full spectrum, and the next block opens with
// S_YOPENB, S_YMID with the pending scale and the block-open half phase.
#include "philox_dev.h"
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

constexpr int br4c(int r) { return ((r & 1) << 3) | ((r & 2) << 1) | ((r & 4) >> 1) | ((r & 8) >> 3); }

template <int N>
constexpr int ythreads() { return 16 * (N / 16); }

template <int N, int SEQ>
__global__ void __launch_bounds__((ythreads<N>())) __attribute__((amdgpu_waves_per_eu(2)))
ysplit_x_kernel(const SweepParams p) {
  using PL = Plan<N>;
  constexpr int T = PL::T, E = PL::E, C = 16, NT = C * T, NW = NT / 64;
  constexpr bool OPEN = SEQ == S_YOPEN || SEQ == S_YSOPEN, NOISE = SEQ == S_YSOPEN || SEQ == S_YSMID;
  constexpr bool ROT = SEQ == S_YROPEN || SEQ == S_YRMID || SEQ == S_YRCLOSEM;
  static_assert(E == 16 && NT % 64 == 0, "16 elements per thread, whole wavefronts");
  // LDS exchanges move real and imaginary parts together (16-byte accesses): half the
  // LDS instructions and barriers of the split exchanges of the other sweeps.  At two
  // groups per CU (the register limit) the 72 KB per group fit; measured at 256^3 the
  // mid-block sweep took 155 us against 161 us with split exchanges.
  __shared__ double2 lds[Lds<N, C, true>::SIZE];
  __shared__ double2 twl[N];
  __shared__ double2 twy[256];
  __shared__ double red[NW];
  const int tid = threadIdx.x;
  for (int i = tid; i < N; i += NT) twl[i] = p.tw[i];
  for (int i = tid; i < 256; i += NT) twy[i] = p.twy[i];
  __syncthreads();

  // the 16 tiles of a z plane go to one XCD (group b runs on XCD b % 8); gridDim = 16 nz
  const unsigned bi = blockIdx.x, xcd = bi & 7u, s8 = bi >> 3;
  const unsigned ar = s8 % 16u, z = (s8 / 16u) * 8u + xcd;
  const unsigned ny = p.gny, nx = p.gnx;
  // pencil layout
  constexpr bool OPEN_SEQ = SEQ == S_YOPEN || SEQ == S_YSOPEN || SEQ == S_YROPEN;
  constexpr bool CLOSE_SEQ = SEQ == S_YCLOSEM || SEQ == S_YRCLOSEM;
  const int c = tid % C, t = tid / C;
  const unsigned colz = ny * (unsigned)t + ny * nx * z;
  // y-split 3D steps: the opening sweep of an outer iteration reads psi in natural order
  // and the closing sweep of its last block writes it; everything in between, block
  // ends inside the iteration included (yblk_in / yblk_out), is in block order in psiy.
  // (In place, the two orders would race: a tile's natural positions are block
  // positions of other tiles.)  A natural-order tile is 16 y values 16 apart, each
  // 16-byte piece in a line shared with 7 other tiles: the block order is the cheaper
  // hand-over.
  const bool nat_in = OPEN_SEQ && !p.yblk_in, nat_out = CLOSE_SEQ && !p.yblk_out;
  const double2* src = nat_in ? p.psi : p.psiy;
  double2* dst = nat_out ? p.psi : p.psiy;
  const unsigned ypos = ysplit_pos(ar, (unsigned)c);
  const unsigned base_in = (nat_in ? ar + 16u * (unsigned)c : ypos) + colz;
  const unsigned base_out = (nat_out ? ar + 16u * (unsigned)c : ypos) + colz;
  const unsigned step = (unsigned)T * ny;
  // row layout: row q = tid is x = tid of the tile's (a, z)
  const unsigned vrow = (z * 16u + ar) * 16u * (unsigned)NT + (unsigned)tid;

  double2 v[E];
#pragma unroll
  for (int m = 0; m < E; ++m) v[m] = src[base_in + (unsigned)m * step];
  // the potential is loaded with the state, before the transform: at two waves per SIMD
  // nothing else would hide its latency at the phase
  // (rotating frame: in the pencil layout, where its phase applies; the row order keeps
  // each pencil's x values contiguous)
  double vv[16];
  if constexpr (ROT) {
    const unsigned vpen = ((z * 16u + ar) * 16u + (unsigned)c) * nx + (unsigned)t;
#pragma unroll
    for (int m = 0; m < 16; ++m) vv[m] = p.V[vpen + (unsigned)(T * m)];
  } else {
#pragma unroll
    for (int b = 0; b < 16; ++b) vv[b] = p.V[vrow + (unsigned)b * (unsigned)NT];
  }

  const bool has_re = p.dr != 0.0;
  const double mu = p.st->mu;
  double2 u[16];   // row layout, b natural (real space) or k1 natural (spectral)

  if constexpr (ROT) {
    // Rotating frame (set_kinop.m:5-7, ssft_kin_step.m:22-25).  Between the sweeps the
    // state is real in x and carries the y factor over b; the x half steps Kx =
    // forward x, x rtab[y][kx] (incl. 1/nx), inverse x run in the pencil layout around
    // the phase, and the y factor over b in the row layout before and after them.
    const unsigned yrow = ar + 16u * (unsigned)c;
    const double2* __restrict__ rrow = p.rtab + (size_t)yrow * N;
    if constexpr (SEQ == S_YROPEN) {
      const double sc = p.st->scale;
#pragma unroll
      for (int m = 0; m < E; ++m) {
        const double2 x = cscale(v[m], sc);
        const double aa = -(vv[m] - mu) - p.g * cabs2(x);
        v[m] = mul_cexp(x, aa * p.dr * 0.5, aa * p.di * 0.5, has_re);   // solve_split.m:55-56
      }
      table_kin<PL, C>(v, t, c, lds, twl, rrow);                   // ssft_kin_step.m:23
    } else {
      row_exchange16<N, C, true>(v, lds);
#pragma unroll
      for (int k1 = 0; k1 < 16; ++k1) u[k1] = cmulc(v[br4c(k1)], twy[ar * (unsigned)k1]);
      DFT<16, +1>::run(u);
#pragma unroll
      for (int b = 0; b < 16; ++b) v[b] = u[b];
      row_exchange16<N, C, false>(v, lds);
      table_kin<PL, C>(v, t, c, lds, twl, rrow);                   // ssft_kin_step.m:25
#pragma unroll
      for (int m = 0; m < E; ++m) {
        const double vm = vv[m] - mu;
        const double aa = -vm - p.g * cabs2(v[m]);                        // solve_split.m:60
        double2 x = mul_cexp(v[m], aa * p.dr, aa * p.di, has_re);          // solve_split.m:61
        if constexpr (SEQ == S_YRCLOSEM) {
          const double bb = vm + p.g * cabs2(x);                           // solve_split.m:63
          x = mul_cexp(x, bb * p.dr * 0.5, bb * p.di * 0.5, has_re);       // solve_split.m:64
        }
        v[m] = x;
      }
      if constexpr (SEQ == S_YRMID) table_kin<PL, C>(v, t, c, lds, twl, rrow);   // next step's 23
    }
    if constexpr (SEQ != S_YRCLOSEM) {
      row_exchange16<N, C, true>(v, lds);
#pragma unroll
      for (int b = 0; b < 16; ++b) u[b] = v[b];
      DFT<16, -1>::run(u);
#pragma unroll
      for (int k1 = 0; k1 < 16; ++k1) v[br4c(k1)] = cmul(u[k1], twy[ar * (unsigned)k1]);
      row_exchange16<N, C, false>(v, lds);
      const unsigned bo = opaque(base_out);   // formed here (see the store of the other sweeps)
#pragma unroll
      for (int m = 0; m < E; ++m) dst[bo + (unsigned)m * step] = v[m];
    } else {
      // block-end sums of the stored real-space state (solve_split.m:65-77), the x
      // kinetic term with the rotating-frame cross term of the pencil's y
      const unsigned bo = opaque(base_out);   // formed here (see the store of the other sweeps)
#pragma unroll
      for (int m = 0; m < E; ++m) dst[bo + (unsigned)m * step] = v[m];
      double accA = 0.0, accB = 0.0, accC = 0.0, accK = 0.0;
#pragma unroll
      for (int m = 0; m < E; ++m) {
        const double n = cabs2(v[m]);
        accA += n;
        accB += vv[m] * n;
        accC += n * n;
      }
      fft_pencil<PL, -1, C>(v, t, c, lds, twl);
      const double rc = p.rcoord[yrow];
#pragma unroll
      for (int m = 0; m < E; ++m) accK += ke_weight(p, t + T * m, rc) * cabs2(v[m]);
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
    return;
  }

  if constexpr (OPEN) {
    // real-space state: slot r holds b = r.  solve_split.m:55-56 with the pending
    // renormalisation (solve_split.m:72-73)
    row_exchange16<N, C, true>(v, lds);
    const double sc = p.st->scale;
#pragma unroll
    for (int b = 0; b < 16; ++b) {
      const double2 x = cscale(v[b], sc);
      const double aa = -(vv[b] - mu) - p.g * cabs2(x);
      u[b] = mul_cexp(x, aa * p.dr * 0.5, aa * p.di * 0.5, has_re);
    }
  }
  // S_YCLOSEB: block-end sums of the real-space state (solve_split.m:65-77), accumulated
  // as the closing phase produces it
  double sumA = 0.0, sumB = 0.0, sumC = 0.0;
  if constexpr (!OPEN) {
    fft_pencil<PL, +1, C>(v, t, c, lds, twl);
    row_exchange16<N, C, true>(v, lds);
    // undo the split twiddle and the y factor over b left by the previous x sweep
#pragma unroll
    for (int k1 = 0; k1 < 16; ++k1) u[k1] = cmulc(v[br4c(k1)], twy[ar * (unsigned)k1]);
    DFT<16, +1>::run(u);
    if constexpr (SEQ == S_YOPENB) {
      // the block-open half phase on the previous block's closing state, with the
      // pending renormalisation (solve_split.m:55-56, 72-73).  S_YCLOSEB stored the
      // forward transforms unscaled (a z sweep would have applied the 1/npts of its
      // kinetic table), so the inverse ones here return 16 nx times the state: the
      // exact power of two 1/(16 nx) joins the scale.
      const double sc = p.st->scale * (1.0 / (16.0 * (double)N));
#pragma unroll
      for (int b = 0; b < 16; ++b) {
        const double2 x = cscale(u[b], sc);
        const double aa = -(vv[b] - mu) - p.g * cabs2(x);
        u[b] = mul_cexp(x, aa * p.dr * 0.5, aa * p.di * 0.5, has_re);
      }
    } else if constexpr (SEQ == S_YCLOSEB) {
#pragma unroll
      for (int b = 0; b < 16; ++b) {
        const double vm = vv[b] - mu;
        const double aa = -vm - p.g * cabs2(u[b]);                  // solve_split.m:60
        double2 x = mul_cexp(u[b], aa * p.dr, aa * p.di, has_re);    // solve_split.m:61
        const double bb = vm + p.g * cabs2(x);                       // solve_split.m:63
        x = mul_cexp(x, bb * p.dr * 0.5, bb * p.di * 0.5, has_re);   // solve_split.m:64
        const double n = cabs2(x);
        sumA += n;
        sumB += vv[b] * n;
        sumC += n * n;
        u[b] = x;
      }
    } else {
#pragma unroll
      for (int b = 0; b < 16; ++b) {
        const double vm = vv[b] - mu;
        const double aa = -vm - p.g * cabs2(u[b]);                  // solve_split.m:60
        double2 x = mul_cexp(u[b], aa * p.dr, aa * p.di, has_re);    // solve_split.m:61
        if constexpr (SEQ == S_YCLOSEM) {
          const double bb = vm + p.g * cabs2(x);                       // solve_split.m:63
          x = mul_cexp(x, bb * p.dr * 0.5, bb * p.di * 0.5, has_re);   // solve_split.m:64
        }
        u[b] = x;
      }
    }
  }
  if constexpr (SEQ == S_YCLOSEB) {
    // block end inside an outer iteration: the kinetic sum comes from the z sweep
    // S_YKEZ over the mixed form this sweep stores, as after any other step
    const double sa = block_sum_w<NW>(sumA, red);
    const double sb = block_sum_w<NW>(sumB, red);
    const double scc = block_sum_w<NW>(sumC, red);
    if (tid == 0) {
      const unsigned wg = blockIdx.x;
      p.partials[SL_A * p.maxwg + wg] = sa;
      p.partials[SL_B * p.maxwg + wg] = sb;
      p.partials[SL_C * p.maxwg + wg] = scc;
    }
  }
  if constexpr (NOISE) {
    // SGPEtask/solve_split.m:59: the noise of this step, one complex normal per point,
    // Philox counter (natural grid index of the point, global step)
    const unsigned pt0 = ar + ny * (unsigned)tid + ny * nx * z;
#pragma unroll
    for (int b = 0; b < 16; ++b) {
      const double2 w = complex_normal(pt0 + 16u * (unsigned)b, p.noise_step, p.noise_seed);
      u[b].x = u[b].x + p.noise_var * w.x;
      u[b].y = u[b].y + p.noise_var * w.y;
    }
  }
  if constexpr (SEQ != S_YCLOSEM) {
    DFT<16, -1>::run(u);
#pragma unroll
    for (int k1 = 0; k1 < 16; ++k1) v[br4c(k1)] = cmul(u[k1], twy[ar * (unsigned)k1]);
    row_exchange16<N, C, false>(v, lds);
    fft_pencil<PL, -1, C>(v, t, c, lds, twl);
    // the closing sweep forms its store offset here, not from the start (with its sums it
    // would otherwise go to scratch); the other sweeps keep it, measured 3 us faster at
    // 256^3 (the mid-block sweep at 245 against 214 registers)
    const unsigned bo = SEQ == S_YCLOSEB ? opaque(base_out) : base_out;
#pragma unroll
    for (int m = 0; m < E; ++m) dst[bo + (unsigned)m * step] = v[m];
  } else {
    // block-end reductions of the real-space state (solve_split.m:65-77), which is
    // stored with slot r = b
    double accA = 0.0, accB = 0.0, accC = 0.0, accK = 0.0;
#pragma unroll
    for (int b = 0; b < 16; ++b) {
      const double n = cabs2(u[b]);
      accA += n;
      accB += vv[b] * n;
      accC += n * n;
      v[b] = u[b];
    }
    row_exchange16<N, C, false>(v, lds);
    const unsigned bo = opaque(base_out);
#pragma unroll
    for (int m = 0; m < E; ++m) dst[bo + (unsigned)m * step] = v[m];
    fft_pencil<PL, -1, C>(v, t, c, lds, twl);
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

hipError_t GPE_CAT(launch_ysplit_x_, GPE_N)(int seq, const SweepParams& p, hipStream_t s, int* nblocks) {
  constexpr int N = GPE_N;
  if (p.gny != 256u || p.gnx != (unsigned)N || p.gnz % 8u != 0u || p.gnz == 0u) return hipErrorInvalidValue;
  const dim3 grid(16u * p.gnz), block(ythreads<N>());
  switch (seq) {
    case S_YOPEN: hipLaunchKernelGGL((ysplit_x_kernel<N, S_YOPEN>), grid, block, 0, s, p); break;
    case S_YMID: hipLaunchKernelGGL((ysplit_x_kernel<N, S_YMID>), grid, block, 0, s, p); break;
    case S_YCLOSEM: hipLaunchKernelGGL((ysplit_x_kernel<N, S_YCLOSEM>), grid, block, 0, s, p); break;
    case S_YSOPEN: hipLaunchKernelGGL((ysplit_x_kernel<N, S_YSOPEN>), grid, block, 0, s, p); break;
    case S_YSMID: hipLaunchKernelGGL((ysplit_x_kernel<N, S_YSMID>), grid, block, 0, s, p); break;
    case S_YROPEN: hipLaunchKernelGGL((ysplit_x_kernel<N, S_YROPEN>), grid, block, 0, s, p); break;
    case S_YRMID: hipLaunchKernelGGL((ysplit_x_kernel<N, S_YRMID>), grid, block, 0, s, p); break;
    case S_YRCLOSEM: hipLaunchKernelGGL((ysplit_x_kernel<N, S_YRCLOSEM>), grid, block, 0, s, p); break;
    case S_YCLOSEB: hipLaunchKernelGGL((ysplit_x_kernel<N, S_YCLOSEB>), grid, block, 0, s, p); break;
    case S_YOPENB: hipLaunchKernelGGL((ysplit_x_kernel<N, S_YOPENB>), grid, block, 0, s, p); break;
    default: return hipErrorInvalidValue;
  }
  if (nblocks) *nblocks = (int)grid.x;
  return hipGetLastError();
}

}  // namespace gpe
