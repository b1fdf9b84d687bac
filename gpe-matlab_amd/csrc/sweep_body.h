// Body of the pencil sweeps, shared by the single-grid kernels
// (sweep_kernels.hip) and the batched ensemble kernels (batch_kernels.hip).
//
// Work-group = 256 threads = C pencils x T threads; every thread holds E
// elements of one pencil in registers (see fft_dev.h).  Grid indices are 32-bit
// (the library limits a grid to 2^27 points), inactive tail pencils load a
// valid pencil and skip the store, and the twiddle table is staged in LDS.
#pragma once
#include <type_traits>

#include "fft_dev.h"
#include "phase_dev.h"
#include "philox_dev.h"
#include "reduce_dev.h"
#include "rowdft_dev.h"
#include "sweep.h"

namespace gpe {

// Sequences that carry the phase step (they run on the x axis).
constexpr bool phase_seq(int s) {
  return s == S_OPEN || s == S_MID || s == S_CLOSE || s == S_CLOSEM || s == S_KEX || s == S_ITPA || s == S_ITPB ||
         s == S_ROPEN || s == S_RMID || s == S_RCLOSEM || s == S_RITPA || s == S_RITPB || s == S_SOPEN ||
         s == S_SMID;
}

// Sequences whose plan the tuner may switch (the solve_split and measurement sweeps);
// every other sequence always runs variant 0.
constexpr bool tunable_seq(int s) {
  return s == S_FWD || s == S_INV || s == S_KIN || s == S_KE || s == S_KEX || s == S_OPEN || s == S_MID ||
         s == S_CLOSE || s == S_CLOSEM || s == S_YKE || s == S_KINE;
}

template <int N> constexpr bool has_variant() { return N == 256 || AltPlan<N>::has; }

// FFT plan of a sweep.  Variant 0: Plan<N>, except 8.8.4 for the phase sweeps of
// 256-point pencils.  Variant 1: AltPlan<N>; for 256 the other of 16.16 and 8.8.4.
template <int N, int SEQ, int VAR = 0> struct SweepPlan {
  using type = typename std::conditional<VAR == 0, Plan<N>, typename AltPlan<N>::type>::type;
};
template <int SEQ, int VAR> struct SweepPlan<256, SEQ, VAR> {
  using type = typename std::conditional<phase_seq(SEQ) != (VAR == 1), Plan256x8, Plan<256>>::type;
};
template <int N, int SEQ, int VAR = 0> constexpr int sweep_cols() { return 256 / SweepPlan<N, SEQ, VAR>::type::T; }

template <int SEQ> struct SeqTraits {
  static constexpr bool LOAD_V = SEQ == S_OPEN || SEQ == S_MID || SEQ == S_CLOSE || SEQ == S_CLOSEM ||
                                 SEQ == S_KEX || SEQ == S_ITPB || SEQ == S_ROPEN || SEQ == S_RMID ||
                                 SEQ == S_RCLOSEM || SEQ == S_RITPB || SEQ == S_SOPEN || SEQ == S_SMID;
  static constexpr bool ROT_KX = SEQ == S_ROPEN || SEQ == S_RMID || SEQ == S_RCLOSEM || SEQ == S_RITPA ||
                                 SEQ == S_RITPB || SEQ == S_RKIN;
  static constexpr bool STORE = !(SEQ == S_KEX || SEQ == S_KE || SEQ == S_YKE || SEQ == S_YKEZ);
};

// Kinetic step of one pencil with a per-(row, index) table: forward, multiply,
// inverse (rtab holds the 1/N of the axis).
template <class PL, int C, bool COLFAST>
__device__ __forceinline__ void table_kin(double2 (&v)[PL::E], int t, int c, double* lds, const double2* twl,
                                          const double2* __restrict__ row) {
  constexpr int T = PL::T, E = PL::E;
  fft_pencil<PL, -1, C, COLFAST>(v, t, c, lds, twl);
#pragma unroll
  for (int m = 0; m < E; ++m) v[m] = cmul(row[t + T * m], v[m]);
  fft_pencil<PL, +1, C, COLFAST>(v, t, c, lds, twl);
}

// The same with complex LDS exchanges (columns consecutive across lanes).
template <class PL, int C>
__device__ __forceinline__ void table_kin(double2 (&v)[PL::E], int t, int c, double2* lds, const double2* twl,
                                          const double2* __restrict__ row) {
  constexpr int T = PL::T, E = PL::E;
  fft_pencil<PL, -1, C>(v, t, c, lds, twl);
#pragma unroll
  for (int m = 0; m < E; ++m) v[m] = cmul(row[t + T * m], v[m]);
  fft_pencil<PL, +1, C>(v, t, c, lds, twl);
}

// Kinetic-energy weight of index i: k_i^2/2/n, plus the rotating-frame cross term.
__device__ __forceinline__ double ke_weight(const SweepParams& p, int i, double rc) {
  return p.rot_w != 0.0 ? fma(p.rot_w * rc, p.k1ax[i], p.keax[i]) : p.keax[i];
}

// One work-group's share of a sweep: pencils tile*C .. tile*C + C-1.  `lds`
// holds Lds<N, C, COLFAST>::SIZE doubles, `twl` 2N entries (twiddles, then the
// S_KIN per-index kinetic factors), `red` 4 doubles;
// reductions go to partials[slot * maxwg + wg].
template <int N, int C, bool COLFAST, int SEQ, int VAR = 0>
__device__ __forceinline__ void sweep_body(const SweepParams& p, double2* __restrict__ psi, unsigned tile,
                                           unsigned wg, double* lds, double2* twl, double* red) {
  using PL = typename SweepPlan<N, SEQ, VAR>::type;
  using TR = SeqTraits<SEQ>;
  constexpr int T = PL::T, E = PL::E;
  static_assert(C * T == 256, "work-group is 256 threads");

  const int tid = threadIdx.x;
  const int c = COLFAST ? tid % C : tid / T;
  const int t = COLFAST ? tid / C : tid % T;
  for (int i = tid; i < N; i += 256) twl[i] = p.tw[i];
  if constexpr (SEQ == S_KIN || SEQ == S_KINE || SEQ == S_KINM || SEQ == S_YKIN || SEQ == S_YKINM) {
    for (int i = tid; i < N; i += 256) twl[N + i] = p.kax[i];
    __syncthreads();  // single-pass plans have no exchange barrier before the first use
  }

  const unsigned cgu = tile * (unsigned)C + (unsigned)c;
  const bool active = cgu < p.ncols;
  const unsigned cg = active ? cgu : p.ncols - 1u;
  unsigned col = (cg % p.A) * p.sa + (cg / p.A) * p.sb;
  if constexpr (SEQ == S_YKIN || SEQ == S_YKINM || SEQ == S_YKEZ) {
    // z pencils of the y-split step: column y + 256 x, y = a + 16 r stored in block order
    col = (col & ~255u) + ysplit_pos(col & 15u, (col >> 4) & 15u);
  }
  const unsigned base = col + (unsigned)t * p.stride;
  const unsigned step = (unsigned)T * p.stride;

  double2 v[E];
  if constexpr (SEQ == S_YKE) {
    // y pencils of a state in block order: value y = a + 16 r of the line is at ysplit_pos(a, r)
    static_assert(N == 256 && !COLFAST, "256-point y pencils, contiguous");
#pragma unroll
    for (int m = 0; m < E; ++m) {
      const unsigned i = (unsigned)(t + T * m);
      v[m] = psi[col + ysplit_pos(i & 15u, i >> 4)];
    }
  } else {
#pragma unroll
    for (int m = 0; m < E; ++m) v[m] = psi[base + (unsigned)m * step];
  }
  double vv[TR::LOAD_V ? E : 1];
  if constexpr (TR::LOAD_V) {
#pragma unroll
    for (int m = 0; m < E; ++m) vv[m] = p.V[base + (unsigned)m * step];
  }

  const bool has_re = p.dr != 0.0;
  double accA = 0.0, accB = 0.0, accC = 0.0, accK = 0.0;
  const unsigned trow = p.trow ? cg % p.trow : 0u;
  const double2* __restrict__ rrow = TR::ROT_KX ? p.rtab + (size_t)trow * N : nullptr;
  const double rc =
      (SEQ == S_KEX || SEQ == S_KE || SEQ == S_YKE || SEQ == S_RCLOSEM) && p.rot_w != 0.0 ? p.rcoord[trow] : 0.0;

  if constexpr (SEQ == S_OPEN) {
    // solve_split.m:55-56 with the pending renormalisation (solve_split.m:72-73) folded in
    const double s = p.st->scale, mu = p.st->mu;
#pragma unroll
    for (int m = 0; m < E; ++m) {
      const double2 x = cscale(v[m], s);
      const double a = -(vv[m] - mu) - p.g * cabs2(x);
      v[m] = mul_cexp(x, a * p.dr * 0.5, a * p.di * 0.5, has_re);
    }
    fft_pencil<PL, -1, C, COLFAST>(v, t, c, lds, twl);
  } else if constexpr (SEQ == S_MID || SEQ == S_CLOSE || SEQ == S_CLOSEM) {
    fft_pencil<PL, +1, C, COLFAST>(v, t, c, lds, twl);
    const double mu = p.st->mu;
#pragma unroll
    for (int m = 0; m < E; ++m) {
      const double vm = vv[m] - mu;
      const double a = -vm - p.g * cabs2(v[m]);                  // solve_split.m:60
      double2 x = mul_cexp(v[m], a * p.dr, a * p.di, has_re);     // solve_split.m:61
      if constexpr (SEQ == S_CLOSE || SEQ == S_CLOSEM) {
        const double b = vm + p.g * cabs2(x);                      // solve_split.m:63
        x = mul_cexp(x, b * p.dr * 0.5, b * p.di * 0.5, has_re);   // solve_split.m:64
      }
      v[m] = x;
    }
    if constexpr (SEQ == S_MID) fft_pencil<PL, -1, C, COLFAST>(v, t, c, lds, twl);
  } else if constexpr (SEQ == S_FWD) {
    if (p.apply_scale) {
      const double s = p.st->scale;
#pragma unroll
      for (int m = 0; m < E; ++m) v[m] = cscale(v[m], s);
    }
    fft_pencil<PL, -1, C, COLFAST>(v, t, c, lds, twl);
  } else if constexpr (SEQ == S_INV) {
    fft_pencil<PL, +1, C, COLFAST>(v, t, c, lds, twl);
  } else if constexpr (SEQ == S_KIN || SEQ == S_KINE) {
    // the per-index factors share the twiddle buffer's second half (staged above)
    const double2 kc = p.kcol[cg];
    fft_pencil<PL, -1, C, COLFAST>(v, t, c, lds, twl);
#pragma unroll
    for (int m = 0; m < E; ++m) v[m] = cmul(cmul(kc, twl[N + t + T * m]), v[m]);
    if constexpr (SEQ == S_KINE) {
      if (active) {
        const double wc = p.kecol[cg];
#pragma unroll
        for (int m = 0; m < E; ++m) accK += (wc + p.keax[t + T * m]) * cabs2(v[m]);
      }
    }
    fft_pencil<PL, +1, C, COLFAST>(v, t, c, lds, twl);
  } else if constexpr (SEQ == S_YKIN || SEQ == S_YKINM) {
    // Every 16 consecutive pencils of the tile are one y block (a = 0..15 of
    // y = a + 16 s).  z transform in the pencil layout, then the tile moves to the row
    // layout (thread q: kz index q / G of the 16 pencils of block g = q % G), where
    // the y factor over a is a 16-point DFT in registers (natural order k2), the
    // kinetic factor of ky = bitrev4(s) + 16 k2 (kcol entry k2 of the block) and kz
    // applies, and the inverse factor runs; back in the pencil layout the inverse z
    // transform follows.
    constexpr int G = C / 16;
    static_assert(COLFAST && C % 16 == 0 && E == 16, "blocks of 16 pencils, 16 elements per thread");
    fft_pencil<PL, -1, C, COLFAST>(v, t, c, lds, twl);
    row_exchange16<N, C, true>(v, lds);
    DFT<16, -1>::run(v);
    const double2 kz = twl[N + tid / G];
    const double2* __restrict__ kc = p.kcol + tile * (unsigned)C + 16u * (unsigned)(tid % G);
    if constexpr (SEQ == S_YKINM) {
      // the cut-off mask of S_KINM (SGPEtask/solve_split.m:39-41): kmax in the same
      // split order as kcol, kz index tid / G
      const int* __restrict__ km = p.kmax + tile * (unsigned)C + 16u * (unsigned)(tid % G);
      const int j = tid / G, aj = j <= N / 2 ? j : N - j;
#pragma unroll
      for (int m = 0; m < E; ++m) {
        const double keep = aj <= km[m] ? 1.0 : 0.0;
        const double2 f = cmul(kc[m], kz);
        v[m] = cmul(make_double2(f.x * keep, f.y * keep), v[m]);
      }
    } else {
#pragma unroll
      for (int m = 0; m < E; ++m) v[m] = cmul(cmul(kc[m], kz), v[m]);
    }
    DFT<16, +1>::run(v);
    row_exchange16<N, C, false>(v, lds);
    fft_pencil<PL, +1, C, COLFAST>(v, t, c, lds, twl);
  } else if constexpr (SEQ == S_YKEZ) {
    // Block end of the y-split step (solve_split.m:74/77, the kinetic part of
    // applyham): the closing x sweep left the state x-spectral with the y factor over b
    // applied, so the forward z transform and the factor over a (the first half of
    // S_YKIN) give the full spectrum; sum (kx^2 + ky^2 + kz^2)/2 |.|^2 / npts over it.
    constexpr int G = C / 16;
    static_assert(COLFAST && C % 16 == 0 && E == 16, "blocks of 16 pencils, 16 elements per thread");
    fft_pencil<PL, -1, C, COLFAST>(v, t, c, lds, twl);
    row_exchange16<N, C, true>(v, lds);
    DFT<16, -1>::run(v);
    const double wz = p.kecol[(size_t)p.ncols + (unsigned)(tid / G)];
    const double* __restrict__ kc = p.kecol + tile * (unsigned)C + 16u * (unsigned)(tid % G);
#pragma unroll
    for (int m = 0; m < E; ++m) accK += (kc[m] + wz) * cabs2(v[m]);
  } else if constexpr (SEQ == S_KEX || SEQ == S_KE || SEQ == S_YKE) {
    if constexpr (SEQ == S_KEX) {
      if (active) {
#pragma unroll
        for (int m = 0; m < E; ++m) {
          const double n = cabs2(v[m]);
          accA += n;
          accB += vv[m] * n;
          accC += n * n;
        }
      }
    }
    fft_pencil<PL, -1, C, COLFAST>(v, t, c, lds, twl);
    if (active) {
#pragma unroll
      for (int m = 0; m < E; ++m) accK += ke_weight(p, t + T * m, rc) * cabs2(v[m]);
    }
  } else if constexpr (SEQ == S_ITPA) {
    fft_pencil<PL, +1, C, COLFAST>(v, t, c, lds, twl);
#pragma unroll
    for (int m = 0; m < E; ++m) {
      const double nl = p.nop[base + (unsigned)m * step];
      v[m] = cscale(v[m], nl);                                     // groundstate_itp.m:42
      accA += active ? cabs2(v[m]) : 0.0;                          // groundstate_itp.m:44
    }
  } else if constexpr (SEQ == S_ITPB) {
    const double mu = p.st->mu_itp;
    const double mu2 = mu * mu;
    const double hdt = -p.dr * 0.5;
#pragma unroll
    for (int m = 0; m < E; ++m) {
      const double n = cabs2(v[m]);                                // nlinop before renormalisation
      const double nl = exp_cw(hdt * (n * p.g * mu2 + vv[m]));     // groundstate_itp.m:76 (36: mu = 1)
      if (active) p.nop[base + (unsigned)m * step] = nl;
      v[m] = cscale(cscale(v[m], mu), nl);                         // groundstate_itp.m:52 then 40
    }
    fft_pencil<PL, -1, C, COLFAST>(v, t, c, lds, twl);
  } else if constexpr (SEQ == S_KINM) {
    // ekk .* mask with mask = kk < ecut (SGPEtask/solve_split.m:39-41).  Along the axis kk
    // grows with |k|, so a column keeps the indices with |k| <= kmax[column] (the host
    // evaluates kk = (kx^2 + ky^2 + kz^2)/2 as the grid sums it to find kmax).
    const double2 kc = p.kcol[cg];
    const int km = p.kmax[cg];
    fft_pencil<PL, -1, C, COLFAST>(v, t, c, lds, twl);
#pragma unroll
    for (int m = 0; m < E; ++m) {
      const int j = t + T * m;
      const int a = j <= N / 2 ? j : N - j;
      // the factor times the 0/1 mask, as ekk .* mask (a select here costs registers)
      const double keep = a <= km ? 1.0 : 0.0;
      const double2 f = cmul(kc, twl[N + j]);
      v[m] = cmul(make_double2(f.x * keep, f.y * keep), v[m]);
    }
    fft_pencil<PL, +1, C, COLFAST>(v, t, c, lds, twl);
  } else if constexpr (SEQ == S_SOPEN || SEQ == S_SMID) {
    // SGPEtask/solve_split.m:55 (open) or 69 (previous step's phase), then 59 (noise)
    if constexpr (SEQ == S_SMID) fft_pencil<PL, +1, C, COLFAST>(v, t, c, lds, twl);
    const double mu = p.st->mu;
    const double f = SEQ == S_SOPEN ? 0.5 : 1.0;
#pragma unroll
    for (int m = 0; m < E; ++m) {
      const double a = -(vv[m] - mu) - p.g * cabs2(v[m]);
      double2 x = mul_cexp(v[m], a * p.dr * f, a * p.di * f, has_re);
      const double2 z = complex_normal(base + (unsigned)m * step, p.noise_step, p.noise_seed);
      x.x = x.x + p.noise_var * z.x;
      x.y = x.y + p.noise_var * z.y;
      v[m] = x;
    }
    fft_pencil<PL, -1, C, COLFAST>(v, t, c, lds, twl);
  } else if constexpr (SEQ == S_RKIN) {
    table_kin<PL, C, COLFAST>(v, t, c, lds, twl, rrow);            // ssft_kin_step.m:24 (2D)
  } else if constexpr (SEQ == S_ROPEN) {
    const double s = p.st->scale, mu = p.st->mu;
#pragma unroll
    for (int m = 0; m < E; ++m) {
      const double2 x = cscale(v[m], s);
      const double a = -(vv[m] - mu) - p.g * cabs2(x);
      v[m] = mul_cexp(x, a * p.dr * 0.5, a * p.di * 0.5, has_re);   // solve_split.m:55-56
    }
    table_kin<PL, C, COLFAST>(v, t, c, lds, twl, rrow);            // ssft_kin_step.m:23
  } else if constexpr (SEQ == S_RMID || SEQ == S_RCLOSEM) {
    table_kin<PL, C, COLFAST>(v, t, c, lds, twl, rrow);            // ssft_kin_step.m:25
    const double mu = p.st->mu;
#pragma unroll
    for (int m = 0; m < E; ++m) {
      const double vm = vv[m] - mu;
      const double a = -vm - p.g * cabs2(v[m]);                  // solve_split.m:60
      double2 x = mul_cexp(v[m], a * p.dr, a * p.di, has_re);     // solve_split.m:61
      if constexpr (SEQ == S_RCLOSEM) {
        const double b = vm + p.g * cabs2(x);                      // solve_split.m:63
        x = mul_cexp(x, b * p.dr * 0.5, b * p.di * 0.5, has_re);   // solve_split.m:64
      }
      v[m] = x;
    }
    if constexpr (SEQ == S_RMID) table_kin<PL, C, COLFAST>(v, t, c, lds, twl, rrow);   // next step's 23
  } else if constexpr (SEQ == S_RITPA) {
    table_kin<PL, C, COLFAST>(v, t, c, lds, twl, rrow);            // ssft_kin_step.m:25
#pragma unroll
    for (int m = 0; m < E; ++m) {
      const double nl = p.nop[base + (unsigned)m * step];
      v[m] = cscale(v[m], nl);                                     // groundstate_itp.m:42
      accA += active ? cabs2(v[m]) : 0.0;                          // groundstate_itp.m:44
    }
  } else if constexpr (SEQ == S_RITPB) {
    const double mu = p.st->mu_itp;
    const double mu2 = mu * mu;
    const double hdt = -p.dr * 0.5;
#pragma unroll
    for (int m = 0; m < E; ++m) {
      const double n = cabs2(v[m]);
      const double nl = exp_cw(hdt * (n * p.g * mu2 + vv[m]));     // groundstate_itp.m:76 (36: mu = 1)
      if (active) p.nop[base + (unsigned)m * step] = nl;
      v[m] = cscale(cscale(v[m], mu), nl);                         // groundstate_itp.m:52 then 40
    }
    table_kin<PL, C, COLFAST>(v, t, c, lds, twl, rrow);            // ssft_kin_step.m:23
  }

  if constexpr (TR::STORE) {
    // the store offset is formed here rather than kept live from the loads
    const unsigned bs = opaque(base);
    if (active) {
#pragma unroll
      for (int m = 0; m < E; ++m) psi[bs + (unsigned)m * step] = v[m];
    }
  }

  if constexpr (SEQ == S_CLOSEM || SEQ == S_RCLOSEM) {
    // block-end reductions of the stored real-space state (solve_split.m:65-77)
    if (active) {
#pragma unroll
      for (int m = 0; m < E; ++m) {
        const double n = cabs2(v[m]);
        accA += n;
        accB += vv[m] * n;
        accC += n * n;
      }
    }
    fft_pencil<PL, -1, C, COLFAST>(v, t, c, lds, twl);
    if (active) {
#pragma unroll
      for (int m = 0; m < E; ++m) accK += ke_weight(p, t + T * m, rc) * cabs2(v[m]);
    }
  }

  // deterministic block reductions
  if constexpr (SEQ == S_KEX || SEQ == S_CLOSEM || SEQ == S_RCLOSEM) {
    const double a = block_sum(accA, red);
    const double b = block_sum(accB, red);
    const double cc = block_sum(accC, red);
    const double k = block_sum(accK, red);
    if (tid == 0) {
      p.partials[SL_A * p.maxwg + wg] = a;
      p.partials[SL_B * p.maxwg + wg] = b;
      p.partials[SL_C * p.maxwg + wg] = cc;
      p.partials[p.slot_ke * p.maxwg + wg] = k;
    }
  } else if constexpr (SEQ == S_KE || SEQ == S_YKE || SEQ == S_YKEZ || SEQ == S_KINE) {
    const double k = block_sum(accK, red);
    const int slot = SEQ == S_KINE && p.psi2 && psi == p.psi2 ? p.slot_ke2 : p.slot_ke;
    if (tid == 0) p.partials[slot * p.maxwg + wg] = k;
  } else if constexpr (SEQ == S_ITPA || SEQ == S_RITPA) {
    const double a = block_sum(accA, red);
    if (tid == 0) p.partials[SL_A * p.maxwg + wg] = a;
  }
}

template <int N, int C, bool COLFAST, int SEQ, int VAR = 0>
__global__ void __launch_bounds__(256) sweep_kernel(const SweepParams p) {
  __shared__ double lds[Lds<N, C, COLFAST>::SIZE];
  __shared__ double2 twl[SEQ == S_KIN || SEQ == S_KINE || SEQ == S_KINM || SEQ == S_YKIN || SEQ == S_YKINM ? 2 * N : N];
  __shared__ double red[4];
  unsigned bid = blockIdx.x, nbk = gridDim.x;
  constexpr bool YZ = SEQ == S_YKIN || SEQ == S_YKINM || SEQ == S_YKEZ;
  double2* psi = YZ ? p.psiy : p.psi;
  if (p.psi2) {   // two arrays: the second half of the grid sweeps psi2 (no reductions)
    nbk >>= 1;
    if (bid >= nbk) {
      bid -= nbk;
      psi = p.psi2;
    }
  }
  const unsigned tile = YZ ? ysplit_ztile(bid, 256u / C)
                        : C < 16                       ? xcd_pair_tile(bid, nbk)
                                                       : bid;
  sweep_body<N, C, COLFAST, SEQ, VAR>(p, psi, tile, tile, lds, twl, red);
}

}  // namespace gpe
