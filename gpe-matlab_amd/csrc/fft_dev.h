// Device-side building blocks of the pencil FFT used by every sweep kernel.
//
// A pencil of length N (complex128) is held by T = N/E threads, E elements per
// thread.  The canonical register layout is "v[m] holds element t + T*m" for
// thread t; a Stockham autosort pass p (radix R, sub-transform size NS = product
// of earlier radices) reads the canonical layout and writes its outputs in place,
// and an LDS exchange between passes restores the canonical layout.  The last
// pass of a transform leaves the canonical layout as well, so an inverse
// transform, a pointwise operation and a forward transform chain in registers
// without any further data movement.
//
// Conventions: forward = e^{-2 pi i jk/N} (DIR = -1), inverse = e^{+2 pi i jk/N}
// (DIR = +1) without the 1/N factor (the caller folds 1/N into a table).
#pragma once
#include <hip/hip_runtime.h>

namespace gpe {

__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// a * conj(b)
__device__ __forceinline__ double2 cmulc(double2 a, double2 b) {
  return make_double2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
__device__ __forceinline__ double2 cscale(double2 a, double s) { return make_double2(a.x * s, a.y * s); }
__device__ __forceinline__ double2 mul_i(double2 a) { return make_double2(-a.y, a.x); }   // i*a
__device__ __forceinline__ double2 mul_mi(double2 a) { return make_double2(a.y, -a.x); }  // -i*a
__device__ __forceinline__ double cabs2(double2 a) { return a.x * a.x + a.y * a.y; }

// An opaque copy: addresses derived from it are recomputed where they are used
// instead of being kept live from the kernel's first loads.
__device__ __forceinline__ unsigned opaque(unsigned x) {
  unsigned y;
  asm volatile("v_mov_b32 %0, %1" : "=v"(y) : "v"(x));
  return y;
}

// The value of the partner lane (lane ^ 1): DPP quad_perm [1, 0, 3, 2].
__device__ __forceinline__ double xchg(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)b, 0xB1, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), 0xB1, 0xF, 0xF, false);
  return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}
__device__ __forceinline__ double2 xchg(double2 v) { return make_double2(xchg(v.x), xchg(v.y)); }

// cos(2*pi*j/48) for j = 0..12 (first quadrant), correctly rounded decimal literals.
__device__ __forceinline__ double q48(int j) {
  switch (j) {
    case 0: return 1.0;
    case 1: return 0.9914448613738104111445575;
    case 2: return 0.9659258262890682867497432;
    case 3: return 0.9238795325112867561281832;
    case 4: return 0.8660254037844386467637232;
    case 5: return 0.7933533402912351645797770;
    case 6: return 0.7071067811865475244008444;
    case 7: return 0.6087614290087206394160975;
    case 8: return 0.5;
    case 9: return 0.3826834323650897717284600;
    case 10: return 0.2588190451025207623488988;
    case 11: return 0.1305261922200515915484062;
    default: return 0.0;
  }
}

// Multiply a by e^{DIR * 2 pi i e / R}; R must divide 48.  With R and e known after
// unrolling, the quadrant logic folds away and the trivial rotations cost no
// multiplications.
template <int DIR>
__device__ __forceinline__ double2 rot(double2 a, int R, int e) {
  int j = ((e % R) + R) % R * (48 / R);
  if (DIR < 0) j = (48 - j) % 48;
  if (j == 0) return a;
  if (j == 12) return mul_i(a);
  if (j == 24) return make_double2(-a.x, -a.y);
  if (j == 36) return mul_mi(a);
  const int q = j / 12, r = j % 12;
  if (r == 6) {
    const double c = 0.7071067811865475244008444;
    // (1+i)c, (-1+i)c, (-1-i)c, (1-i)c for q = 0..3
    if (q == 0) return make_double2((a.x - a.y) * c, (a.x + a.y) * c);
    if (q == 1) return make_double2(-(a.x + a.y) * c, (a.x - a.y) * c);
    if (q == 2) return make_double2((a.y - a.x) * c, -(a.x + a.y) * c);
    return make_double2((a.x + a.y) * c, (a.y - a.x) * c);
  }
  double cr, ci;
  if (q == 0) { cr = q48(r); ci = q48(12 - r); }
  else if (q == 1) { cr = -q48(12 - r); ci = q48(r); }
  else if (q == 2) { cr = -q48(r); ci = -q48(12 - r); }
  else { cr = q48(12 - r); ci = -q48(r); }
  return make_double2(a.x * cr - a.y * ci, a.x * ci + a.y * cr);
}

// ---------------------------------------------------------------- radix DFTs
template <int R, int DIR> struct DFT;

template <int DIR> struct DFT<1, DIR> {
  __device__ __forceinline__ static void run(double2*) {}
};

template <int DIR> struct DFT<2, DIR> {
  __device__ __forceinline__ static void run(double2* x) {
    double2 a = x[0], b = x[1];
    x[0] = cadd(a, b);
    x[1] = csub(a, b);
  }
};

template <int DIR> struct DFT<3, DIR> {
  __device__ __forceinline__ static void run(double2* x) {
    const double s3 = 0.8660254037844386467637232;  // sin(2 pi / 3)
    double2 t1 = cadd(x[1], x[2]);
    double2 t2 = make_double2(x[0].x - 0.5 * t1.x, x[0].y - 0.5 * t1.y);
    double2 d = csub(x[1], x[2]);
    // DIR * i * s3 * d
    double2 u = (DIR < 0) ? make_double2(d.y * s3, -d.x * s3) : make_double2(-d.y * s3, d.x * s3);
    x[0] = cadd(x[0], t1);
    x[1] = cadd(t2, u);
    x[2] = csub(t2, u);
  }
};

template <int DIR> struct DFT<4, DIR> {
  __device__ __forceinline__ static void run(double2* x) {
    double2 s02 = cadd(x[0], x[2]), d02 = csub(x[0], x[2]);
    double2 s13 = cadd(x[1], x[3]), d13 = csub(x[1], x[3]);
    double2 r = (DIR < 0) ? mul_mi(d13) : mul_i(d13);
    x[0] = cadd(s02, s13);
    x[2] = csub(s02, s13);
    x[1] = cadd(d02, r);
    x[3] = csub(d02, r);
  }
};

// DFT of size R1*R2 in registers: n = n1 + R1 n2, k = k2 + R2 k1.
template <int R1, int R2, int DIR>
__device__ __forceinline__ void dft_composite(double2* x) {
  constexpr int R = R1 * R2;
  double2 z[R];
#pragma unroll
  for (int n1 = 0; n1 < R1; ++n1) {
    double2 a[R2];
#pragma unroll
    for (int n2 = 0; n2 < R2; ++n2) a[n2] = x[n1 + R1 * n2];
    DFT<R2, DIR>::run(a);
#pragma unroll
    for (int k2 = 0; k2 < R2; ++k2) z[n1 * R2 + k2] = rot<DIR>(a[k2], R, n1 * k2);
  }
#pragma unroll
  for (int k2 = 0; k2 < R2; ++k2) {
    double2 b[R1];
#pragma unroll
    for (int n1 = 0; n1 < R1; ++n1) b[n1] = z[n1 * R2 + k2];
    DFT<R1, DIR>::run(b);
#pragma unroll
    for (int k1 = 0; k1 < R1; ++k1) x[k2 + R2 * k1] = b[k1];
  }
}

template <int DIR> struct DFT<6, DIR> {
  __device__ __forceinline__ static void run(double2* x) { dft_composite<2, 3, DIR>(x); }
};
template <int DIR> struct DFT<8, DIR> {
  __device__ __forceinline__ static void run(double2* x) { dft_composite<2, 4, DIR>(x); }
};
template <int DIR> struct DFT<12, DIR> {
  __device__ __forceinline__ static void run(double2* x) { dft_composite<3, 4, DIR>(x); }
};
template <int DIR> struct DFT<16, DIR> {
  __device__ __forceinline__ static void run(double2* x) { dft_composite<4, 4, DIR>(x); }
};

// ---------------------------------------------------------------- plans
// Plan<N>: E elements per thread, T = N/E threads per pencil, P passes of radix
// R0..R3 (each radix divides E).
template <int N> struct Plan;
#define GPE_PLAN(N_, E_, P_, A_, B_, C_, D_)                                     \
  template <> struct Plan<N_> {                                                  \
    static constexpr int N = N_, E = E_, T = N_ / E_, P = P_;                    \
    static constexpr int R0 = A_, R1 = B_, R2 = C_, R3 = D_;                     \
    __device__ __host__ static constexpr int radix(int p) {                      \
      return p == 0 ? A_ : p == 1 ? B_ : p == 2 ? C_ : D_;                       \
    }                                                                            \
    __device__ __host__ static constexpr int ns(int p) {                         \
      return p == 0 ? 1 : p == 1 ? A_ : p == 2 ? A_ * B_ : A_ * B_ * C_;          \
    }                                                                            \
  };
GPE_PLAN(8, 8, 1, 8, 1, 1, 1)
GPE_PLAN(12, 12, 1, 12, 1, 1, 1)
GPE_PLAN(16, 16, 1, 16, 1, 1, 1)
GPE_PLAN(24, 12, 2, 12, 2, 1, 1)
GPE_PLAN(32, 16, 2, 16, 2, 1, 1)
GPE_PLAN(48, 12, 3, 4, 4, 3, 1)
GPE_PLAN(64, 16, 2, 16, 4, 1, 1)
GPE_PLAN(96, 12, 3, 12, 4, 2, 1)
GPE_PLAN(128, 16, 2, 16, 8, 1, 1)
GPE_PLAN(192, 12, 3, 12, 4, 4, 1)
GPE_PLAN(256, 16, 2, 16, 16, 1, 1)
GPE_PLAN(384, 12, 4, 12, 4, 4, 2)
GPE_PLAN(512, 16, 3, 16, 16, 2, 1)
#undef GPE_PLAN

// 256 points as 8.8.4 with 8 elements per thread (32 threads per pencil): twice the
// LDS exchanges of 16.16, half the registers.  The x sweeps that carry the phase
// step use it (see SweepPlan): their exp/sincos work then runs at 4 waves per SIMD
// instead of 2 (mid_x 118 -> 110 us at 256^3); the transform-only and kinetic sweeps
// keep 16.16, where the extra exchanges cost more than the occupancy returns.
struct Plan256x8 {
  static constexpr int N = 256, E = 8, T = 32, P = 3;
  static constexpr int R0 = 8, R1 = 8, R2 = 4, R3 = 1;
  __device__ __host__ static constexpr int radix(int p) { return p == 0 ? 8 : p == 1 ? 8 : p == 2 ? 4 : 1; }
  __device__ __host__ static constexpr int ns(int p) { return p == 0 ? 1 : p == 1 ? 8 : p == 2 ? 64 : 256; }
};

// Alternative plans measured by the plan tuner (gpe_tune): fewer elements per thread
// (more threads per pencil, fewer pencils per work-group) for 64, 128 and 512, more
// for 384.  Each radix divides E.
#define GPE_PLAN_NAMED(NAME, N_, E_, P_, A_, B_, C_)                               \
  struct NAME {                                                                    \
    static constexpr int N = N_, E = E_, T = N_ / E_, P = P_;                      \
    static constexpr int R0 = A_, R1 = B_, R2 = C_, R3 = 1;                        \
    __device__ __host__ static constexpr int radix(int p) {                        \
      return p == 0 ? A_ : p == 1 ? B_ : p == 2 ? C_ : 1;                          \
    }                                                                              \
    __device__ __host__ static constexpr int ns(int p) {                           \
      return p == 0 ? 1 : p == 1 ? A_ : p == 2 ? A_ * B_ : A_ * B_ * C_;            \
    }                                                                              \
  };
GPE_PLAN_NAMED(Plan64x8, 64, 8, 2, 8, 8, 1)
GPE_PLAN_NAMED(Plan128x8, 128, 8, 3, 8, 4, 4)
GPE_PLAN_NAMED(Plan384x24, 384, 24, 3, 12, 8, 4)
GPE_PLAN_NAMED(Plan512x8, 512, 8, 3, 8, 8, 8)
#undef GPE_PLAN_NAMED

// Second plan of a length (plan variant 1), where one exists (256 has its own rule).
template <int N> struct AltPlan { static constexpr bool has = false; using type = Plan<N>; };
template <> struct AltPlan<64> { static constexpr bool has = true; using type = Plan64x8; };
template <> struct AltPlan<128> { static constexpr bool has = true; using type = Plan128x8; };
template <> struct AltPlan<384> { static constexpr bool has = true; using type = Plan384x24; };
template <> struct AltPlan<512> { static constexpr bool has = true; using type = Plan512x8; };

// One Stockham pass on the canonical layout.  Butterfly j = t + T*b takes
// elements j + r*N/R (register b + r*B), twiddles them by W_{NS R}^{r (j mod NS)}
// and applies a radix-R DFT; output r of butterfly j belongs at
// (j / NS) * NS * R + j mod NS + r * NS.
template <class PL, int p, int DIR>
__device__ __forceinline__ void stockham_pass(double2 (&v)[PL::E], int t, const double2* __restrict__ tw) {
  constexpr int R = PL::radix(p);
  constexpr int NS = PL::ns(p);
  constexpr int B = PL::E / R;
  constexpr int N = PL::N;
#pragma unroll
  for (int b = 0; b < B; ++b) {
    double2 x[R];
#pragma unroll
    for (int r = 0; r < R; ++r) x[r] = v[b + r * B];
    if (NS > 1) {
      const int j = t + PL::T * b;
      const int k = j % NS;
#pragma unroll
      for (int r = 1; r < R; ++r) {
        double2 w = tw[r * k * (N / (NS * R))];
        x[r] = (DIR < 0) ? cmul(x[r], w) : cmulc(x[r], w);
      }
    }
    DFT<R, DIR>::run(x);
#pragma unroll
    for (int r = 0; r < R; ++r) v[b + r * B] = x[r];
  }
}

// Output index of register (b + r*B) after pass p (see stockham_pass).
template <class PL, int p>
__device__ __forceinline__ int pass_out_index(int t, int b, int r) {
  constexpr int R = PL::radix(p);
  constexpr int NS = PL::ns(p);
  const int j = t + PL::T * b;
  return (j / NS) * NS * R + (j % NS) + r * NS;
}

// LDS exchange helpers.  COLFAST: element i of column c lives at i*C + c (lanes
// are consecutive columns, conflict-free).  TFAST: column c owns a padded row
// of N + N/16 slots (lanes are consecutive t within one column).
// COLFAST with C = 8 or 4: the 32 doubles of one 64-bank line hold RL = 32/C rows, and
// a row's slot in its line is XOR-ed with (i / 8) mod RL.  Without it the first
// Stockham pass of an 8-element plan writes rows 8t + r from the lane groups t of a
// wave, which all fall in the same slot (8- or 16-way bank conflicts); with it they
// spread over the line (2-way, the minimum for 512 bytes per instruction).
template <int N, int C, bool COLFAST>
struct Lds {
  static constexpr int PITCH = N + N / 16 + 1;
  static constexpr int SIZE = COLFAST ? N * C : C * PITCH;
  static constexpr int RL = 32 / C;
  __device__ __forceinline__ static int at(int c, int i) {
    if (COLFAST && (C == 8 || C == 4)) {
      const unsigned u = (unsigned)i;   // rows are non-negative: shifts and masks, no signed division
      return (int)((u / RL) * 32u + (((u % RL) ^ ((u >> 3) % RL)) * (unsigned)C)) + c;
    }
    return COLFAST ? i * C + c : c * PITCH + i + (i >> 4);
  }
};

// Restore the canonical layout after pass p via LDS (doubles, re and im staged
// separately so the LDS footprint is N*C doubles).  Every slot index is XOR-ed with the
// lane's `sw` (0 for a plain sweep); a non-zero sw below 32 keeps the index inside its
// 32-slot line, so it needs a footprint of whole lines.
template <class PL, int p, int C, bool COLFAST>
__device__ __forceinline__ void exchange(double2 (&v)[PL::E], int t, int c, double* lds, unsigned sw = 0) {
  using L = Lds<PL::N, C, COLFAST>;
  constexpr int R = PL::radix(p);
  constexpr int B = PL::E / R;
#pragma unroll
  for (int half = 0; half < 2; ++half) {
    __syncthreads();
#pragma unroll
    for (int b = 0; b < B; ++b)
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const int idx = pass_out_index<PL, p>(t, b, r);
        lds[L::at(c, idx) ^ sw] = half ? v[b + r * B].y : v[b + r * B].x;
      }
    __syncthreads();
#pragma unroll
    for (int m = 0; m < PL::E; ++m) {
      const double val = lds[L::at(c, t + PL::T * m) ^ sw];
      if (half) v[m].y = val; else v[m].x = val;
    }
  }
}

// The same exchange with real and imaginary parts moved together (16-byte LDS accesses at
// the same slots, XOR-ed with sw): half the LDS instructions and barriers, twice the
// footprint (N C complex values).
template <class PL, int p, int C, bool COLFAST>
__device__ __forceinline__ void exchange(double2 (&v)[PL::E], int t, int c, double2* lds, unsigned sw = 0) {
  using L = Lds<PL::N, C, COLFAST>;
  constexpr int R = PL::radix(p);
  constexpr int B = PL::E / R;
  __syncthreads();
#pragma unroll
  for (int b = 0; b < B; ++b)
#pragma unroll
    for (int r = 0; r < R; ++r) lds[L::at(c, pass_out_index<PL, p>(t, b, r)) ^ sw] = v[b + r * B];
  __syncthreads();
#pragma unroll
  for (int m = 0; m < PL::E; ++m) v[m] = lds[L::at(c, t + PL::T * m) ^ sw];
}

// Full transform of one pencil held in canonical layout, complex LDS exchanges.
template <class PL, int DIR, int C, bool COLFAST = true>
__device__ __forceinline__ void fft_pencil(double2 (&v)[PL::E], int t, int c, double2* lds,
                                           const double2* __restrict__ tw, unsigned sw = 0) {
  stockham_pass<PL, 0, DIR>(v, t, tw);
  if constexpr (PL::P > 1) {
    exchange<PL, 0, C, COLFAST>(v, t, c, lds, sw);
    stockham_pass<PL, 1, DIR>(v, t, tw);
  }
  if constexpr (PL::P > 2) {
    exchange<PL, 1, C, COLFAST>(v, t, c, lds, sw);
    stockham_pass<PL, 2, DIR>(v, t, tw);
  }
  if constexpr (PL::P > 3) {
    exchange<PL, 2, C, COLFAST>(v, t, c, lds, sw);
    stockham_pass<PL, 3, DIR>(v, t, tw);
  }
}

// Full transform of one pencil held in canonical layout.
template <class PL, int DIR, int C, bool COLFAST>
__device__ __forceinline__ void fft_pencil(double2 (&v)[PL::E], int t, int c, double* lds,
                                           const double2* __restrict__ tw, unsigned sw = 0) {
  stockham_pass<PL, 0, DIR>(v, t, tw);
  if constexpr (PL::P > 1) {
    exchange<PL, 0, C, COLFAST>(v, t, c, lds, sw);
    stockham_pass<PL, 1, DIR>(v, t, tw);
  }
  if constexpr (PL::P > 2) {
    exchange<PL, 1, C, COLFAST>(v, t, c, lds, sw);
    stockham_pass<PL, 2, DIR>(v, t, tw);
  }
  if constexpr (PL::P > 3) {
    exchange<PL, 2, C, COLFAST>(v, t, c, lds, sw);
    stockham_pass<PL, 3, DIR>(v, t, tw);
  }
}

}  // namespace gpe
