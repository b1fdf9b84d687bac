// 16-point factors of the y transform of the two-sweep 3D step.
//
// The 3D sweeps of a 256-point y axis split the y transform as 256 = 16 x 16
// (y = a + 16 b): the factor over b runs inside the x sweep and the factor over a
// inside the z sweep.  In both, the 16 values a factor combines belong to 16
// neighbouring pencils of the work-group's tile.  The tile moves through LDS from its
// pencil layout to a row layout in which one thread holds the 16 values, and the
// factor is a 16-point DFT in registers (DFT<16>, fft_dev.h); a second exchange
// moves it back.
#pragma once
#include <hip/hip_runtime.h>

#include "fft_dev.h"

namespace gpe {

__device__ __forceinline__ int bitrev4(int r) {
  return ((r & 1) << 3) | ((r & 2) << 1) | ((r & 4) >> 1) | ((r & 8) >> 3);
}

// Storage order of the 256 y values of every (x, z) column in the second state array
// of the two-sweep step, which holds the state while a block runs (written by the
// block's opening x sweep, read by its closing one; at block ends psi holds the state
// in natural order y = a + 16 r).  An x-sweep tile holds one a
// and every r, a z-sweep tile every a and one r, so no order puts a whole 128-byte line
// (8 values) into one tile of both sweeps.  Bit i of YMASK says whether bit i of the
// position within a line is the next bit of r (1) or of a (0); the lines then follow
// in the order of the remaining bits of a, then of r (YMASK = 0 is the natural order).
// YMASK = 2 (bits a0 r0 a1): a line holds a 4-value block of a times two r, so an x
// tile touches each of its lines at two values of one 64-byte half and a z tile at two
// pairs of neighbours.  Measured at 256^3 over the eight masks' classes: the natural
// order gives 210 + 105 us per step (x sweep + z sweep), 4 a per line with a fastest
// 189 + 118, r fastest 161 + 126, this order 161.5 + 121; 2 a per line 147 + 185 or
// 152 + 160, 8 r per line 144 + 257.
constexpr unsigned YMASK = 2u;
constexpr unsigned YR_BITS = (YMASK & 1u) + ((YMASK >> 1) & 1u) + ((YMASK >> 2) & 1u), YA_BITS = 3u - YR_BITS;
static_assert(YMASK < 8u, "a line holds 8 values");

__device__ __forceinline__ unsigned ysplit_pos(unsigned a, unsigned r) {
  unsigned off = 0u, ia = 0u, ir = 0u;
#pragma unroll
  for (unsigned i = 0; i < 3u; ++i) {
    if ((YMASK >> i) & 1u) off |= ((r >> ir++) & 1u) << i;
    else off |= ((a >> ia++) & 1u) << i;
  }
  return off + 8u * ((a >> YA_BITS) + (16u >> YA_BITS) * (r >> YR_BITS));
}

// Tile of work-group b in the y-split z sweep, which has per = 256 / C tiles for each x
// (C pencils per tile) and nx per work-groups.  The 8 XCDs get the work-groups round
// robin; the per tiles of one x, which share the lines of the block order, go to one
// XCD (nx is a multiple of 8).
__device__ __forceinline__ unsigned ysplit_ztile(unsigned b, unsigned per) {
  const unsigned s = b >> 3;
  return ((s / per) * 8u + (b & 7u)) * per + s % per;
}

// A tile of C pencils of length N (256 threads, T = N/16 threads per pencil, 16
// elements each) is G = C/16 blocks of 16 pencils.  It moves between the canonical
// pencil layout -- thread (c = tid % C, t = tid / C) holds element t + T m of pencil
// c = 16 g + a in v[m] -- and the row layout -- thread q = tid holds element z of
// pencil 16 g + a in v[a], q = G z + g -- through LDS, real and imaginary parts in
// turn (C T threads).  Slot of row q, pencil a: 16 q + (a ^ h(q)), h(q) = (q mod 16) ^ (q / 16 mod
// 2).  Every 16 lanes then touch 16 distinct 8-byte bank pairs and every 32 lanes 32:
// the pencil-layout writes cover the 16 values of a for one row q (two rows of
// opposite parity per 32 lanes: two t for G = 1, two g otherwise), the row-layout
// reads consecutive rows q, whose (q mod 2, h(q)) are distinct over 32 rows.  `lds`
// holds C N doubles.
__device__ __forceinline__ int row_slot(int a, int q) { return 16 * q + (a ^ ((q & 15) ^ ((q >> 4) & 1))); }

template <int N, int C, bool TO_ROW>
__device__ __forceinline__ void row_exchange16(double2 (&v)[16], double* lds) {
  constexpr int G = C / 16, T = N / 16;
  static_assert(C % 16 == 0 && (C * T) % 64 == 0, "whole blocks of 16 pencils, whole wavefronts");
  const int tid = threadIdx.x, c = tid % C, t = tid / C, g = c / 16, a = c % 16;
#pragma unroll
  for (int half = 0; half < 2; ++half) {
    __syncthreads();
#pragma unroll
    for (int m = 0; m < 16; ++m) {
      const int slot = TO_ROW ? row_slot(a, (t + T * m) * G + g) : row_slot(m, tid);
      lds[slot] = half ? v[m].y : v[m].x;
    }
    __syncthreads();
#pragma unroll
    for (int m = 0; m < 16; ++m) {
      const double val = lds[TO_ROW ? row_slot(m, tid) : row_slot(a, (t + T * m) * G + g)];
      if (half) v[m].y = val; else v[m].x = val;
    }
  }
}

// The same exchange with real and imaginary parts moved together (16-byte LDS accesses at
// the same slots): half the LDS instructions and barriers, `lds` holds C N complex values.
template <int N, int C, bool TO_ROW>
__device__ __forceinline__ void row_exchange16(double2 (&v)[16], double2* lds) {
  constexpr int G = C / 16, T = N / 16;
  static_assert(C % 16 == 0 && (C * T) % 64 == 0, "whole blocks of 16 pencils, whole wavefronts");
  const int tid = threadIdx.x, c = tid % C, t = tid / C, g = c / 16, a = c % 16;
  __syncthreads();
#pragma unroll
  for (int m = 0; m < 16; ++m) lds[TO_ROW ? row_slot(a, (t + T * m) * G + g) : row_slot(m, tid)] = v[m];
  __syncthreads();
#pragma unroll
  for (int m = 0; m < 16; ++m) v[m] = lds[TO_ROW ? row_slot(m, tid) : row_slot(a, (t + T * m) * G + g)];
}

}  // namespace gpe
