// DFT of length 16 across the 16 lanes of a DPP row (lanes 16q .. 16q+15 of a wave).
//
// The 3D sweeps of a 256-point y axis split the y transform as 256 = 16 x 16
// (y = a + 16 b): one 16-point factor runs inside the x sweep and the other inside
// the z sweep.  In both, the 16 values a factor combines sit in the 16 lanes of one
// row, one per lane (one value per register m), so the factor is a radix-2 network
// across lanes: partner values arrive by DPP lane permutations (no LDS traffic).
//
// Forward: decimation in frequency, lane r holds x_r on entry and X_{bitrev4(r)} on
// exit, X_k = sum_r x_r W16^{rk}, W16 = e^{-2 pi i/16}.  Inverse: decimation in
// time with conjugate factors, lane r holds X_{bitrev4(r)} on entry and
// x_r = sum_k X_k W16^{-rk} (no 1/16) on exit, so inverse(forward(x)) = 16 x.
#pragma once
#include <hip/hip_runtime.h>

#include "fft_dev.h"

namespace gpe {

// DPP lane permutation of a double within each row (all rows and banks enabled).
template <int CTRL>
__device__ __forceinline__ double dpp_d(double x) {
  const long long b = __double_as_longlong(x);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)(b & 0xffffffffll), CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, 0xF, 0xF, false);
  return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}

// Value of lane r ^ H of the same row.
template <int H> __device__ __forceinline__ double row_xor(double x);
template <> __device__ __forceinline__ double row_xor<1>(double x) { return dpp_d<0xB1>(x); }   // quad_perm 1,0,3,2
template <> __device__ __forceinline__ double row_xor<2>(double x) { return dpp_d<0x4E>(x); }   // quad_perm 2,3,0,1
// row_half_mirror (r ^ 7), then quad_perm 3,2,1,0 (r ^ 3): r ^ 4
template <> __device__ __forceinline__ double row_xor<4>(double x) { return dpp_d<0x1B>(dpp_d<0x141>(x)); }
template <> __device__ __forceinline__ double row_xor<8>(double x) { return dpp_d<0x128>(x); }  // row_ror 8

template <int H>
__device__ __forceinline__ double2 row_xor2(double2 v) {
  return make_double2(row_xor<H>(v.x), row_xor<H>(v.y));
}

__device__ __forceinline__ int bitrev4(int r) {
  return ((r & 1) << 3) | ((r & 2) << 1) | ((r & 4) >> 1) | ((r & 8) >> 3);
}

// Per-lane factors of the three twiddled stages (half sizes 8, 4, 2): on the upper
// lane of a pair, W_{2H}^{r mod H} = W16^{(r mod H) 8/H}; 1 on the lower lane.
// tw holds W16^q at tw[q * stride16] (e.g. a staged twiddle table in LDS).
// HOLD = true keeps the three factors in registers; false reads each at its stage.
template <bool HOLD>
struct RowFactorsT {
  const double2* tw;
  int r, stride16;
  double2 wh[3];
  __device__ __forceinline__ void init(int lane, const double2* table, int stride) {
    tw = table;
    r = lane;
    stride16 = stride;
    if (HOLD) {
      wh[0] = load<8>();
      wh[1] = load<4>();
      wh[2] = load<2>();
    }
  }
  template <int H>
  __device__ __forceinline__ double2 load() const {
    return (r & H) ? tw[((r % H) * (8 / H)) * stride16] : make_double2(1.0, 0.0);
  }
  template <int H>
  __device__ __forceinline__ double2 w() const {
    if (HOLD) return wh[H == 8 ? 0 : H == 4 ? 1 : 2];
    return load<H>();
  }
  // +1 on the lower lane of the pair at half size H, -1 on the upper
  template <int H>
  __device__ __forceinline__ double s() const { return (r & H) ? -1.0 : 1.0; }
};
using RowFactors = RowFactorsT<false>;

// Decimation-in-frequency stage of half size H: lower lane a + b, upper lane (a - b) W.
template <int E, int H, class RF>
__device__ __forceinline__ void row_dif_stage(double2 (&v)[E], const RF& f) {
  const double sg = f.template s<H>();
  const double2 w = H > 1 ? f.template w<H>() : make_double2(1.0, 0.0);
#pragma unroll
  for (int m = 0; m < E; ++m) {
    const double2 p = row_xor2<H>(v[m]);
    const double2 d = make_double2(fma(sg, v[m].x, p.x), fma(sg, v[m].y, p.y));
    v[m] = (H > 1) ? cmul(d, w) : d;
  }
}

// Decimation-in-time stage of half size H with conjugate factors: the upper lane's
// value times conj(W) is exchanged; lower lane a + b conj(W), upper lane a - b conj(W).
template <int E, int H, class RF>
__device__ __forceinline__ void row_dit_stage(double2 (&v)[E], const RF& f) {
  const double sg = f.template s<H>();
  const double2 w = H > 1 ? f.template w<H>() : make_double2(1.0, 0.0);
#pragma unroll
  for (int m = 0; m < E; ++m) {
    const double2 y = (H > 1) ? cmulc(v[m], w) : v[m];
    const double2 p = row_xor2<H>(y);
    v[m] = make_double2(fma(sg, y.x, p.x), fma(sg, y.y, p.y));
  }
}

template <int E, class RF>
__device__ __forceinline__ void row_dft16_fwd(double2 (&v)[E], const RF& f) {
  row_dif_stage<E, 8, RF>(v, f);
  row_dif_stage<E, 4, RF>(v, f);
  row_dif_stage<E, 2, RF>(v, f);
  row_dif_stage<E, 1, RF>(v, f);
}

template <int E, class RF>
__device__ __forceinline__ void row_dft16_inv(double2 (&v)[E], const RF& f) {
  row_dit_stage<E, 1, RF>(v, f);
  row_dit_stage<E, 2, RF>(v, f);
  row_dit_stage<E, 4, RF>(v, f);
  row_dit_stage<E, 8, RF>(v, f);
}

}  // namespace gpe
