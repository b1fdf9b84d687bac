// Deterministic work-group reductions (fixed association order, so repeated runs
// are bit-identical).  Work-groups are 256 threads = 4 wavefronts of 64.
#pragma once
#include <hip/hip_runtime.h>

namespace gpe {

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
  return v;
}

// Sum over the 256-thread work-group; the result is valid on thread 0.
// `red` is a __shared__ scratch of at least 4 doubles.
__device__ __forceinline__ double block_sum(double v, double* red) {
  v = wave_sum(v);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) red[w] = v;
  __syncthreads();
  double r = 0.0;
  if (threadIdx.x == 0) r = ((red[0] + red[1]) + red[2]) + red[3];
  return r;
}

__device__ __forceinline__ double block_max(double v, double* red) {
  v = wave_max(v);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) red[w] = v;
  __syncthreads();
  double r = 0.0;
  if (threadIdx.x == 0) r = fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
  return r;
}

}  // namespace gpe
