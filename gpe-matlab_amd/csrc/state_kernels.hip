// Block-boundary kernels: second-stage reductions, device-side scalar updates
// (norm, renormalisation, chemical potential) and small elementwise passes.
#include <algorithm>
#include <array>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

#include "fft_dev.h"
#include "phase_dev.h"
#include "reduce_dev.h"
#include "finalize_dev.h"
#include "state.h"
#include "sweep_body.h"

namespace gpe {

__global__ void __launch_bounds__(256) finalize_kernel(const FinParams f) {
  __shared__ double red[4];
  finalize_core(f, red);
}

// groundstate_itp.m:61-62: max | |s psi| - |psi_old| |, then psi_old = s psi.
__global__ void __launch_bounds__(256) itp_check_kernel(const double2* __restrict__ psi, double2* __restrict__ old,
                                                        const DevState* st, long long n, double* partials,
                                                        int maxwg) {
  __shared__ double red[4];
  const double s = st->mu_itp;
  double m = 0.0;
  for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long long)gridDim.x * 256) {
    const double2 x = cscale(psi[i], s);
    const double2 o = old[i];
    m = fmax(m, fabs(hypot(x.x, x.y) - hypot(o.x, o.y)));
    old[i] = x;
  }
  const double r = block_max(m, red);
  if (threadIdx.x == 0) partials[SL_MAX * maxwg + blockIdx.x] = r;
}

// psi *= s (s read from the device state: 0 -> scale, 1 -> mu_itp); resets the
// pending scale when it materialises it.
__global__ void __launch_bounds__(256) scale_kernel(double2* __restrict__ psi, DevState* st, int which,
                                                    long long n) {
  const double s = which == 0 ? st->scale : which == 1 ? st->mu_itp : st->sc[which - 2];
  for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long long)gridDim.x * 256)
    psi[i] = cscale(psi[i], s);
}

__global__ void reset_scale_kernel(DevState* st, int which) {
  if (threadIdx.x == 0) {
    if (which == 0) st->scale = 1.0; else st->mu_itp = 1.0;
  }
}

// out[i] = (exp(er) cos(ei), exp(er) sin(ei)) through the sweep kernels' phase code.
__global__ void phase_selftest_kernel(const double* er, const double* ei, double2* out, int n) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i < n) out[i] = mul_cexp(make_double2(1.0, 0.0), er[i], ei[i], true);
}

hipError_t launch_phase_selftest(const double* er, const double* ei, double2* out, int n, hipStream_t s) {
  hipLaunchKernelGGL(phase_selftest_kernel, dim3((n + 255) / 256), dim3(256), 0, s, er, ei, out, n);
  return hipGetLastError();
}

hipError_t launch_finalize(const FinParams& f, hipStream_t s) {
  hipLaunchKernelGGL(finalize_kernel, dim3(1), dim3(256), 0, s, f);
  return hipGetLastError();
}

hipError_t launch_itp_check(const double2* psi, double2* old, const DevState* st, long long n, double* partials,
                            int maxwg, int nblocks, hipStream_t s) {
  hipLaunchKernelGGL(itp_check_kernel, dim3(nblocks), dim3(256), 0, s, psi, old, st, n, partials, maxwg);
  return hipGetLastError();
}

hipError_t launch_scale(double2* psi, DevState* st, int which, long long n, int nblocks, hipStream_t s,
                        bool reset) {
  hipLaunchKernelGGL(scale_kernel, dim3(nblocks), dim3(256), 0, s, psi, st, which, n);
  if (reset) hipLaunchKernelGGL(reset_scale_kernel, dim3(1), dim3(64), 0, s, st, which);
  return hipGetLastError();
}

// ------------------------------------------------------------ length dispatch
#define GPE_DECL(N) hipError_t launch_sweep_##N(int, bool, int, const SweepParams&, hipStream_t, int*);
GPE_DECL(8) GPE_DECL(12) GPE_DECL(16) GPE_DECL(24) GPE_DECL(32) GPE_DECL(48) GPE_DECL(64) GPE_DECL(96)
GPE_DECL(128) GPE_DECL(192) GPE_DECL(256) GPE_DECL(384) GPE_DECL(512)
#undef GPE_DECL

bool sweep_length_supported(int n) {
  switch (n) {
    case 8: case 12: case 16: case 24: case 32: case 48: case 64: case 96: case 128: case 192: case 256:
    case 384: case 512:
      return true;
    default:
      return false;
  }
}

int plan_variants(int n) {
  switch (n) {
    case 64: case 128: case 256: case 384: case 512: return 2;
    default: return sweep_length_supported(n) ? 1 : 0;
  }
}

int sweep_cols_per_block_phase(int n, int var) {
  if (n == 256) return var ? sweep_cols<256, S_MID, 1>() : sweep_cols<256, S_MID, 0>();
  return sweep_cols_per_block(n, var);
}

int sweep_cols_per_block(int n, int var) {
  if (var == 1 && plan_variants(n) == 2) {
    switch (n) {
      case 64: return sweep_cols<64, S_KIN, 1>();
      case 128: return sweep_cols<128, S_KIN, 1>();
      case 256: return sweep_cols<256, S_KIN, 1>();
      case 384: return sweep_cols<384, S_KIN, 1>();
      case 512: return sweep_cols<512, S_KIN, 1>();
      default: break;
    }
  }
  switch (n) {
    case 8: return 256 / Plan<8>::T;
    case 12: return 256 / Plan<12>::T;
    case 16: return 256 / Plan<16>::T;
    case 24: return 256 / Plan<24>::T;
    case 32: return 256 / Plan<32>::T;
    case 48: return 256 / Plan<48>::T;
    case 64: return 256 / Plan<64>::T;
    case 96: return 256 / Plan<96>::T;
    case 128: return 256 / Plan<128>::T;
    case 192: return 256 / Plan<192>::T;
    case 256: return 256 / Plan<256>::T;
    case 384: return 256 / Plan<384>::T;
    case 512: return 256 / Plan<512>::T;
    default: return 0;
  }
}

int sweep_cols_min(int n) {
  int c = 1 << 30;
  for (int v = 0; v < plan_variants(n); ++v)
    c = std::min(c, std::min(sweep_cols_per_block(n, v), sweep_cols_per_block_phase(n, v)));
  return c;
}

// Process-wide plan wisdom: the measured variant per (dimensions, axis, length).
namespace {
std::mutex g_wisdom_mu;
std::map<std::tuple<int, int, int>, int> g_wisdom;
}  // namespace

int wisdom_get(int ndims, int axis, int n) {
  std::lock_guard<std::mutex> lk(g_wisdom_mu);
  auto it = g_wisdom.find({ndims, axis, n});
  return it == g_wisdom.end() ? -1 : it->second;
}

void wisdom_put(int ndims, int axis, int n, int var) {
  std::lock_guard<std::mutex> lk(g_wisdom_mu);
  g_wisdom[{ndims, axis, n}] = var;
}

bool wisdom_put_if_absent(int ndims, int axis, int n, int var) {
  std::lock_guard<std::mutex> lk(g_wisdom_mu);
  return g_wisdom.emplace(std::make_tuple(ndims, axis, n), var).second;
}

void wisdom_clear() {
  std::lock_guard<std::mutex> lk(g_wisdom_mu);
  g_wisdom.clear();
}

std::vector<std::array<int, 4>> wisdom_entries() {
  std::lock_guard<std::mutex> lk(g_wisdom_mu);
  std::vector<std::array<int, 4>> out;
  for (auto& kv : g_wisdom)
    out.push_back({std::get<0>(kv.first), std::get<1>(kv.first), std::get<2>(kv.first), kv.second});
  return out;
}

#define GPE_DECL(N) hipError_t launch_spinor_sweep_##N(int, const SpinorParams&, hipStream_t, int*);
GPE_DECL(8) GPE_DECL(12) GPE_DECL(16) GPE_DECL(24) GPE_DECL(32) GPE_DECL(48) GPE_DECL(64) GPE_DECL(96)
GPE_DECL(128) GPE_DECL(192) GPE_DECL(256) GPE_DECL(384) GPE_DECL(512)
#undef GPE_DECL

hipError_t launch_spinor_sweep(int n, int seq, const SpinorParams& p, hipStream_t s, int* nblocks) {
  switch (n) {
#define GPE_C(N) case N: return launch_spinor_sweep_##N(seq, p, s, nblocks);
    GPE_C(8) GPE_C(12) GPE_C(16) GPE_C(24) GPE_C(32) GPE_C(48) GPE_C(64) GPE_C(96) GPE_C(128) GPE_C(192)
    GPE_C(256) GPE_C(384) GPE_C(512)
#undef GPE_C
    default:
      return hipErrorInvalidValue;
  }
}

#define GPE_DECL(N) hipError_t launch_ysplit_x_##N(int, const SweepParams&, hipStream_t, int*);
GPE_DECL(64) GPE_DECL(128) GPE_DECL(256)
#undef GPE_DECL

bool ysplit_x_supported(int n) { return n == 64 || n == 128 || n == 256; }

hipError_t launch_ysplit_x(int n, int seq, const SweepParams& p, hipStream_t s, int* nblocks) {
  switch (n) {
    case 64: return launch_ysplit_x_64(seq, p, s, nblocks);
    case 128: return launch_ysplit_x_128(seq, p, s, nblocks);
    case 256: return launch_ysplit_x_256(seq, p, s, nblocks);
    default: return hipErrorInvalidValue;
  }
}

// Potential in the row order of the y-split x sweeps:
// Vp[((z 16 + a) 16 + b) nx + x] = V(y = a + 16 b, x, z), one thread per output element.
__global__ void __launch_bounds__(256) ysplit_permute_v_kernel(const double* __restrict__ V, double* __restrict__ Vp,
                                                               unsigned nx, unsigned total) {
  const unsigned i = blockIdx.x * 256u + threadIdx.x;
  if (i >= total) return;
  const unsigned x = i % nx, b = (i / nx) % 16u, tile = i / (16u * nx);
  const unsigned a = tile % 16u, z = tile / 16u;
  Vp[i] = V[a + 16u * b + 256u * x + 256u * nx * z];
}

hipError_t launch_ysplit_permute_v(const double* V, double* Vp, unsigned nx, unsigned nz, hipStream_t s) {
  const unsigned total = 256u * nx * nz;
  hipLaunchKernelGGL(ysplit_permute_v_kernel, dim3((total + 255u) / 256u), dim3(256), 0, s, V, Vp, nx, total);
  return hipGetLastError();
}

// One tile of 16 pencils x 256 points (entry a + 16 z is point z of pencil a), loaded in
// the pencil layout of the sweeps; the 16-point DFT over a (forward into out_f,
// inverse without 1/16 into out_i) runs in the row layout, and the results return
// through the reverse exchange.
__global__ void __launch_bounds__(256) rowdft_selftest_kernel(const double2* in, double2* out_f, double2* out_i) {
  __shared__ double lds[4096];
  const int tid = threadIdx.x, c = tid & 15, t = tid >> 4;
  double2 f[16], g[16];
#pragma unroll
  for (int m = 0; m < 16; ++m) f[m] = in[c + 16 * (t + 16 * m)];
  row_exchange16<256, 16, true>(f, lds);
#pragma unroll
  for (int m = 0; m < 16; ++m) g[m] = f[m];
  DFT<16, -1>::run(f);
  DFT<16, +1>::run(g);
  row_exchange16<256, 16, false>(f, lds);
  row_exchange16<256, 16, false>(g, lds);
#pragma unroll
  for (int m = 0; m < 16; ++m) {
    out_f[c + 16 * (t + 16 * m)] = f[m];
    out_i[c + 16 * (t + 16 * m)] = g[m];
  }
}

hipError_t launch_rowdft_selftest(const double2* in, double2* out_f, double2* out_i, hipStream_t s) {
  hipLaunchKernelGGL(rowdft_selftest_kernel, dim3(1), dim3(256), 0, s, in, out_f, out_i);
  return hipGetLastError();
}

// splitmix64 finaliser of the index: a uniform double in [0, 1)
__device__ __forceinline__ double hash_unit(unsigned long long i) {
  unsigned long long x = i * 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  x ^= x >> 31;
  return (double)(x >> 11) * 0x1.0p-53;
}

__global__ void __launch_bounds__(256) tune_fill_kernel(double2* z, long long nz, double* v, long long nv) {
  const long long stride = (long long)gridDim.x * 256;
  for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < nz || i < nv; i += stride) {
    const double u = hash_unit((unsigned long long)i);
    if (i < nz) {
      double sn, cs;
      sincospi(2.0 * u, &sn, &cs);
      z[i] = make_double2(cs, sn);
    }
    if (v && i < nv) v[i] = 10.0 * u;
  }
}

hipError_t launch_tune_fill(double2* z, long long nz, double* v, long long nv, hipStream_t s) {
  const long long n = nz > nv ? nz : nv;
  if (n <= 0) return hipSuccess;
  const long long want = (n + 255) / 256;
  const unsigned blocks = (unsigned)(want < 8192 ? want : 8192);
  hipLaunchKernelGGL(tune_fill_kernel, dim3(blocks), dim3(256), 0, s, z, nz, v, v ? nv : 0ll);
  return hipGetLastError();
}

__device__ __forceinline__ double scaled(double x, double s) { return x * s; }
__device__ __forceinline__ double2 scaled(double2 x, double s) { return make_double2(x.x * s, x.y * s); }

// One 32 x 32 tile of the (y, last) matrix of one middle index per work-group
// (32 x 8 threads): loads run along the input's contiguous index and stores along
// the output's, through a padded LDS tile.
template <class T, bool TO_COL>
__global__ void __launch_bounds__(256) reorder_kernel(const T* __restrict__ in, T* __restrict__ out, unsigned ny,
                                                      unsigned nmid, unsigned nlast, double scale) {
  __shared__ T tile[32][33];
  const unsigned tx = threadIdx.x & 31u, ty = threadIdx.x >> 5;
  const unsigned z0 = blockIdx.x * 32u, y0 = blockIdx.y * 32u, im = blockIdx.z;
  const bool sc = scale != 1.0;
  if constexpr (TO_COL) {
    for (unsigned r = ty; r < 32u; r += 8u) {   // C order: z fastest
      const unsigned iy = y0 + r, iz = z0 + tx;
      if (iy < ny && iz < nlast) tile[r][tx] = in[((size_t)iy * nmid + im) * nlast + iz];
    }
    __syncthreads();
    for (unsigned r = ty; r < 32u; r += 8u) {   // column-major: y fastest
      const unsigned iz = z0 + r, iy = y0 + tx;
      if (iy < ny && iz < nlast) out[iy + (size_t)ny * (im + (size_t)nmid * iz)] = tile[tx][r];
    }
  } else {
    for (unsigned r = ty; r < 32u; r += 8u) {
      const unsigned iz = z0 + r, iy = y0 + tx;
      if (iy < ny && iz < nlast) tile[tx][r] = in[iy + (size_t)ny * (im + (size_t)nmid * iz)];
    }
    __syncthreads();
    for (unsigned r = ty; r < 32u; r += 8u) {
      const unsigned iy = y0 + r, iz = z0 + tx;
      if (iy < ny && iz < nlast) {
        const T v = tile[r][tx];
        out[((size_t)iy * nmid + im) * nlast + iz] = sc ? scaled(v, scale) : v;
      }
    }
  }
}

hipError_t launch_reorder(const void* in, void* out, int elem, unsigned ny, unsigned nmid, unsigned nlast,
                          bool to_colmajor, double scale, hipStream_t s) {
  if (ny == 0u || nmid == 0u || nlast == 0u || nmid > 65535u || (elem != 8 && elem != 16)) return hipErrorInvalidValue;
  const dim3 grid((nlast + 31u) / 32u, (ny + 31u) / 32u, nmid), block(256);
  if (elem == 16) {
    const double2* i = static_cast<const double2*>(in);
    double2* o = static_cast<double2*>(out);
    if (to_colmajor)
      hipLaunchKernelGGL((reorder_kernel<double2, true>), grid, block, 0, s, i, o, ny, nmid, nlast, scale);
    else
      hipLaunchKernelGGL((reorder_kernel<double2, false>), grid, block, 0, s, i, o, ny, nmid, nlast, scale);
  } else {
    const double* i = static_cast<const double*>(in);
    double* o = static_cast<double*>(out);
    if (to_colmajor)
      hipLaunchKernelGGL((reorder_kernel<double, true>), grid, block, 0, s, i, o, ny, nmid, nlast, scale);
    else
      hipLaunchKernelGGL((reorder_kernel<double, false>), grid, block, 0, s, i, o, ny, nmid, nlast, scale);
  }
  return hipGetLastError();
}

hipError_t launch_sweep(int n, int seq, bool colfast, const SweepParams& p, hipStream_t s, int* nblocks, int var) {
  switch (n) {
#define GPE_C(N) case N: return launch_sweep_##N(seq, colfast, var, p, s, nblocks);
    GPE_C(8) GPE_C(12) GPE_C(16) GPE_C(24) GPE_C(32) GPE_C(48) GPE_C(64) GPE_C(96) GPE_C(128) GPE_C(192)
    GPE_C(256) GPE_C(384) GPE_C(512)
#undef GPE_C
    default:
      return hipErrorInvalidValue;
  }
}

}  // namespace gpe
