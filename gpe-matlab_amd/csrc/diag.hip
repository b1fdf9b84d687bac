// Vortex-core detection on a 2D slice of the resident state (lib/util/detect_core.m:1-66)
// so that a user_callback can track cores without copying psi to the host.
//
// Five launches on the handle's stream over the ny x nx slice (column-major, y
// contiguous), each work-group owning 1024 consecutive points:
//   prep  : |psi|^2 = abs(psi).^2, the MATLAB del2 of it (centred second differences / 2
//           per axis, linear extrapolation onto the edge points, sum / 2) -> |del2|,
//           angle(psi), and a per-group maximum of |del2|            (detect_core.m:16-19)
//   max   : the maximum over the groups
//   flag  : the phase winding of the plaquette (i, j), (i-1, j), (i-1, j-1), (i, j-1)
//           with circular wrap, counted as #(dif > pi) - #(dif < -pi), kept where
//           |del2| > 0.01 max                                        (detect_core.m:20-27)
//   scan  : exclusive offsets of the per-group counts of positive / negative cores
//   emit  : the cores in column-major (find) order, each one's position either at the
//           plaquette centre or refined to the zero of the bilinear interpolants of
//           Re psi and Im psi on the plaquette                       (detect_core.m:29-64)
// The refinement solves the 2x2 generalised eigenproblem A v = d B v of
// detect_core.m:56-59 in closed form: det(A - d B) = 0 is a quadratic in d and v is
// the null vector of A - d B, taken from its row of larger magnitude.
#include <algorithm>
#include <mutex>

#include "handle.h"

using namespace gpe_detail;

namespace gpe {

constexpr int DC_T = 256;              // threads per work-group
constexpr int DC_PPT = 4;              // points per thread
constexpr int DC_BLK = DC_T * DC_PPT;  // points per work-group
constexpr double DC_PI = 3.141592653589793;

struct DcSlice {
  const double2* psi;  // slice base, ny * nx points
  double scale;        // pending real scale of the stored state
  int ny, nx;
  double hx, hy;       // spacings along columns (x) and rows (y)
};

__device__ inline double dc_f(const DcSlice& s, int i, int j) {
  const double2 v = s.psi[(size_t)j * s.ny + i];
  const double a = hypot(v.x * s.scale, v.y * s.scale);
  return a * a;
}

// MATLAB del2 per axis: ((f+ - f0)/h - (f0 - f-)/h) / (2h) inside, linear
// extrapolation of those values onto the first and last points (n > 3).
__device__ inline double dc_d2(double fm, double f0, double fp, double h) {
  return ((fp - f0) / h - (f0 - fm) / h) / (2.0 * h);
}

template <bool ALONG_X>
__device__ inline double dc_axis(const DcSlice& s, int i, int j) {
  const int n = ALONG_X ? s.nx : s.ny;
  const int k = ALONG_X ? j : i;
  const double h = ALONG_X ? s.hx : s.hy;
  auto f = [&](int kk) { return ALONG_X ? dc_f(s, i, kk) : dc_f(s, kk, j); };
  if (k == 0) {
    const double f0 = f(0), f1 = f(1), f2 = f(2), f3 = f(3);
    return 2.0 * dc_d2(f0, f1, f2, h) - dc_d2(f1, f2, f3, h);
  }
  if (k == n - 1) {
    const double a = f(n - 4), b = f(n - 3), c = f(n - 2), d = f(n - 1);
    return 2.0 * dc_d2(b, c, d, h) - dc_d2(a, b, c, h);
  }
  return dc_d2(f(k - 1), f(k), f(k + 1), h);
}

__device__ inline double dc_wave_max(double v) {
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  return v;
}

__device__ inline int dc_wave_sum(int v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

__global__ void __launch_bounds__(DC_T) dc_prep_kernel(DcSlice s, double* __restrict__ an, double* __restrict__ dd,
                                                       double* __restrict__ pmax) {
  __shared__ double red[DC_T / 64];
  const long long n = (long long)s.ny * s.nx;
  const long long base = (long long)blockIdx.x * DC_BLK;
  double m = 0.0;
  for (int k = 0; k < DC_PPT; ++k) {
    const long long p = base + k * DC_T + threadIdx.x;
    if (p >= n) break;
    const int i = (int)(p % s.ny), j = (int)(p / s.ny);
    const double gx = dc_axis<true>(s, i, j);
    const double gy = dc_axis<false>(s, i, j);
    const double v = fabs((gx + gy) / 2.0);
    dd[p] = v;
    m = fmax(m, v);
    const double2 c = s.psi[p];
    an[p] = atan2(c.y * s.scale, c.x * s.scale);
  }
  m = dc_wave_max(m);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    double r = red[0];
    for (int w = 1; w < DC_T / 64; ++w) r = fmax(r, red[w]);
    pmax[blockIdx.x] = r;
  }
}

__global__ void __launch_bounds__(DC_T) dc_max_kernel(const double* __restrict__ pmax, int nb, double* __restrict__ dmax) {
  __shared__ double red[DC_T / 64];
  double m = 0.0;
  for (int b = threadIdx.x; b < nb; b += DC_T) m = fmax(m, pmax[b]);
  m = dc_wave_max(m);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    double r = red[0];
    for (int w = 1; w < DC_T / 64; ++w) r = fmax(r, red[w]);
    *dmax = r;
  }
}

__global__ void __launch_bounds__(DC_T) dc_flag_kernel(const double* __restrict__ an, const double* __restrict__ dd,
                                                       const double* __restrict__ dmax, int ny, int nx,
                                                       signed char* __restrict__ q, int* __restrict__ cnt) {
  __shared__ int red[2][DC_T / 64];
  const long long n = (long long)ny * nx;
  const long long base = (long long)blockIdx.x * DC_BLK;
  const double thr = 0.01 * dmax[0];
  int cp = 0, cm = 0;
  for (int k = 0; k < DC_PPT; ++k) {
    const long long p = base + k * DC_T + threadIdx.x;
    if (p >= n) break;
    const int i = (int)(p % ny), j = (int)(p / ny);
    const int im = i == 0 ? ny - 1 : i - 1, jm = j == 0 ? nx - 1 : j - 1;
    const double a1 = an[p];
    const double a2 = an[(size_t)j * ny + im];
    const double a3 = an[(size_t)jm * ny + im];
    const double a4 = an[(size_t)jm * ny + i];
    const double d1 = a1 - a4, d2 = a2 - a1, d3 = a3 - a2, d4 = a4 - a3;
    int c = (d1 > DC_PI) + (d2 > DC_PI) + (d3 > DC_PI) + (d4 > DC_PI) -
            ((d1 < -DC_PI) + (d2 < -DC_PI) + (d3 < -DC_PI) + (d4 < -DC_PI));
    if (!(dd[p] > thr)) c = 0;
    q[p] = (signed char)c;
    cp += c > 0;
    cm += c < 0;
  }
  cp = dc_wave_sum(cp);
  cm = dc_wave_sum(cm);
  if ((threadIdx.x & 63) == 0) {
    red[0][threadIdx.x >> 6] = cp;
    red[1][threadIdx.x >> 6] = cm;
  }
  __syncthreads();
  if (threadIdx.x < 2) {
    int r = 0;
    for (int w = 0; w < DC_T / 64; ++w) r += red[threadIdx.x][w];
    cnt[2 * blockIdx.x + threadIdx.x] = r;
  }
}

// Exclusive offsets of the (positive, negative) counts over the groups, and the totals.
// One work-group; thread t scans a contiguous chunk of groups.
__global__ void __launch_bounds__(DC_T) dc_scan_kernel(const int* __restrict__ cnt, int nb, int* __restrict__ offs,
                                                       int* __restrict__ tot) {
  __shared__ int sp[DC_T], sm[DC_T];
  const int chunk = (nb + DC_T - 1) / DC_T;
  const int b0 = threadIdx.x * chunk;
  const int b1 = min(nb, b0 + chunk);
  int lp = 0, lm = 0;
  for (int b = b0; b < b1; ++b) {
    lp += cnt[2 * b];
    lm += cnt[2 * b + 1];
  }
  sp[threadIdx.x] = lp;
  sm[threadIdx.x] = lm;
  __syncthreads();
  for (int o = 1; o < DC_T; o <<= 1) {   // inclusive Hillis-Steele scan
    const int vp = threadIdx.x >= o ? sp[threadIdx.x - o] : 0;
    const int vm = threadIdx.x >= o ? sm[threadIdx.x - o] : 0;
    __syncthreads();
    sp[threadIdx.x] += vp;
    sm[threadIdx.x] += vm;
    __syncthreads();
  }
  int rp = sp[threadIdx.x] - lp, rm = sm[threadIdx.x] - lm;
  for (int b = b0; b < b1; ++b) {
    offs[2 * b] = rp;
    offs[2 * b + 1] = rm;
    rp += cnt[2 * b];
    rm += cnt[2 * b + 1];
  }
  if (threadIdx.x == DC_T - 1) {
    tot[0] = sp[DC_T - 1];
    tot[1] = sm[DC_T - 1];
  }
}

struct DcCoords {
  const double* xv;  // x of each column (rx(1,:))
  const double* yv;  // y of each row (ry(:,1))
  double dx, dy;     // detect_core.m:14-15
};

// s = v(1)/v(2) for the null vector of M = A - d B, d = dr + i di, from the row of
// larger magnitude; returns the real part (MATLAB compares complex values by it).
__device__ inline double dc_ratio(const double A[4], const double B[4], double dr, double di) {
  const double m11r = A[0] - dr * B[0], m11i = -di * B[0];
  const double m12r = A[1] - dr * B[1], m12i = -di * B[1];
  const double m21r = A[2] - dr * B[2], m21i = -di * B[2];
  const double m22r = A[3] - dr * B[3], m22i = -di * B[3];
  const double row1 = fabs(m11r) + fabs(m11i) + fabs(m12r) + fabs(m12i);
  const double row2 = fabs(m21r) + fabs(m21i) + fabs(m22r) + fabs(m22i);
  // v = (-M12, M11) from row 1, (-M22, M21) from row 2
  const double nr = row1 >= row2 ? -m12r : -m22r, ni = row1 >= row2 ? -m12i : -m22i;
  const double dr2 = row1 >= row2 ? m11r : m21r, di2 = row1 >= row2 ? m11i : m21i;
  if (di == 0.0) return nr / dr2;
  const double den = dr2 * dr2 + di2 * di2;
  return (nr * dr2 + ni * di2) / den;
}

// detect_core.m:54-64 at (0-based) row ix >= 1, column iy >= 1
__device__ inline void dc_refine(const DcSlice& s, const DcCoords& xy, int ix, int iy, double* out) {
  auto ld = [&](int i, int j) {
    const double2 v = s.psi[(size_t)j * s.ny + i];
    return make_double2(v.x * s.scale, v.y * s.scale);
  };
  const double2 p00 = ld(ix, iy), p10 = ld(ix - 1, iy), p01 = ld(ix, iy - 1), p11 = ld(ix - 1, iy - 1);
  const double A[4] = {-(p10.x - p00.x), -p00.x, -(p10.y - p00.y), -p00.y};
  const double B[4] = {p00.x - p10.x - p01.x + p11.x, p01.x - p00.x, p00.y - p10.y - p01.y + p11.y, p01.y - p00.y};
  // det(A - d B) = a d^2 + b d + c
  const double a = B[0] * B[3] - B[1] * B[2];
  const double b = -(A[0] * B[3] + A[3] * B[0] - A[1] * B[2] - A[2] * B[1]);
  const double c = A[0] * A[3] - A[1] * A[2];
  double dr[2], di[2];
  int nroot = 0;
  if (a == 0.0) {
    if (b != 0.0) {
      dr[0] = -c / b;
      di[0] = 0.0;
      nroot = 1;
    }
  } else {
    const double disc = b * b - 4.0 * a * c;
    if (disc >= 0.0) {
      const double qq = -0.5 * (b + copysign(sqrt(disc), b));
      if (qq != 0.0) {
        const double r1 = qq / a, r2 = c / qq;
        dr[0] = fmin(r1, r2);
        dr[1] = fmax(r1, r2);
        di[0] = di[1] = 0.0;
        nroot = 2;
      }
    } else {
      dr[0] = -b / (2.0 * a);
      di[0] = sqrt(-disc) / (2.0 * a);
      nroot = 1;   // the conjugate root has the same real parts
    }
  }
  out[0] = 0.0;
  out[1] = 0.0;
  for (int r = 0; r < nroot; ++r) {
    const double d = dr[r];
    const double sr = dc_ratio(A, B, dr[r], di[r]);
    if (d > 0.0 && d < 1.0 && sr > 0.0 && sr < 1.0) {
      out[0] = xy.xv[iy] - xy.dy * d;
      out[1] = xy.yv[ix] - xy.dx * sr;
      return;
    }
  }
}

__global__ void __launch_bounds__(DC_T) dc_emit_kernel(DcSlice s, DcCoords xy, const signed char* __restrict__ q,
                                                       const int* __restrict__ offs, int refine, int cap,
                                                       double* __restrict__ outp, double* __restrict__ outm) {
  __shared__ int wp[DC_T / 64], wm[DC_T / 64];
  const long long n = (long long)s.ny * s.nx;
  const long long base = (long long)blockIdx.x * DC_BLK;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const unsigned long long below = (1ull << lane) - 1ull;
  int runp = offs[2 * blockIdx.x], runm = offs[2 * blockIdx.x + 1];
  for (int k = 0; k < DC_PPT; ++k) {
    const long long p = base + k * DC_T + threadIdx.x;
    const int c = p < n ? q[p] : 0;
    const unsigned long long bp = __ballot(c > 0), bm = __ballot(c < 0);
    if (lane == 0) {
      wp[wave] = __popcll(bp);
      wm[wave] = __popcll(bm);
    }
    __syncthreads();
    int prep = 0, prem = 0, totp = 0, totm = 0;
    for (int w = 0; w < DC_T / 64; ++w) {
      if (w < wave) {
        prep += wp[w];
        prem += wm[w];
      }
      totp += wp[w];
      totm += wm[w];
    }
    if (c != 0) {
      const bool pos = c > 0;
      const int r = pos ? runp + prep + __popcll(bp & below) : runm + prem + __popcll(bm & below);
      if (r < cap) {
        const int i = (int)(p % s.ny), j = (int)(p / s.ny);
        double* o = (pos ? outp : outm) + 2 * (size_t)r;
        if (refine > 0) {
          double t[2];
          dc_refine(s, xy, i < 1 ? 1 : i, j < 1 ? 1 : j, t);
          o[0] = t[0];
          o[1] = t[1];
        } else {
          o[0] = xy.xv[j] - xy.dx / 2.0;
          o[1] = xy.yv[i] - (pos ? xy.dy : xy.dx) / 2.0;   // detect_core.m:39 and 48
        }
      }
    }
    runp += totp;
    runm += totm;
    __syncthreads();
  }
}

}  // namespace gpe

namespace {

// Scratch of gpe_detect_core_noalloc: the slice's intermediate fields (angle, del2,
// plaquette flags, group maxima, counts, offsets) and the compacted output, in a
// fixed device array instead of the handle's grown allocation.  Each GPU the library
// runs on holds its own copy until the process ends.  Host code never uses the
// array's name as an address: dc_scratch() looks up the device address once per
// device with hipGetSymbolAddress and every launch and copy takes that pointer.
// No part needs clearing between calls: every launch writes each value the next one
// reads.  Calls that overlap on one GPU (two threads, or handles on other streams)
// share the array and race.
constexpr size_t DC_SCRATCH_BYTES = (size_t)GPE_DC_NOALLOC_SCRATCH_MIB << 20;
static __device__ __align__(256) char scratch[DC_SCRATCH_BYTES];

constexpr int DC_MAX_DEVICES = 64;
std::mutex dc_scratch_mu;
char* dc_scratch_ptr[DC_MAX_DEVICES] = {};

// Device address of `scratch` on the current device h->device (looked up on the
// first call on that device).
int dc_scratch(gpe_handle* h, char** out) {
  if (h->device < 0 || h->device >= DC_MAX_DEVICES) return fail(h, GPE_EINVAL, "device %d out of range", h->device);
  std::lock_guard<std::mutex> lock(dc_scratch_mu);
  if (!dc_scratch_ptr[h->device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(h, GPE_EHIP, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
    dc_scratch_ptr[h->device] = static_cast<char*>(p);
  }
  *out = dc_scratch_ptr[h->device];
  return GPE_OK;
}

size_t dc_align(size_t b) { return (b + 255) & ~(size_t)255; }

// Runs the five launches on the slice `psi` (device) and copies the results out.
// fixed == nullptr: the handle's scratch (dc_buf), grown with hipMalloc when the
// slice needs more; otherwise the fixed_bytes at `fixed`, and a slice that needs more
// is refused (nothing is allocated on this path, on the device or the host).
int dc_run(gpe_handle* h, const double2* psi, bool upload, const double* phi_host, double scale, int ny, int nx,
           const double* x, const double* y, int refine, int cap, double* resp, int* np, double* resm, int* nm,
           char* fixed = nullptr, size_t fixed_bytes = 0) {
  if (!x || !y || !np || !nm || (cap > 0 && (!resp || !resm)) || cap < 0)
    return fail(h, GPE_EINVAL, "null coordinates / outputs or negative capacity");
  if (ny < 4 || nx < 4) return fail(h, GPE_EINVAL, "detect_core needs at least 4 points per axis");
  if (ny % 2 || nx % 2) return fail(h, GPE_EINVAL, "detect_core needs even sizes (rx(1,end/2) must be an index)");
  const long long n = (long long)ny * nx;
  if (n > (1ll << 31) - gpe::DC_BLK) return fail(h, GPE_EINVAL, "slice too large");
  const int nb = (int)((n + gpe::DC_BLK - 1) / gpe::DC_BLK);
  // scratch layout: phi | an | dd | q | xy | pmax | dmax | cnt | offs | out (tot, resp, resm)
  size_t off = 0;
  const size_t o_phi = off;
  off += upload ? dc_align((size_t)n * sizeof(double2)) : 0;
  const size_t o_an = off;
  off += dc_align((size_t)n * sizeof(double));
  const size_t o_dd = off;
  off += dc_align((size_t)n * sizeof(double));
  const size_t o_q = off;
  off += dc_align((size_t)n);
  const size_t o_xy = off;
  off += dc_align((size_t)(nx + ny) * sizeof(double));
  const size_t o_pmax = off;
  off += dc_align((size_t)nb * sizeof(double) + sizeof(double));
  const size_t o_cnt = off;
  off += dc_align((size_t)2 * nb * sizeof(int));
  const size_t o_offs = off;
  off += dc_align((size_t)2 * nb * sizeof(int));
  const size_t o_out = off;
  const size_t out_bytes = 16 + (size_t)4 * cap * sizeof(double);
  off += dc_align(out_bytes);
  HIPCHK(h, hipSetDevice(h->device));
  char* B = fixed;
  if (fixed) {
    if (off > fixed_bytes)
      return fail(h, GPE_EINVAL,
                  "gpe_detect_core_noalloc: a %d x %d slice with capacity %d needs %zu bytes of scratch, more than "
                  "the %zu it holds",
                  ny, nx, cap, off, fixed_bytes);
  } else {
    if (off > h->dc_bytes) {
      (void)hipFree(h->dc_buf);
      h->dc_buf = nullptr;
      h->dc_bytes = 0;
      HIPCHK(h, hipMalloc(&h->dc_buf, off));
      h->dc_bytes = off;
    }
    B = h->dc_buf;
  }
  if (upload) {
    HIPCHK(h, hipMemcpyAsync(B + o_phi, phi_host, (size_t)n * sizeof(double2), hipMemcpyHostToDevice, h->stream));
    psi = reinterpret_cast<const double2*>(B + o_phi);
  }
  double* xyd = reinterpret_cast<double*>(B + o_xy);
  HIPCHK(h, hipMemcpyAsync(xyd, x, (size_t)nx * sizeof(double), hipMemcpyHostToDevice, h->stream));
  HIPCHK(h, hipMemcpyAsync(xyd + nx, y, (size_t)ny * sizeof(double), hipMemcpyHostToDevice, h->stream));
  gpe::DcSlice s{psi, scale, ny, nx, 0.0, 0.0};
  gpe::DcCoords c{xyd, xyd + nx, x[nx / 2] - x[nx / 2 - 1], y[ny / 2] - y[ny / 2 - 1]};   // detect_core.m:14-15
  s.hx = c.dx;
  s.hy = c.dy;
  double* an = reinterpret_cast<double*>(B + o_an);
  double* dd = reinterpret_cast<double*>(B + o_dd);
  signed char* q = reinterpret_cast<signed char*>(B + o_q);
  double* pmax = reinterpret_cast<double*>(B + o_pmax);
  double* dmax = pmax + nb;
  int* cnt = reinterpret_cast<int*>(B + o_cnt);
  int* offs = reinterpret_cast<int*>(B + o_offs);
  int* tot = reinterpret_cast<int*>(B + o_out);
  double* outp = reinterpret_cast<double*>(B + o_out + 16);
  double* outm = outp + 2 * (size_t)cap;
  hipLaunchKernelGGL(gpe::dc_prep_kernel, dim3(nb), dim3(gpe::DC_T), 0, h->stream, s, an, dd, pmax);
  HIPCHK(h, hipGetLastError());
  hipLaunchKernelGGL(gpe::dc_max_kernel, dim3(1), dim3(gpe::DC_T), 0, h->stream, pmax, nb, dmax);
  HIPCHK(h, hipGetLastError());
  hipLaunchKernelGGL(gpe::dc_flag_kernel, dim3(nb), dim3(gpe::DC_T), 0, h->stream, an, dd, dmax, ny, nx, q, cnt);
  HIPCHK(h, hipGetLastError());
  hipLaunchKernelGGL(gpe::dc_scan_kernel, dim3(1), dim3(gpe::DC_T), 0, h->stream, cnt, nb, offs, tot);
  HIPCHK(h, hipGetLastError());
  hipLaunchKernelGGL(gpe::dc_emit_kernel, dim3(nb), dim3(gpe::DC_T), 0, h->stream, s, c, q, offs, refine, cap, outp,
                     outm);
  HIPCHK(h, hipGetLastError());
  if (fixed) {   // counts first, then exactly the rows that fit, straight into the caller's arrays
    int t[4] = {0, 0, 0, 0};
    HIPCHK(h, hipMemcpyAsync(t, B + o_out, sizeof t, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    *np = t[0];
    *nm = t[1];
    const size_t kp = (size_t)std::min(t[0], cap), km = (size_t)std::min(t[1], cap);
    if (kp) HIPCHK(h, hipMemcpyAsync(resp, outp, 2 * kp * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    if (km) HIPCHK(h, hipMemcpyAsync(resm, outm, 2 * km * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    return GPE_OK;
  }
  std::vector<char> host(out_bytes);
  HIPCHK(h, hipMemcpyAsync(host.data(), B + o_out, out_bytes, hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  int t[2];
  std::memcpy(t, host.data(), sizeof t);
  *np = t[0];
  *nm = t[1];
  const double* hp = reinterpret_cast<const double*>(host.data() + 16);
  if (cap > 0) {
    std::memcpy(resp, hp, (size_t)2 * std::min(t[0], cap) * sizeof(double));
    std::memcpy(resm, hp + 2 * (size_t)cap, (size_t)2 * std::min(t[1], cap) * sizeof(double));
  }
  return GPE_OK;
}

}  // namespace

extern "C" {

int gpe_detect_core(gpe_handle* h, int comp, int zslice, const double* x, const double* y, int refine, int cap,
                    double* resp, int* np, double* resm, int* nm) {
  if (!h) return GPE_EINVAL;
  if (comp < 0 || comp > 1 || !h->psi[comp]) return fail(h, GPE_EINVAL, "bad component");
  const int nz = h->ndims == 3 ? h->nz : 1;
  if (zslice < 0 || zslice >= nz) return fail(h, GPE_EINVAL, "z slice %d outside 0..%d", zslice, nz - 1);
  HIPCHK(h, hipSetDevice(h->device));
  DevState st;
  int rc;
  if ((rc = read_state(h, &st))) return rc;
  const double2* psi = h->psi[comp] + (size_t)zslice * h->nx * h->ny;
  return dc_run(h, psi, false, nullptr, st.scale, h->ny, h->nx, x, y, refine, cap, resp, np, resm, nm);
}

int gpe_detect_core_noalloc(gpe_handle* h, int comp, int zslice, const double* x, const double* y, int refine,
                            int cap, double* resp, int* np, double* resm, int* nm) {
  if (!h) return GPE_EINVAL;
  if (comp < 0 || comp > 1 || !h->psi[comp]) return fail(h, GPE_EINVAL, "bad component");
  const int nz = h->ndims == 3 ? h->nz : 1;
  if (zslice < 0 || zslice >= nz) return fail(h, GPE_EINVAL, "z slice %d outside 0..%d", zslice, nz - 1);
  HIPCHK(h, hipSetDevice(h->device));
  char* base = nullptr;
  int rc;
  if ((rc = dc_scratch(h, &base))) return rc;
  DevState st;
  if ((rc = read_state(h, &st))) return rc;
  const double2* psi = h->psi[comp] + (size_t)zslice * h->nx * h->ny;
  return dc_run(h, psi, false, nullptr, st.scale, h->ny, h->nx, x, y, refine, cap, resp, np, resm, nm, base,
                DC_SCRATCH_BYTES);
}

int gpe_detect_core_array(gpe_handle* h, const double* phi, int ny, int nx, const double* x, const double* y,
                          int refine, int cap, double* resp, int* np, double* resm, int* nm) {
  if (!h) return GPE_EINVAL;
  if (!phi) return fail(h, GPE_EINVAL, "null phi");
  return dc_run(h, nullptr, true, phi, 1.0, ny, nx, x, y, refine, cap, resp, np, resm, nm);
}

}  // extern "C"
