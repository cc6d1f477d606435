// Density filters, projections and optimality-criteria (OC) updates.
//
// Two reference variants, both with volume, move and Emin as parameters:
//  - cone: clipped (non-periodic) cone filter, eta-projection, OC with
//    exponent 0.3, lower bound 0.001 and relative-tolerance bisection
//    (exp/asg_sctop_reproduction_20261105/outer.cu:8-39, reproduction.py:69-118);
//  - gaussian: periodic separable Gaussian in scipy's accumulation order,
//    tanh projection, sqrt OC with fixed bisection count
//    (exp/asg_gpu_resident_20261103/outer.cu:7-37, outer.py:26-44).
// Floating-point contraction is disabled so products and sums round like the
// reference's --fmad=false builds and numpy.
#include <cmath>

#include "layout.h"
#include "pgj_internal.h"

#pragma clang fp contract(off)

namespace pgj {

// bounds layout (cone):     [lo, hi, running (1/0), bisection count]
// bounds layout (gaussian): [lo, hi]

__global__ __launch_bounds__(256) void cone_filter_kernel(const double* x, const int* off, const double* w, int m, int n,
                                                          double* y, const double* bounds, int conditional) {
  const int64_t N = int64_t(n) * n * n;
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= N || (conditional && bounds[2] == 0.0)) return;
  const int xx = int(i % n), yy = int((i / n) % n), zz = int(i / (int64_t(n) * n));
  double sum = 0.0;
  for (int j = 0; j < m; ++j) {
    const int xj = xx + off[3 * j], yj = yy + off[3 * j + 1], zj = zz + off[3 * j + 2];
    if (xj >= 0 && xj < n && yj >= 0 && yj < n && zj >= 0 && zj < n) sum += w[j] * x[(int64_t(zj) * n + yj) * n + xj];
  }
  y[i] = sum;
}

// eta-projection: rho = (tanh(b eta) + tanh(b (t/hs - eta))) / den, with volume partials.
__global__ __launch_bounds__(256) void cone_project_kernel(const double* t, const double* hs, double* ph, double* dp,
                                                           double* a, const double* bounds, int64_t N, double beta,
                                                           double eta, double emin, int conditional, double* vpart) {
  __shared__ double red[8];
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (conditional && bounds[2] == 0.0) return;
  double rho = 0.0;
  if (i < N) {
    const double den = tanh(beta * eta) + tanh(beta * (1 - eta));
    const double q = tanh(beta * (t[i] / hs[i] - eta));
    rho = (tanh(beta * eta) + q) / den;
    ph[i] = rho;
    dp[i] = beta * (1 - q * q) / den;
    a[i] = emin + (1 - emin) * rho * rho * rho;
  }
  if (vpart) {
    const double s = block_sum(rho, red);
    if (threadIdx.x == 0) vpart[blockIdx.x] = s;
  }
}

// g_in = dp (3 (1-Emin) ph^2) e / hs ; dv_in = dp / hs
__global__ __launch_bounds__(256) void cone_prepare_kernel(const double* dp, const double* ph, const double* e,
                                                           const double* hs, double emin, int64_t N, double* gin,
                                                           double* dvin) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const double c = 3 * (1 - emin) * (ph[i] * ph[i]);
  gin[i] = dp[i] * c * e[i] / hs[i];
  dvin[i] = dp[i] / hs[i];
}

// g /= n^3 ; ratio = -g / dv ; reset bounds for the bisection.
__global__ __launch_bounds__(256) void cone_ratio_kernel(double* g, const double* dv, int64_t N, double* ratio,
                                                         double* bounds) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i == 0) {
    bounds[0] = 0.0;
    bounds[1] = 1e9;
    bounds[2] = 1.0;
    bounds[3] = 0.0;
  }
  if (i >= N) return;
  const double gv = g[i] / double(N);
  g[i] = gv;
  ratio[i] = -gv / dv[i];
}

__global__ __launch_bounds__(256) void cone_candidate_kernel(const double* x, const double* ratio, const double* bounds,
                                                             double move, double* y, int64_t N) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= N || bounds[2] == 0.0) return;
  const double mid = 0.5 * (bounds[0] + bounds[1]);
  const double v = x[i] * pow(fmax(1e-10, ratio[i] / mid), 0.3);
  y[i] = fmax(0.001, fmax(x[i] - move, fmin(1.0, fmin(x[i] + move, v))));
}

__global__ void cone_bisect_kernel(const double* vpart, int nblk, int64_t N, double* b, double target) {
  if (b[2] == 0.0) return;
  const double s = wave_reduce_partials(vpart, nblk, 1);
  if (threadIdx.x == 0) {
    const double volume = s / double(N);
    const double mid = 0.5 * (b[0] + b[1]);
    if (volume > target) b[0] = mid;
    else b[1] = mid;
    b[3] += 1.0;
    b[2] = ((b[1] - b[0]) / (b[1] + b[0]) > 1e-4 && b[1] > 1e-40) ? 1.0 : 0.0;
  }
}

// max |a - b| partials and finish (order independent)
__global__ __launch_bounds__(256) void absmax_partials(const double* a, const double* b, int64_t N, double* part) {
  __shared__ double red[8];
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  double v = i < N ? fabs(a[i] - b[i]) : 0.0;
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) red[wave] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = 0.0;
    for (int k = 0; k < (int)(blockDim.x >> 6); ++k) m = fmax(m, red[k]);
    part[blockIdx.x] = m;
  }
}

__global__ void absmax_finish(const double* part, int nblk, double* out) {
  double v = 0.0;
  for (int k = threadIdx.x; k < nblk; k += 64) v = fmax(v, part[k]);
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
  if (threadIdx.x == 0) out[0] = v;
}

// ------------------------------------------------------------------ gaussian
// Periodic correlation along one axis in scipy's symmetric order:
// y = w[r] x[i] + sum_{k=r..1} (x[i-k] + x[i+k]) w[r+k].
__global__ __launch_bounds__(256) void gauss_axis_kernel(const double* x, const double* w, int radius, int n, int axis,
                                                         double* y) {
  const int64_t N = int64_t(n) * n * n;
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const int64_t stride = axis == 0 ? int64_t(n) * n : (axis == 1 ? n : 1);
  const int coord = int((i / stride) % n);
  double v = x[i] * w[radius];
  for (int k = radius; k >= 1; --k) {
    int lo = (coord - k) % n;
    if (lo < 0) lo += n;
    const int hi = (coord + k) % n;
    v += (x[i + int64_t(lo - coord) * stride] + x[i + int64_t(hi - coord) * stride]) * w[radius + k];
  }
  y[i] = v;
}

__global__ __launch_bounds__(256) void gauss_project_kernel(const double* t, double beta, double* ph, double* dp,
                                                            double* a, int64_t N, double emin, double* vpart) {
  __shared__ double red[8];
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  double p = 0.0;
  if (i < N) {
    const double den = 2 * tanh(beta / 2);
    const double q = tanh(beta * (t[i] - 0.5));
    p = (tanh(beta / 2) + q) / den;
    ph[i] = p;
    dp[i] = beta * (1 - q * q) / den;
    a[i] = emin + (1 - emin) * (p * p * p);
  }
  if (vpart) {
    const double s = block_sum(p, red);
    if (threadIdx.x == 0) vpart[blockIdx.x] = s;
  }
}

// g_in = dp p (1-Emin) ph^2 e  (left-to-right as the reference)
__global__ __launch_bounds__(256) void gauss_prepare_kernel(const double* dp, const double* ph, const double* e,
                                                            double emin, double penal, int64_t N, double* gin) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= N) return;
  gin[i] = dp[i] * penal * (1 - emin) * (ph[i] * ph[i]) * e[i];
}

// ratio = max(g,1e-30)/max(dv,1e-30), block maxima of ratio
__global__ __launch_bounds__(256) void gauss_ratio_kernel(const double* g, const double* dv, int64_t N, double* ratio,
                                                          double* part) {
  __shared__ double red[8];
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  double v = 0.0;
  if (i < N) {
    v = fmax(g[i], 1e-30) / fmax(dv[i], 1e-30);
    ratio[i] = v;
  }
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) red[wave] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = 0.0;
    for (int k = 0; k < (int)(blockDim.x >> 6); ++k) m = fmax(m, red[k]);
    part[blockIdx.x] = m;
  }
}

__global__ void gauss_bounds_kernel(const double* part, int nblk, double* bounds) {
  double v = 0.0;
  for (int k = threadIdx.x; k < nblk; k += 64) v = fmax(v, part[k]);
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
  if (threadIdx.x == 0) {
    bounds[0] = 0.0;
    bounds[1] = v * 1e4;
  }
}

__global__ __launch_bounds__(256) void gauss_candidate_kernel(const double* x, const double* ratio, const double* bounds,
                                                              double move, double* y, int64_t N) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const double lam = fmax((bounds[0] + bounds[1]) * 0.5, 1e-30);
  const double lo = fmax(0.0, x[i] - move), hi = fmin(1.0, x[i] + move);
  y[i] = fmin(hi, fmax(lo, x[i] * sqrt(ratio[i] / lam)));
}

__global__ void gauss_bisect_kernel(const double* vpart, int nblk, int64_t N, double* b, double target) {
  const double s = wave_reduce_partials(vpart, nblk, 1);
  if (threadIdx.x == 0) {
    const double mid = (b[0] + b[1]) * 0.5;
    if (s / double(N) > target) b[0] = mid;
    else b[1] = mid;
  }
}

}  // namespace pgj

using namespace pgj;

namespace {
int blocks_of(int64_t N) { return int((N + 255) / 256); }
}

extern "C" size_t pgj_oc_ws_bytes(int n) {
  const int64_t N = int64_t(n) * n * n;
  return align256(size_t(blocks_of(N)) * 8);
}

extern "C" int pgj_cone_filter(int n, const double* x, const int32_t* offsets, const double* weights, int m,
                               double* y, const double* bounds, int conditional, void* stream) {
  PGJ_REQUIRE(n >= 2 && x && offsets && weights && y && m >= 1, "invalid cone filter arguments");
  PGJ_REQUIRE(!conditional || bounds, "conditional filtering needs bounds");
  const int64_t N = int64_t(n) * n * n;
  hipLaunchKernelGGL(cone_filter_kernel, dim3(blocks_of(N)), dim3(256), 0, as_stream(stream), x, offsets, weights, m,
                     n, y, bounds, conditional);
  PGJ_LAUNCHED("cone_filter");
  return PGJ_OK;
}

extern "C" int pgj_cone_project(int n, const double* t, const double* hs, double* ph, double* dp, double* a,
                                const double* bounds, double beta, double eta, double emin, int conditional,
                                double* ws, void* stream) {
  PGJ_REQUIRE(n >= 2 && t && hs && ph && dp && a, "invalid projection arguments");
  PGJ_REQUIRE(!conditional || bounds, "conditional projection needs bounds");
  const int64_t N = int64_t(n) * n * n;
  hipLaunchKernelGGL(cone_project_kernel, dim3(blocks_of(N)), dim3(256), 0, as_stream(stream), t, hs, ph, dp, a, bounds,
                     N, beta, eta, emin, conditional, ws);
  PGJ_LAUNCHED("cone_project");
  return PGJ_OK;
}

extern "C" int pgj_cone_sensitivity(int n, const double* dp, const double* ph, const double* energy, const double* hs,
                                    double emin, double* gin, double* dvin, void* stream) {
  PGJ_REQUIRE(n >= 2 && dp && ph && energy && hs && gin && dvin, "invalid sensitivity arguments");
  const int64_t N = int64_t(n) * n * n;
  hipLaunchKernelGGL(cone_prepare_kernel, dim3(blocks_of(N)), dim3(256), 0, as_stream(stream), dp, ph, energy, hs, emin,
                     N, gin, dvin);
  PGJ_LAUNCHED("cone_prepare");
  return PGJ_OK;
}

extern "C" int pgj_cone_ratio(int n, double* g, const double* dv, double* ratio, double* bounds, void* stream) {
  PGJ_REQUIRE(n >= 2 && g && dv && ratio && bounds, "invalid ratio arguments");
  const int64_t N = int64_t(n) * n * n;
  hipLaunchKernelGGL(cone_ratio_kernel, dim3(blocks_of(N)), dim3(256), 0, as_stream(stream), g, dv, N, ratio, bounds);
  PGJ_LAUNCHED("cone_ratio");
  return PGJ_OK;
}

extern "C" int pgj_cone_candidate(int n, const double* x, const double* ratio, const double* bounds, double move,
                                  double* y, void* stream) {
  PGJ_REQUIRE(n >= 2 && x && ratio && bounds && y, "invalid candidate arguments");
  const int64_t N = int64_t(n) * n * n;
  hipLaunchKernelGGL(cone_candidate_kernel, dim3(blocks_of(N)), dim3(256), 0, as_stream(stream), x, ratio, bounds, move,
                     y, N);
  PGJ_LAUNCHED("cone_candidate");
  return PGJ_OK;
}

extern "C" int pgj_cone_bisect(int n, const double* ws, double* bounds, double volume, void* stream) {
  PGJ_REQUIRE(n >= 2 && ws && bounds, "invalid bisection arguments");
  const int64_t N = int64_t(n) * n * n;
  hipLaunchKernelGGL(cone_bisect_kernel, dim3(1), dim3(64), 0, as_stream(stream), ws, blocks_of(N), N, bounds, volume);
  PGJ_LAUNCHED("cone_bisect");
  return PGJ_OK;
}

extern "C" int pgj_absmax_diff(int n, const double* a, const double* b, double* ws, double* out, void* stream) {
  PGJ_REQUIRE(n >= 2 && a && b && ws && out, "invalid absmax arguments");
  const int64_t N = int64_t(n) * n * n;
  hipStream_t s = as_stream(stream);
  hipLaunchKernelGGL(absmax_partials, dim3(blocks_of(N)), dim3(256), 0, s, a, b, N, ws);
  PGJ_LAUNCHED("absmax_partials");
  hipLaunchKernelGGL(absmax_finish, dim3(1), dim3(64), 0, s, ws, blocks_of(N), out);
  PGJ_LAUNCHED("absmax_finish");
  return PGJ_OK;
}

extern "C" int pgj_gauss_filter(int n, const double* x, const double* weights, int radius, double* tmp1, double* tmp2,
                                double* y, void* stream) {
  PGJ_REQUIRE(n >= 2 && x && weights && tmp1 && tmp2 && y && radius >= 0 && radius <= 4 * n,
              "invalid gaussian filter arguments");
  const int64_t N = int64_t(n) * n * n;
  hipStream_t s = as_stream(stream);
  hipLaunchKernelGGL(gauss_axis_kernel, dim3(blocks_of(N)), dim3(256), 0, s, x, weights, radius, n, 0, tmp1);
  hipLaunchKernelGGL(gauss_axis_kernel, dim3(blocks_of(N)), dim3(256), 0, s, tmp1, weights, radius, n, 1, tmp2);
  hipLaunchKernelGGL(gauss_axis_kernel, dim3(blocks_of(N)), dim3(256), 0, s, tmp2, weights, radius, n, 2, y);
  PGJ_LAUNCHED("gauss_axis");
  return PGJ_OK;
}

extern "C" int pgj_gauss_project(int n, const double* t, double beta, double* ph, double* dp, double* a, double emin,
                                 double* ws, void* stream) {
  PGJ_REQUIRE(n >= 2 && t && ph && dp && a, "invalid projection arguments");
  const int64_t N = int64_t(n) * n * n;
  hipLaunchKernelGGL(gauss_project_kernel, dim3(blocks_of(N)), dim3(256), 0, as_stream(stream), t, beta, ph, dp, a, N,
                     emin, ws);
  PGJ_LAUNCHED("gauss_project");
  return PGJ_OK;
}

extern "C" int pgj_gauss_sensitivity(int n, const double* dp, const double* ph, const double* energy, double emin,
                                     double penal, double* gin, void* stream) {
  PGJ_REQUIRE(n >= 2 && dp && ph && energy && gin, "invalid sensitivity arguments");
  const int64_t N = int64_t(n) * n * n;
  hipLaunchKernelGGL(gauss_prepare_kernel, dim3(blocks_of(N)), dim3(256), 0, as_stream(stream), dp, ph, energy, emin,
                     penal, N, gin);
  PGJ_LAUNCHED("gauss_prepare");
  return PGJ_OK;
}

extern "C" int pgj_gauss_ratio(int n, const double* g, const double* dv, double* ratio, double* ws, double* bounds,
                               void* stream) {
  PGJ_REQUIRE(n >= 2 && g && dv && ratio && ws && bounds, "invalid ratio arguments");
  const int64_t N = int64_t(n) * n * n;
  hipStream_t s = as_stream(stream);
  hipLaunchKernelGGL(gauss_ratio_kernel, dim3(blocks_of(N)), dim3(256), 0, s, g, dv, N, ratio, ws);
  PGJ_LAUNCHED("gauss_ratio");
  hipLaunchKernelGGL(gauss_bounds_kernel, dim3(1), dim3(64), 0, s, ws, blocks_of(N), bounds);
  PGJ_LAUNCHED("gauss_bounds");
  return PGJ_OK;
}

extern "C" int pgj_gauss_candidate(int n, const double* x, const double* ratio, const double* bounds, double move,
                                   double* y, void* stream) {
  PGJ_REQUIRE(n >= 2 && x && ratio && bounds && y, "invalid candidate arguments");
  const int64_t N = int64_t(n) * n * n;
  hipLaunchKernelGGL(gauss_candidate_kernel, dim3(blocks_of(N)), dim3(256), 0, as_stream(stream), x, ratio, bounds,
                     move, y, N);
  PGJ_LAUNCHED("gauss_candidate");
  return PGJ_OK;
}

extern "C" int pgj_gauss_bisect(int n, const double* ws, double* bounds, double volume, void* stream) {
  PGJ_REQUIRE(n >= 2 && ws && bounds, "invalid bisection arguments");
  const int64_t N = int64_t(n) * n * n;
  hipLaunchKernelGGL(gauss_bisect_kernel, dim3(1), dim3(64), 0, as_stream(stream), ws, blocks_of(N), N, bounds, volume);
  PGJ_LAUNCHED("gauss_bisect");
  return PGJ_OK;
}
