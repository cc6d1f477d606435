// On-device choice of the Green-Jacobi exponent alpha of a design, from
// reductions over the SIMP field a (and the physical density for the gray
// selector), followed by the scale s = exp(-alpha q / 2) with alpha read from
// device memory, so selection and scaling need no host round trip.
//
// Selectors (reference: exp/asg_density_selector_20261105/run.py:28-67,
// exp/asg_gpu_resident_20261103/resident.py:156-174,
// solver/experiments/global_alpha_compact_law/law.py:9-18):
//   fixed / green / gj  alpha = value / 0 / 1
//   eta                 alpha = anchor of the Emin decade (host-side)
//   contrast            s = clamp(log10(max a / min a), 0, 9); piecewise-linear
//                       in s through (4, b4), (6, b6), (9, b9), clamp [0, 1],
//                       rounded to 1/16; 0 when s <= 1e-12
//   gray                contrast base + gamma (g_ref - g), g = clamp((v - <rho^2>)
//                       / (v (1 - v) + 1e-12), 0, 1), v = <rho>; same rounding
//   compact8(q)         law features of a in FP32: ell = log a, q = (ell - lo) /
//                       max(hi - lo, 1e-12); per axis d = |q(x + e) - q(x)|
//                       (periodic): area = mean over axes of <d > 1e-6>, jump =
//                       sum <d^2> / max(sum <d>, 1e-20), contrast = (hi - lo) /
//                       ln 1e4, complexity = area n; alpha = 1 - 0.75
//                       sqrt(max(area jump, 0)) / sqrt(max(contrast, 1e-30)) -
//                       (2.5 / max(complexity, 1e-30))^8, clamped to [0, 1]
//                       (0 unless complexity, contrast > 1e-12), 0 below 0.375;
//                       compact8q rounds to 1/16.
// The FP32 law features accumulate in FP64 in a fixed order (the reference's
// torch FP32 means sum in a library-chosen order), so they agree to FP32
// rounding, not bitwise.
#include <cmath>

#include "hex8_kernels.h"
#include "layout.h"
#include "pgj_internal.h"

namespace pgj {

constexpr int kSelThreads = 256;
// per-block partials of the statistics pass: min a, max a, min ell, max ell, sum rho, sum rho^2
constexpr int kStatVals = 6;
// per-block partials of the law pass: per axis sum d, sum d^2, count(d > 1e-6)
constexpr int kLawVals = 9;

__device__ inline double block_min(double v, double* scratch) {
  for (int o = 32; o >= 1; o >>= 1) v = fmin(v, __shfl_xor(v, o, 64));
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
  __syncthreads();
  if (lane == 0) scratch[wave] = v;
  __syncthreads();
  double t = scratch[0];
  for (int w = 1; w < nw; ++w) t = fmin(t, scratch[w]);
  __syncthreads();
  return t;
}

__device__ inline double block_max(double v, double* scratch) { return -block_min(-v, scratch); }

__global__ __launch_bounds__(kSelThreads) void alpha_stats_kernel(const double* a, const double* rho, int64_t N,
                                                                  double* part) {
  __shared__ double red[kSelThreads / 64];
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  double amin = INFINITY, amax = -INFINITY, lmin = INFINITY, lmax = -INFINITY, r1 = 0.0, r2 = 0.0;
  if (i < N) {
    const double v = a[i];
    const float ell = logf(float(v));
    amin = amax = v;
    lmin = lmax = double(ell);
    if (rho) {
      const double r = rho[i];
      r1 = r;
      r2 = r * r;
    }
  }
  double* p = part + int64_t(blockIdx.x) * kStatVals;
  const double t0 = block_min(amin, red), t1 = block_max(amax, red);
  const double t2 = block_min(lmin, red), t3 = block_max(lmax, red);
  const double t4 = block_sum(r1, red), t5 = block_sum(r2, red);
  if (threadIdx.x == 0) {
    p[0] = t0;
    p[1] = t1;
    p[2] = t2;
    p[3] = t3;
    p[4] = t4;
    p[5] = t5;
  }
}

// min / max of column k of the statistics partials (every lane obtains the same value)
__device__ inline double wave_partial_min(const double* part, int nblk, int k) {
  const int lane = threadIdx.x & 63;
  double v = INFINITY;
  for (int b = lane; b < nblk; b += 64) v = fmin(v, part[int64_t(b) * kStatVals + k]);
  for (int o = 32; o >= 1; o >>= 1) v = fmin(v, __shfl_xor(v, o, 64));
  return v;
}

__device__ inline double wave_partial_max(const double* part, int nblk, int k) {
  const int lane = threadIdx.x & 63;
  double v = -INFINITY;
  for (int b = lane; b < nblk; b += 64) v = fmax(v, part[int64_t(b) * kStatVals + k]);
  for (int o = 32; o >= 1; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  return v;
}

__global__ __launch_bounds__(kSelThreads) void alpha_law_kernel(const double* a, int nz, int ny, int nx,
                                                                const double* stat, int nblk, double* part) {
  __shared__ double red[kSelThreads / 64];
  __shared__ float lo_s, den_s;
  if (threadIdx.x < 64) {
    const float lo = float(wave_partial_min(stat, nblk, 2)), hi = float(wave_partial_max(stat, nblk, 3));
    const float span = hi - lo;
    if (threadIdx.x == 0) {
      lo_s = lo;
      den_s = span > 1e-12f ? span : 1e-12f;
    }
  }
  __syncthreads();
  const float lo = lo_s, den = den_s;
  const int64_t N = int64_t(nx) * ny * nz;
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  double v[kLawVals];
  for (int k = 0; k < kLawVals; ++k) v[k] = 0.0;
  if (i < N) {
    const int x = int(i % nx), y = int((i / nx) % ny), z = int(i / (int64_t(nx) * ny));
    auto q = [&](int zz, int yy, int xx) {
      return (logf(float(a[(int64_t(zz) * ny + yy) * nx + xx])) - lo) / den;
    };
    const float q0 = q(z, y, x);
    // axis order of the roll: 0 = z, 1 = y, 2 = x (the field is (z, y, x))
    const float nb[3] = {q(wrapi(z + 1, nz), y, x), q(z, wrapi(y + 1, ny), x), q(z, y, wrapi(x + 1, nx))};
    for (int ax = 0; ax < 3; ++ax) {
      const float d = fabsf(nb[ax] - q0);
      v[3 * ax] = double(d);
      v[3 * ax + 1] = double(d * d);
      v[3 * ax + 2] = d > 1e-6f ? 1.0 : 0.0;
    }
  }
  double* p = part + int64_t(blockIdx.x) * kLawVals;
  for (int k = 0; k < kLawVals; ++k) {
    const double t = block_sum(v[k], red);
    if (threadIdx.x == 0) p[k] = t;
  }
}

__device__ inline double round16(double v) { return floor(v * 16.0 + 0.5) / 16.0; }

__device__ inline double contrast_base(double s, const double* anchors) {
  const double b4 = anchors[0], b6 = anchors[1], b9 = anchors[2];
  return s < 4.0 ? b4 * s / 4.0 : (s < 6.0 ? b4 + (b6 - b4) * (s - 4.0) / 2.0 : b6 + (b9 - b6) * (s - 6.0) / 3.0);
}

struct AlphaParams {
  double v[6];
};

// One wave: features out[0..6] = (s, <rho>, <rho^2>, area, jump, law contrast, complexity); alpha.
__global__ void alpha_finish_kernel(int method, AlphaParams prm, const double* stat, int nblk, const double* law,
                                    int nlaw, int64_t N, int n_axis0, double* out, double* alpha) {
  const double lo = wave_partial_min(stat, nblk, 0), hi = wave_partial_max(stat, nblk, 1);
  const double llo = wave_partial_min(stat, nblk, 2), lhi = wave_partial_max(stat, nblk, 3);
  double r1 = 0.0, r2 = 0.0;
  {
    const int lane = threadIdx.x & 63;
    for (int b = lane; b < nblk; b += 64) {
      r1 += stat[int64_t(b) * kStatVals + 4];
      r2 += stat[int64_t(b) * kStatVals + 5];
    }
    r1 = wave_sum(r1);
    r2 = wave_sum(r2);
  }
  double sums[kLawVals];
  for (int k = 0; k < kLawVals; ++k) sums[k] = 0.0;
  if (law) {
    const int off[kLawVals] = {0, 1, 2, 3, 4, 5, 6, 7, 8};
    wave_reduce_partials_n<kLawVals>(law, off, nlaw, kLawVals, sums);
  }
  const double s = fmin(fmax(log10(hi / lo), 0.0), 9.0);
  const double mr = r1 / double(N), mr2 = r2 / double(N);
  // law features in the reference's float32 flow (law.py:10-19): per-axis float32 means (each the float
  // rounding of an FP64 sum), float32 totals over the axes in order z, y, x, float32 ratios; widened after
  auto f32 = [](double v) { return double(float(v)); };
  double total1 = 0.0, total2 = 0.0, act = 0.0;
  for (int ax = 0; ax < 3; ++ax) {
    total1 = f32(total1 + f32(sums[3 * ax] / double(N)));
    total2 = f32(total2 + f32(sums[3 * ax + 1] / double(N)));
    act = f32(act + f32(sums[3 * ax + 2] / double(N)));
  }
  const double area = f32(act / 3.0);
  const double jump = f32(total2 / fmax(total1, 1e-20));
  const float lspan = float(lhi) - float(llo);
  const double lcontrast = double(lspan / float(log(10000.0)));
  const double complexity = f32(area * double(n_axis0));
  double al = 0.0;
  switch (method) {
    case 0: al = prm.v[0]; break;                 // fixed (also eta: anchor chosen on the host)
    case 1: al = 0.0; break;                      // green
    case 2: al = 1.0; break;                      // gj
    case 4: {                                     // contrast
      const double value = contrast_base(s, prm.v);
      al = s > 1e-12 ? round16(fmin(fmax(value, 0.0), 1.0)) : 0.0;
      break;
    }
    case 5: {                                     // gray
      const double g = fmin(fmax((mr - mr2) / (mr * (1.0 - mr) + 1e-12), 0.0), 1.0);
      const double value = contrast_base(s, prm.v) + prm.v[3] * (prm.v[4] - g);
      al = s > 1e-12 ? round16(fmin(fmax(value, 0.0), 1.0)) : 0.0;
      break;
    }
    default: {                                    // compact8 (6), compact8q (7)
      double value = 1.0 - 0.75 * sqrt(fmax(area * jump, 0.0)) / sqrt(fmax(lcontrast, 1e-30)) -
                     pow(2.5 / fmax(complexity, 1e-30), 8.0);
      value = (complexity > 1e-12 && lcontrast > 1e-12) ? fmin(fmax(value, 0.0), 1.0) : 0.0;
      value = value < 0.375 ? 0.0 : value;
      al = method == 7 ? round16(value) : value;
    }
  }
  if (threadIdx.x == 0) {
    if (out) {
      out[0] = s;
      out[1] = mr;
      out[2] = mr2;
      out[3] = area;
      out[4] = jump;
      out[5] = lcontrast;
      out[6] = complexity;
    }
    *alpha = al;
  }
}

}  // namespace pgj

using namespace pgj;

namespace {
int64_t sel_blocks(const pgj_grid& g) {
  const int64_t N = int64_t(g.nx) * g.ny * g.nz;
  return (N + kSelThreads - 1) / kSelThreads;
}
}  // namespace

extern "C" size_t pgj_alpha_ws_bytes(const pgj_grid* g) {
  if (!g) return 0;
  return align256(size_t(sel_blocks(*g)) * (kStatVals + kLawVals) * 8);
}

extern "C" int pgj_alpha_select(const pgj_grid* g, int method, const double* params, const double* a, const double* rho,
                                double* ws, double* features, double* alpha, void* stream) {
  if (int e = check_grid(g)) return e;
  PGJ_REQUIRE(method >= 0 && method <= 7 && method != 3, "alpha method must be 0..2 or 4..7, got %d", method);
  PGJ_REQUIRE(a && ws && alpha, "invalid alpha selection arguments");
  PGJ_REQUIRE(method != 5 || rho, "the gray selector needs the physical density");
  PGJ_REQUIRE(method == 1 || method == 2 || params, "alpha method %d needs parameters", method);
  AlphaParams prm{};
  const int np = method == 0 ? 1 : (method == 4 ? 3 : (method == 5 ? 5 : 0));
  for (int k = 0; k < np; ++k) prm.v[k] = params[k];
  if (method == 0) PGJ_REQUIRE(prm.v[0] >= 0.0 && prm.v[0] <= 1.0, "alpha must be in [0, 1], got %g", prm.v[0]);
  const int64_t N = int64_t(g->nx) * g->ny * g->nz;
  const int nblk = int(sel_blocks(*g));
  double* stat = ws;
  double* law = ws + int64_t(nblk) * kStatVals;
  hipStream_t s = as_stream(stream);
  hipLaunchKernelGGL(alpha_stats_kernel, dim3(nblk), dim3(kSelThreads), 0, s, a, rho, N, stat);
  PGJ_LAUNCHED("alpha_stats_kernel");
  const bool need_law = method >= 6;
  if (need_law) {
    hipLaunchKernelGGL(alpha_law_kernel, dim3(nblk), dim3(kSelThreads), 0, s, a, g->nz, g->ny, g->nx, stat, nblk,
                       law);
    PGJ_LAUNCHED("alpha_law_kernel");
  }
  hipLaunchKernelGGL(alpha_finish_kernel, dim3(1), dim3(64), 0, s, method, prm, stat, nblk, need_law ? law : nullptr,
                     nblk, N, g->nz, features, alpha);
  PGJ_LAUNCHED("alpha_finish_kernel");
  return PGJ_OK;
}
