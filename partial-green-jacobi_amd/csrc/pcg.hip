// Vector kernels of the generic batched PCG (solve_pcg_batched) on (B, n)
// row-major FP64 vectors: dot-product partials reduced in a fixed order, the
// fused alpha/u/r update with |r|^2 partials, the per-batch relative residual /
// best-iterate / convergence-latch finalization, the fused (optionally Jacobi)
// z / r.z / beta / p update. Semantics follow
// solver/src/runtime/fi/source/femlib/batched/pcg.py:75-245 (eager loop),
// :636-669 (compiled pre/post steps), :677-727 (latch kernels, best copy).
#include <cmath>

#include "layout.h"
#include "pgj_internal.h"

namespace pgj {

constexpr int kVecThreads = 256;
constexpr double kPcgEps = 1e-30;  // pcg.py:112 (eps clamp of pAp and rz)

inline int vec_blocks(int64_t n) {
  const int64_t b = (n + kVecThreads * 4 - 1) / (kVecThreads * 4);
  return int(b < 1024 ? b : 1024);
}

#define PGJ_GRID_LOOP(i, n) \
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < (n); i += int64_t(gridDim.x) * blockDim.x)

// r = rhs - Ax (Ax may be null: r = rhs); rhs_part[b][blk] = partial |rhs_b|^2.
__global__ __launch_bounds__(kVecThreads) void pcg_start(const double* rhs, const double* Ax, double* r, int64_t n,
                                                         double* rhs_part) {
  __shared__ double red[kVecThreads / 64];
  const int64_t off = int64_t(blockIdx.y) * n;
  double s = 0.0;
  PGJ_GRID_LOOP(i, n) {
    const double f = rhs[off + i];
    r[off + i] = Ax ? f - Ax[off + i] : f;
    s = fma(f, f, s);
  }
  const double t = block_sum(s, red);
  if (threadIdx.x == 0) rhs_part[int64_t(blockIdx.y) * gridDim.x + blockIdx.x] = t;
}

// rhs_norm[b] = max(sqrt(sum of partials), 1e-30), one wavefront per batch.
__global__ void pcg_norms(int B, const double* part, int nblk, double* out) {
  const int wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
  for (int b = blockIdx.x * nw + wave; b < B; b += gridDim.x * nw) {
    const double s = wave_reduce_partials(part + int64_t(b) * nblk, nblk, 1);
    if ((threadIdx.x & 63) == 0) out[b] = fmax(sqrt(s), 1e-30);
  }
}

// part[b][blk] = partial x_b . y_b; with minv, z = minv * x is stored and y is ignored (Jacobi r.z).
__global__ __launch_bounds__(kVecThreads) void pcg_dot(const double* x, const double* y, const double* minv, double* z,
                                                       int64_t n, int64_t minv_stride, double* part) {
  __shared__ double red[kVecThreads / 64];
  const int64_t off = int64_t(blockIdx.y) * n;
  const int64_t moff = int64_t(blockIdx.y) * minv_stride;
  double s = 0.0;
  PGJ_GRID_LOOP(i, n) {
    const double xv = x[off + i];
    double yv;
    if (minv) {
      yv = xv * minv[moff + i];
      z[off + i] = yv;
    } else {
      yv = y[off + i];
    }
    s = fma(xv, yv, s);
  }
  const double t = block_sum(s, red);
  if (threadIdx.x == 0) part[int64_t(blockIdx.y) * gridDim.x + blockIdx.x] = t;
}

// Per batch: reduce partials to a scalar (one wavefront each).
__global__ void pcg_reduce(int B, const double* part, int nblk, double* out) {
  const int wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
  for (int b = blockIdx.x * nw + wave; b < B; b += gridDim.x * nw) {
    const double s = wave_reduce_partials(part + int64_t(b) * nblk, nblk, 1);
    if ((threadIdx.x & 63) == 0) out[b] = s;
  }
}

// alpha = rz / max(pAp, eps), 0 where pAp is not finite or <= 0; u += alpha p;
// r -= alpha Ap; rr_part = partial |r|^2 of the updated residual.
__global__ __launch_bounds__(kVecThreads) void pcg_update_ur(double* u, double* r, const double* p, const double* Ap,
                                                             int64_t n, const double* pap_part, int nblk,
                                                             const double* rz, double* rr_part) {
  __shared__ double red[kVecThreads / 64];
  __shared__ double s_alpha;
  const int b = blockIdx.y;
  if (threadIdx.x < 64) {
    const double pap = wave_reduce_partials(pap_part + int64_t(b) * nblk, nblk, 1);
    if (threadIdx.x == 0) {
      const bool bad = !isfinite(pap) || pap <= 0.0;
      s_alpha = bad ? 0.0 : rz[b] / fmax(pap, kPcgEps);
    }
  }
  __syncthreads();
  const double alpha = s_alpha;
  const int64_t off = int64_t(b) * n;
  double s = 0.0;
  PGJ_GRID_LOOP(i, n) {
    u[off + i] = fma(alpha, p[off + i], u[off + i]);
    const double rn = fma(-alpha, Ap[off + i], r[off + i]);
    r[off + i] = rn;
    s = fma(rn, rn, s);
  }
  const double t = block_sum(s, red);
  if (threadIdx.x == 0) rr_part[int64_t(b) * gridDim.x + blockIdx.x] = t;
}

// res = sqrt(|r|^2) / rhs_norm; improved = res < best_res (latched: and no stop
// recorded yet); best_res updated. Latched: state[0] (stop iteration) is set to
// `iteration` the first time every res < tol; state[1] to 1 the first time
// every res < near_tol (near_tol > 0 only). One block.
__global__ void pcg_finalize(int B, const double* rr_part, int nblk, const double* rhs_norm, double* res,
                             double* best_res, uint8_t* improved, int64_t* state, double tol, double near_tol,
                             int64_t iteration, int latched) {
  const int wave = threadIdx.x >> 6, nw = blockDim.x >> 6, lane = threadIdx.x & 63;
  const int64_t stopped = latched ? state[0] : 0;
  int ok = 1, near = 1;
  for (int b = wave; b < B; b += nw) {
    const double rr = wave_reduce_partials(rr_part + int64_t(b) * nblk, nblk, 1);
    const double rv = sqrt(rr) / rhs_norm[b];
    const bool imp = rv < best_res[b] && stopped == 0;
    ok &= rv < tol;
    near &= rv < near_tol;
    if (lane == 0) {
      res[b] = rv;
      improved[b] = imp;
      if (imp) best_res[b] = rv;
    }
  }
  const int all_ok = __syncthreads_and(ok);
  const int all_near = __syncthreads_and(near);
  if (threadIdx.x == 0 && latched) {
    if (stopped == 0 && all_ok) state[0] = iteration;
    if (near_tol > 0.0 && state[1] == 0 && all_near) state[1] = 1;
  }
}

__global__ __launch_bounds__(kVecThreads) void pcg_copy_best(const double* u, double* best_u, const uint8_t* improved,
                                                             int64_t n) {
  const int b = blockIdx.y;
  if (!improved[b]) return;
  const int64_t off = int64_t(b) * n;
  PGJ_GRID_LOOP(i, n) best_u[off + i] = u[off + i];
}

// beta = rz_new / max(rz_old, eps); p = z + beta p (rz_old read here, replaced by pcg_reduce afterwards).
__global__ __launch_bounds__(kVecThreads) void pcg_update_p(const double* z, double* p, int64_t n,
                                                            const double* rz_part, int nblk, const double* rz) {
  __shared__ double s_beta;
  const int b = blockIdx.y;
  if (threadIdx.x < 64) {
    const double rzn = wave_reduce_partials(rz_part + int64_t(b) * nblk, nblk, 1);
    if (threadIdx.x == 0) s_beta = rzn / fmax(rz[b], kPcgEps);
  }
  __syncthreads();
  const double beta = s_beta;
  const int64_t off = int64_t(b) * n;
  PGJ_GRID_LOOP(i, n) p[off + i] = fma(beta, p[off + i], z[off + i]);
}

__global__ __launch_bounds__(kVecThreads) void pcg_copy(const double* x, double* y, int64_t total) {
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < total; i += int64_t(gridDim.x) * blockDim.x)
    y[i] = x[i];
}

inline dim3 batch_grid(int B) { return dim3((B + 3) / 4); }

}  // namespace pgj

using namespace pgj;

namespace {
bool pcg_shape_ok(int B, int64_t n) { return B >= 1 && B <= 65535 && n >= 1 && n <= (int64_t(1) << 40); }
}  // namespace

extern "C" size_t pgj_pcg_ws_bytes(int B, int64_t n) {
  if (!pcg_shape_ok(B, n)) return 0;
  return align256(size_t(B) * vec_blocks(n) * 8 * 3);
}

extern "C" int pgj_pcg_start(int B, int64_t n, const double* rhs, const double* Ax, double* r, double* rhs_norm,
                             double* ws, void* stream) {
  PGJ_REQUIRE(pcg_shape_ok(B, n) && rhs && r && rhs_norm && ws, "invalid pcg start arguments");
  hipStream_t s = as_stream(stream);
  const int nblk = vec_blocks(n);
  hipLaunchKernelGGL(pcg_start, dim3(nblk, B), dim3(kVecThreads), 0, s, rhs, Ax, r, n, ws);
  PGJ_LAUNCHED("pcg_start");
  hipLaunchKernelGGL(pcg_norms, batch_grid(B), dim3(256), 0, s, B, ws, nblk, rhs_norm);
  PGJ_LAUNCHED("pcg_norms");
  return PGJ_OK;
}

extern "C" int pgj_pcg_init(int B, int64_t n, const double* r, const double* z, double* p, double* rz, double* ws,
                            void* stream) {
  PGJ_REQUIRE(pcg_shape_ok(B, n) && r && z && p && rz && ws, "invalid pcg init arguments");
  hipStream_t s = as_stream(stream);
  const int nblk = vec_blocks(n);
  hipLaunchKernelGGL(pcg_dot, dim3(nblk, B), dim3(kVecThreads), 0, s, r, z, nullptr, nullptr, n, int64_t(0), ws);
  PGJ_LAUNCHED("pcg_dot");
  hipLaunchKernelGGL(pcg_reduce, batch_grid(B), dim3(256), 0, s, B, ws, nblk, rz);
  PGJ_LAUNCHED("pcg_reduce");
  const int64_t total = int64_t(B) * n;
  const int64_t g = (total + kVecThreads * 4 - 1) / (kVecThreads * 4);
  hipLaunchKernelGGL(pcg_copy, dim3(unsigned(g < 4096 ? g : 4096)), dim3(kVecThreads), 0, s, z, p, total);
  PGJ_LAUNCHED("pcg_copy");
  return PGJ_OK;
}

extern "C" int pgj_pcg_pre_step(int B, int64_t n, double* u, double* r, const double* p, const double* Ap,
                                const double* rz, const double* rhs_norm, double* best_u, double* best_res,
                                double* res, uint8_t* improved, int64_t* state, double tol, double near_tol,
                                int64_t iteration, int latched, double* ws, void* stream) {
  PGJ_REQUIRE(pcg_shape_ok(B, n) && u && r && p && Ap && rz && rhs_norm && best_u && best_res && res && improved &&
                  state && ws,
              "invalid pcg pre-step arguments");
  hipStream_t s = as_stream(stream);
  const int nblk = vec_blocks(n);
  double* pap = ws;
  double* rr = ws + size_t(B) * nblk;
  hipLaunchKernelGGL(pcg_dot, dim3(nblk, B), dim3(kVecThreads), 0, s, p, Ap, nullptr, nullptr, n, int64_t(0), pap);
  PGJ_LAUNCHED("pcg_dot");
  hipLaunchKernelGGL(pcg_update_ur, dim3(nblk, B), dim3(kVecThreads), 0, s, u, r, p, Ap, n, pap, nblk, rz, rr);
  PGJ_LAUNCHED("pcg_update_ur");
  hipLaunchKernelGGL(pcg_finalize, dim3(1), dim3(1024), 0, s, B, rr, nblk, rhs_norm, res, best_res, improved, state,
                     tol, near_tol, iteration, latched);
  PGJ_LAUNCHED("pcg_finalize");
  hipLaunchKernelGGL(pcg_copy_best, dim3(nblk, B), dim3(kVecThreads), 0, s, u, best_u, improved, n);
  PGJ_LAUNCHED("pcg_copy_best");
  return PGJ_OK;
}

extern "C" int pgj_pcg_post_step(int B, int64_t n, const double* r, double* z, const double* minv,
                                 int64_t minv_stride, double* p, double* rz, double* ws, void* stream) {
  PGJ_REQUIRE(pcg_shape_ok(B, n) && r && z && p && rz && ws && (minv_stride == 0 || minv_stride == n),
              "invalid pcg post-step arguments");
  hipStream_t s = as_stream(stream);
  const int nblk = vec_blocks(n);
  double* part = ws + size_t(2) * B * nblk;
  hipLaunchKernelGGL(pcg_dot, dim3(nblk, B), dim3(kVecThreads), 0, s, r, z, minv, z, n, minv_stride, part);
  PGJ_LAUNCHED("pcg_dot");
  hipLaunchKernelGGL(pcg_update_p, dim3(nblk, B), dim3(kVecThreads), 0, s, z, p, n, part, nblk, rz);
  PGJ_LAUNCHED("pcg_update_p");
  hipLaunchKernelGGL(pcg_reduce, batch_grid(B), dim3(256), 0, s, B, part, nblk, rz);
  PGJ_LAUNCHED("pcg_reduce");
  return PGJ_OK;
}
