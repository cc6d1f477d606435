// Vector kernels of the generic batched PCG (solve_pcg_batched) on (B, n)
// row-major FP64 or FP32 vectors (sums and scalars in FP64): dot-product partials reduced in a fixed order, the
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
// Vectors are T (FP64 or FP32); sums, partials and per-batch scalars are FP64.
template <typename T>
__global__ __launch_bounds__(kVecThreads) void pcg_start(const T* rhs, const T* Ax, T* r, int64_t n,
                                                         double* rhs_part) {
  __shared__ double red[kVecThreads / 64];
  const int64_t off = int64_t(blockIdx.y) * n;
  double s = 0.0;
  PGJ_GRID_LOOP(i, n) {
    const T f = rhs[off + i];
    r[off + i] = Ax ? T(f - Ax[off + i]) : f;
    s = fma(double(f), double(f), s);
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
template <typename T>
__global__ __launch_bounds__(kVecThreads) void pcg_dot(const T* x, const T* y, const T* minv, T* z, int64_t n,
                                                       int64_t minv_stride, double* part) {
  __shared__ double red[kVecThreads / 64];
  const int64_t off = int64_t(blockIdx.y) * n;
  const int64_t moff = int64_t(blockIdx.y) * minv_stride;
  double s = 0.0;
  PGJ_GRID_LOOP(i, n) {
    const T xv = x[off + i];
    T yv;
    if (minv) {
      yv = xv * minv[moff + i];
      z[off + i] = yv;
    } else {
      yv = y[off + i];
    }
    s = fma(double(xv), double(yv), s);
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
template <typename T>
__global__ __launch_bounds__(kVecThreads) void pcg_update_ur(T* u, T* r, const T* p, const T* Ap, int64_t n,
                                                             const double* pap_part, int nblk, const double* rz,
                                                             double* rr_part) {
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
  const T alpha = T(s_alpha);
  const int64_t off = int64_t(b) * n;
  double s = 0.0;
  PGJ_GRID_LOOP(i, n) {
    u[off + i] = fma(alpha, p[off + i], u[off + i]);
    const T rn = fma(-alpha, Ap[off + i], r[off + i]);
    r[off + i] = rn;
    s = fma(double(rn), double(rn), s);
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

template <typename T>
__global__ __launch_bounds__(kVecThreads) void pcg_copy_best(const T* u, T* best_u, const uint8_t* improved,
                                                             int64_t n) {
  const int b = blockIdx.y;
  if (!improved[b]) return;
  const int64_t off = int64_t(b) * n;
  PGJ_GRID_LOOP(i, n) best_u[off + i] = u[off + i];
}

// beta = rz_new / max(rz_old, eps); p = z + beta p (rz_old read here, replaced by pcg_reduce afterwards).
template <typename T>
__global__ __launch_bounds__(kVecThreads) void pcg_update_p(const T* z, T* p, int64_t n, const double* rz_part,
                                                            int nblk, const double* rz) {
  __shared__ double s_beta;
  const int b = blockIdx.y;
  if (threadIdx.x < 64) {
    const double rzn = wave_reduce_partials(rz_part + int64_t(b) * nblk, nblk, 1);
    if (threadIdx.x == 0) s_beta = rzn / fmax(rz[b], kPcgEps);
  }
  __syncthreads();
  const T beta = T(s_beta);
  const int64_t off = int64_t(b) * n;
  PGJ_GRID_LOOP(i, n) p[off + i] = fma(beta, p[off + i], z[off + i]);
}

template <typename T>
__global__ __launch_bounds__(kVecThreads) void pcg_copy(const T* x, T* y, int64_t total) {
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

namespace {
template <typename T>
int pcg_start_t(int B, int64_t n, const T* rhs, const T* Ax, T* r, double* rhs_norm, double* ws, void* stream) {
  PGJ_REQUIRE(pcg_shape_ok(B, n) && rhs && r && rhs_norm && ws, "invalid pcg start arguments");
  hipStream_t s = as_stream(stream);
  const int nblk = vec_blocks(n);
  hipLaunchKernelGGL(pcg_start<T>, dim3(nblk, B), dim3(kVecThreads), 0, s, rhs, Ax, r, n, ws);
  PGJ_LAUNCHED("pcg_start");
  hipLaunchKernelGGL(pcg_norms, batch_grid(B), dim3(256), 0, s, B, ws, nblk, rhs_norm);
  PGJ_LAUNCHED("pcg_norms");
  return PGJ_OK;
}

template <typename T>
int pcg_init_t(int B, int64_t n, const T* r, const T* z, T* p, double* rz, double* ws, void* stream) {
  PGJ_REQUIRE(pcg_shape_ok(B, n) && r && z && p && rz && ws, "invalid pcg init arguments");
  hipStream_t s = as_stream(stream);
  const int nblk = vec_blocks(n);
  hipLaunchKernelGGL(pcg_dot<T>, dim3(nblk, B), dim3(kVecThreads), 0, s, r, z, nullptr, nullptr, n, int64_t(0), ws);
  PGJ_LAUNCHED("pcg_dot");
  hipLaunchKernelGGL(pcg_reduce, batch_grid(B), dim3(256), 0, s, B, ws, nblk, rz);
  PGJ_LAUNCHED("pcg_reduce");
  const int64_t total = int64_t(B) * n;
  const int64_t g = (total + kVecThreads * 4 - 1) / (kVecThreads * 4);
  hipLaunchKernelGGL(pcg_copy<T>, dim3(unsigned(g < 4096 ? g : 4096)), dim3(kVecThreads), 0, s, z, p, total);
  PGJ_LAUNCHED("pcg_copy");
  return PGJ_OK;
}

template <typename T>
int pcg_pre_step_t(int B, int64_t n, T* u, T* r, const T* p, const T* Ap, const double* rz, const double* rhs_norm,
                   T* best_u, double* best_res, double* res, uint8_t* improved, int64_t* state, double tol,
                   double near_tol, int64_t iteration, int latched, double* ws, void* stream) {
  PGJ_REQUIRE(pcg_shape_ok(B, n) && u && r && p && Ap && rz && rhs_norm && best_u && best_res && res && improved &&
                  state && ws,
              "invalid pcg pre-step arguments");
  hipStream_t s = as_stream(stream);
  const int nblk = vec_blocks(n);
  double* pap = ws;
  double* rr = ws + size_t(B) * nblk;
  hipLaunchKernelGGL(pcg_dot<T>, dim3(nblk, B), dim3(kVecThreads), 0, s, p, Ap, nullptr, nullptr, n, int64_t(0), pap);
  PGJ_LAUNCHED("pcg_dot");
  hipLaunchKernelGGL(pcg_update_ur<T>, dim3(nblk, B), dim3(kVecThreads), 0, s, u, r, p, Ap, n, pap, nblk, rz, rr);
  PGJ_LAUNCHED("pcg_update_ur");
  hipLaunchKernelGGL(pcg_finalize, dim3(1), dim3(1024), 0, s, B, rr, nblk, rhs_norm, res, best_res, improved, state,
                     tol, near_tol, iteration, latched);
  PGJ_LAUNCHED("pcg_finalize");
  hipLaunchKernelGGL(pcg_copy_best<T>, dim3(nblk, B), dim3(kVecThreads), 0, s, u, best_u, improved, n);
  PGJ_LAUNCHED("pcg_copy_best");
  return PGJ_OK;
}

template <typename T>
int pcg_post_step_t(int B, int64_t n, const T* r, T* z, const T* minv, int64_t minv_stride, T* p, double* rz,
                    double* ws, void* stream) {
  PGJ_REQUIRE(pcg_shape_ok(B, n) && r && z && p && rz && ws && (minv_stride == 0 || minv_stride == n),
              "invalid pcg post-step arguments");
  hipStream_t s = as_stream(stream);
  const int nblk = vec_blocks(n);
  double* part = ws + size_t(2) * B * nblk;
  hipLaunchKernelGGL(pcg_dot<T>, dim3(nblk, B), dim3(kVecThreads), 0, s, r, z, minv, z, n, minv_stride, part);
  PGJ_LAUNCHED("pcg_dot");
  hipLaunchKernelGGL(pcg_update_p<T>, dim3(nblk, B), dim3(kVecThreads), 0, s, z, p, n, part, nblk, rz);
  PGJ_LAUNCHED("pcg_update_p");
  hipLaunchKernelGGL(pcg_reduce, batch_grid(B), dim3(256), 0, s, B, part, nblk, rz);
  PGJ_LAUNCHED("pcg_reduce");
  return PGJ_OK;
}
}  // namespace

extern "C" int pgj_pcg_start(int B, int64_t n, const double* rhs, const double* Ax, double* r, double* rhs_norm,
                             double* ws, void* stream) {
  return pcg_start_t(B, n, rhs, Ax, r, rhs_norm, ws, stream);
}

extern "C" int pgj_pcg_init(int B, int64_t n, const double* r, const double* z, double* p, double* rz, double* ws,
                            void* stream) {
  return pcg_init_t(B, n, r, z, p, rz, ws, stream);
}

extern "C" int pgj_pcg_pre_step(int B, int64_t n, double* u, double* r, const double* p, const double* Ap,
                                const double* rz, const double* rhs_norm, double* best_u, double* best_res,
                                double* res, uint8_t* improved, int64_t* state, double tol, double near_tol,
                                int64_t iteration, int latched, double* ws, void* stream) {
  return pcg_pre_step_t(B, n, u, r, p, Ap, rz, rhs_norm, best_u, best_res, res, improved, state, tol, near_tol,
                        iteration, latched, ws, stream);
}

extern "C" int pgj_pcg_post_step(int B, int64_t n, const double* r, double* z, const double* minv,
                                 int64_t minv_stride, double* p, double* rz, double* ws, void* stream) {
  return pcg_post_step_t(B, n, r, z, minv, minv_stride, p, rz, ws, stream);
}

extern "C" int pgj_pcg32_start(int B, int64_t n, const float* rhs, const float* Ax, float* r, double* rhs_norm,
                               double* ws, void* stream) {
  return pcg_start_t(B, n, rhs, Ax, r, rhs_norm, ws, stream);
}

extern "C" int pgj_pcg32_init(int B, int64_t n, const float* r, const float* z, float* p, double* rz, double* ws,
                              void* stream) {
  return pcg_init_t(B, n, r, z, p, rz, ws, stream);
}

extern "C" int pgj_pcg32_pre_step(int B, int64_t n, float* u, float* r, const float* p, const float* Ap,
                                  const double* rz, const double* rhs_norm, float* best_u, double* best_res,
                                  double* res, uint8_t* improved, int64_t* state, double tol, double near_tol,
                                  int64_t iteration, int latched, double* ws, void* stream) {
  return pcg_pre_step_t(B, n, u, r, p, Ap, rz, rhs_norm, best_u, best_res, res, improved, state, tol, near_tol,
                        iteration, latched, ws, stream);
}

extern "C" int pgj_pcg32_post_step(int B, int64_t n, const float* r, float* z, const float* minv,
                                   int64_t minv_stride, float* p, double* rz, double* ws, void* stream) {
  return pcg_post_step_t(B, n, r, z, minv, minv_stride, p, rz, ws, stream);
}
