// Shared helpers for the libpgj HIP sources: error reporting, launch checks,
// index arithmetic and wavefront/block reductions with a fixed summation order.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>

#include "../../include/pgj.h"

namespace pgj {

void set_error(const char* fmt, ...);
void clear_error();

#define PGJ_REQUIRE(cond, ...)          \
  do {                                  \
    if (!(cond)) {                      \
      ::pgj::set_error(__VA_ARGS__);    \
      return PGJ_ERR_ARG;               \
    }                                   \
  } while (0)

#define PGJ_LAUNCHED(name)                                                        \
  do {                                                                            \
    hipError_t e_ = hipGetLastError();                                            \
    if (e_ != hipSuccess) {                                                       \
      ::pgj::set_error("%s launch failed: %s", name, hipGetErrorString(e_));      \
      return PGJ_ERR_HIP;                                                         \
    }                                                                             \
  } while (0)

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

// Complex value type of a real type: double2 for FP64, float2 for FP32.
template <typename T> struct Complex;
template <> struct Complex<double> { using type = double2; };
template <> struct Complex<float> { using type = float2; };
template <typename T> using cplx = typename Complex<T>::type;

__device__ __forceinline__ int wrapi(int v, int n) { return v < 0 ? v + n : (v >= n ? v - n : v); }

// Design of load l when consecutive groups of `group` loads share one material
// field a, scale s and spectrum s_hat (group 0: every load shares design 0).
__host__ __device__ __forceinline__ int design_of(int l, int group) { return group > 0 ? l / group : 0; }

// Sum over the 64 lanes of a wavefront with a fixed butterfly order; every lane gets the total.
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Block-wide sum with fixed order (lanes, then waves in index order). `scratch`
// must hold blockDim.x/64 doubles. All threads receive the total.
__device__ __forceinline__ double block_sum(double v, double* scratch) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
  v = wave_sum(v);
  __syncthreads();
  if (lane == 0) scratch[wave] = v;
  __syncthreads();
  double t = 0.0;
  for (int w = 0; w < nw; ++w) t += scratch[w];
  __syncthreads();
  return t;
}

// Block-wide sums of K values at once (fixed order, one pair of barriers).
// `scratch` must hold K * blockDim.x/64 doubles; every thread receives the totals.
template <int K>
__device__ __forceinline__ void block_sums(double (&v)[K], double* scratch) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
#pragma unroll
  for (int k = 0; k < K; ++k) v[k] = wave_sum(v[k]);
  __syncthreads();
  if (lane == 0)
#pragma unroll
    for (int k = 0; k < K; ++k) scratch[k * nw + wave] = v[k];
  __syncthreads();
#pragma unroll
  for (int k = 0; k < K; ++k) {
    double t = 0.0;
    for (int w = 0; w < nw; ++w) t += scratch[k * nw + w];
    v[k] = t;
  }
  __syncthreads();
}

// Deterministic reduction of `count` partials at `p` (stride `stride`) by one
// wavefront: lane-strided accumulation followed by wave_sum. Every lane of
// every calling wavefront obtains bit-identical results. The loop is unrolled
// so that a lane's loads are in flight together; the additions keep their order.
__device__ __forceinline__ double wave_reduce_partials(const double* p, int count, int stride) {
  const int lane = threadIdx.x & 63;
  double v = 0.0;
#pragma unroll 8
  for (int i = lane; i < count; i += 64) v += p[(int64_t)i * stride];
  return wave_sum(v);
}

// K reductions of `count` entries of stride `stride` in one lane-strided pass
// (reduction k reads p[i * stride + off[k]]), so their loads are in flight
// together. Each accumulates in the order of wave_reduce_partials, so the
// results are bit-identical to K separate calls.
template <int K>
__device__ __forceinline__ void wave_reduce_partials_n(const double* p, const int (&off)[K], int count, int stride,
                                                       double (&out)[K]) {
  const int lane = threadIdx.x & 63;
  double v[K];
#pragma unroll
  for (int k = 0; k < K; ++k) v[k] = 0.0;
#pragma unroll 8
  for (int i = lane; i < count; i += 64) {
    const double* q = p + (int64_t)i * stride;
#pragma unroll
    for (int k = 0; k < K; ++k) v[k] += q[off[k]];
  }
#pragma unroll
  for (int k = 0; k < K; ++k) out[k] = wave_sum(v[k]);
}

// Two reductions over different partial arrays in one pass (bit-identical to
// two wave_reduce_partials calls).
__device__ __forceinline__ void wave_reduce_partials_2(const double* p, int pc, int ps, const double* q, int qc, int qs,
                                                       double& a, double& b) {
  const int lane = threadIdx.x & 63;
  double va = 0.0, vb = 0.0;
  const int n = pc > qc ? pc : qc;
#pragma unroll 8
  for (int i = lane; i < n; i += 64) {
    if (i < pc) va += p[(int64_t)i * ps];
    if (i < qc) vb += q[(int64_t)i * qs];
  }
  a = wave_sum(va);
  b = wave_sum(vb);
}

}  // namespace pgj
