// Mixed-precision form: FP32 copies of the plan and fields, and the two ends
// of one FP64 iterative-refinement step around the FP32 inner resident solver.
//   begin: per load, s_l = |r_l| from the FP64 certification (0 once the load
//          is certified); r32 = r64 / s_l; d32 = 0; inner target |r32| <= tol.
//   end:   u64 += s_l * d32 (FP64 accumulation).
// The residual, the correction accumulation, certification and every
// reduction stay FP64; only the inner PCG vectors, the HEX8 apply and the
// spectral Green multiply run in FP32.
#include <cmath>

#include "fft_kernels.h"
#include "hex8_kernels.h"

namespace pgj {

__global__ void convert_f32(const double* x, float* y, int64_t n) {
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x)
    y[i] = float(x[i]);
}

// The inner target is rtol, or 1/2 of the reduction the outer target still
// needs if that is looser (the last refinement step then stops as soon as
// the FP64 certification can pass instead of over-solving).
__global__ void inner_scale(StateView outer, StateView inner, double* rscale, double rtol) {
  const int l = blockIdx.x;
  if (threadIdx.x != 0) return;
  const double abs_r = outer.absolute[l];
  const bool open = isfinite(abs_r) && abs_r * abs_r > outer.target2[l];
  rscale[l] = open ? abs_r : 0.0;
  inner.norm[l] = open ? 1.0 : 0.0;
  const double need = open ? 0.5 * sqrt(outer.target2[l]) / abs_r : 0.0;
  const double t = fmax(rtol, fmin(need, 0.5));
  inner.target2[l] = t * t;
  inner.it[l] = 0;
  inner.cap[l] = outer.cap[l];
  inner.bad[l] = 0;
  inner.active[l] = 0;
  inner.first[l] = 0;
  inner.stopped[l] = 0;
  inner.valid[l] = 0;
}

__global__ __launch_bounds__(256) void inner_convert(const double* r64, const double* rscale, float* r32, float* u32,
                                                     int64_t n3) {
  const int l = blockIdx.y;
  const double s = rscale[l];
  const double inv = s > 0.0 ? 1.0 / s : 0.0;
  const int64_t off = int64_t(l) * n3;
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n3; i += int64_t(gridDim.x) * blockDim.x) {
    r32[off + i] = float(r64[off + i] * inv);
    u32[off + i] = 0.0f;
  }
}

__global__ __launch_bounds__(256) void inner_accumulate(const float* u32, const double* rscale, double* u64,
                                                        int64_t n3) {
  const int l = blockIdx.y;
  const double s = rscale[l];
  if (s == 0.0) return;
  const int64_t off = int64_t(l) * n3;
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n3; i += int64_t(gridDim.x) * blockDim.x)
    u64[off + i] = fma(s, double(u32[off + i]), u64[off + i]);
}

inline unsigned vec_grid(int64_t n) {
  const int64_t b = (n + 1023) / 1024;
  return unsigned(b < 2048 ? b : 2048);
}

FFTArgsT<float> fft_args32(const pgj_grid& g, const void* plan32) {
  FFTArgsT<float> A{};
  A.nz = g.nz; A.ny = g.ny; A.nx = g.nx;
  A.N = int64_t(g.nx) * g.ny * g.nz;
  A.f = fft_geom(g);
  PlanLayout pl = plan_layout(g, 4);
  const char* b = static_cast<const char*>(plan32);
  A.twx = reinterpret_cast<const float2*>(b + pl.tw_x);
  A.twy = reinterpret_cast<const float2*>(b + pl.tw_y);
  A.twz = reinterpret_cast<const float2*>(b + pl.tw_z);
  A.sym = reinterpret_cast<const float*>(b + pl.sym);
  return A;
}

}  // namespace pgj

using namespace pgj;

extern "C" size_t pgj_plan32_bytes(const pgj_grid* g) { return plan_layout(*g, 4).bytes; }

extern "C" int pgj_to_f32(int64_t n, const double* x, float* y, void* stream) {
  PGJ_REQUIRE(n >= 0 && (n == 0 || (x && y)), "invalid conversion arguments");
  if (n == 0) return PGJ_OK;
  hipLaunchKernelGGL(convert_f32, dim3(vec_grid(n)), dim3(256), 0, as_stream(stream), x, y, n);
  PGJ_LAUNCHED("convert_f32");
  return PGJ_OK;
}

extern "C" int pgj_plan32_init(const pgj_grid* g, const void* plan, void* plan32, void* stream) {
  PGJ_REQUIRE(g && plan && plan32, "invalid plan32 arguments");
  PGJ_REQUIRE(static_cube(*g), "the FP32 inner solver needs a cubic n in 16, 24, 32, 48, 64, 96, 128");
  const PlanLayout p8 = plan_layout(*g, 8), p4 = plan_layout(*g, 4);
  const FFTGeom f = fft_geom(*g);
  const char* src = static_cast<const char*>(plan);
  char* dst = static_cast<char*>(plan32);
  struct Part { size_t from, to; int64_t count; };
  const Part parts[4] = {{p8.tw_x, p4.tw_x, 2 * int64_t(g->nx)}, {p8.tw_y, p4.tw_y, 2 * int64_t(g->ny)},
                         {p8.tw_z, p4.tw_z, 2 * int64_t(g->nz)}, {p8.sym, p4.sym, 6 * f.nf}};
  for (const Part& q : parts)
    if (int e = pgj_to_f32(q.count, reinterpret_cast<const double*>(src + q.from), reinterpret_cast<float*>(dst + q.to),
                           stream))
      return e;
  return PGJ_OK;
}

extern "C" int pgj_precond_apply_mixed(const pgj_grid* g, int L, const void* plan32, const double* r, const float* s,
                                       const void* s_hat, double* z, void* ws, void* stream) {
  PGJ_REQUIRE(g && static_cube(*g), "the FP32 preconditioner apply needs a cubic n in 16, 24, 32, 48, 64, 96, 128");
  PGJ_REQUIRE(L >= 1 && L <= 65535 && plan32 && r && z && ws, "invalid FP32 preconditioner arguments");
  PGJ_REQUIRE((s == nullptr) == (s_hat == nullptr), "s and s_hat must be given together");
  // the compile-time FP32 passes on the (L, 3N) node-interleaved vectors, r rounded to FP32 as it is read and z
  // written in FP64; the FP64 preconditioner workspace holds the complex64 spectrum with room to spare and the
  // same FP64 partial rows
  FFTArgsT<float> A = fft_args32(*g, plan32);
  const PrecondLayout pl = precond_layout(*g, L);
  char* w = static_cast<char*>(ws);
  A.rin64 = r;
  A.s = s;
  A.shat = static_cast<const float2*>(s_hat);
  A.spec = reinterpret_cast<float2*>(w + pl.spec);
  A.pr = reinterpret_cast<double*>(w + pl.pr);
  A.pg = reinterpret_cast<double*>(w + pl.pg);
  A.zout64 = z;
  A.soa = 0;
  hipStream_t st = as_stream(stream);
  if (int e = launch_fxy_static(A, L, kFxyPre, st)) return e;
  if (int e = launch_fz_static(A, L, kFzPre, st)) return e;
  return launch_fxyi_static(A, L, st, kFxyiPre);
}

extern "C" size_t pgj_inner_ws_bytes(const pgj_grid* g, int L) { return resident_layout(*g, L, 8).bytes; }

extern "C" int pgj_inner_begin(const pgj_grid* g, int L, int l0, int nl, const double* r64, const void* state64,
                               double rtol, float* r32, float* u32, void* state32, void* ws32, void* stream) {
  PGJ_REQUIRE(g && L >= 1 && L <= kMaxResidentLoads && l0 >= 0 && nl >= 1 && l0 + nl <= L,
              "invalid inner slice [%d, %d) of %d loads", l0, l0 + nl, L);
  PGJ_REQUIRE(r64 && state64 && r32 && u32 && state32 && ws32 && rtol > 0.0 && rtol < 1.0,
              "invalid inner begin arguments");
  hipStream_t s = as_stream(stream);
  const int64_t n3 = 3 * int64_t(g->nx) * g->ny * g->nz;
  const ResidentLayout lay = resident_layout(*g, L, 8);
  double* rscale = reinterpret_cast<double*>(static_cast<char*>(ws32) + lay.rscale) + l0;
  hipLaunchKernelGGL(inner_scale, dim3(nl), dim3(64), 0, s, state_view(const_cast<void*>(state64), L, l0),
                     state_view(state32, L, l0), rscale, rtol);
  PGJ_LAUNCHED("inner_scale");
  hipLaunchKernelGGL(inner_convert, dim3(vec_grid(n3), nl), dim3(256), 0, s, r64 + l0 * n3, rscale, r32 + l0 * n3,
                     u32 + l0 * n3, n3);
  PGJ_LAUNCHED("inner_convert");
  return PGJ_OK;
}

extern "C" int pgj_inner_end(const pgj_grid* g, int L, int l0, int nl, const float* u32, const void* ws32,
                             double* u64, void* stream) {
  PGJ_REQUIRE(g && L >= 1 && L <= kMaxResidentLoads && l0 >= 0 && nl >= 1 && l0 + nl <= L,
              "invalid inner slice [%d, %d) of %d loads", l0, l0 + nl, L);
  PGJ_REQUIRE(u32 && ws32 && u64, "invalid inner end arguments");
  const int64_t n3 = 3 * int64_t(g->nx) * g->ny * g->nz;
  const ResidentLayout lay = resident_layout(*g, L, 8);
  const double* rscale = reinterpret_cast<const double*>(static_cast<const char*>(ws32) + lay.rscale) + l0;
  hipLaunchKernelGGL(inner_accumulate, dim3(vec_grid(n3), nl), dim3(256), 0, as_stream(stream), u32 + l0 * n3,
                     rscale, u64 + l0 * n3, n3);
  PGJ_LAUNCHED("inner_accumulate");
  return PGJ_OK;
}
