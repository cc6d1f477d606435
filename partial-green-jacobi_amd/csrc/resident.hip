// Resident multi-load PCG: set-up, certification, initialization and the
// fused iteration (four kernels per iteration, no host round trip):
//   1. hex8_rowmarch<PCG>: q = K p, p.q partials
//   2. fxy<PCG>          : alpha = rz / p.q; r -= alpha q; forward xy FFT of s*r
//   3. fz<PCG>           : masks, latches, iteration count; z-FFT, projection, Green, Parseval partials
//   4. fxyi<PCG>         : inverse FFT, z = P(s*w), next direction p = z + beta p (z alone after init),
//                          u += alpha p (also for a load stopped by this iteration's test);
//                          publishes the new active mask. z itself is never stored.
// Per-load scalar semantics follow exp/asg_gpu_resident_20261103/resident.py:39-58 and 119-136.
#include <cmath>

#include "fft_kernels.h"
#include "hex8_kernels.h"

namespace pgj {

__global__ __launch_bounds__(256) void square_partials(const double* v, int64_t N, double* part) {
  __shared__ double red[8];
  const int l = blockIdx.y;
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  double s = 0.0;
  if (i < N)
    for (int d = 0; d < 3; ++d) {
      const double x = v[int64_t(l) * 3 * N + 3 * i + d];
      s = fma(x, x, s);
    }
  const double t = block_sum(s, red);
  if (threadIdx.x == 0) part[(int64_t(l) * gridDim.x + blockIdx.x) * 4 + 3] = t;
}

__global__ void setup_finish(const double* part, int nblk, double rtol, double atol, int64_t cap, StateView st) {
  const int l = blockIdx.x;
  const double s = wave_reduce_partials(part + int64_t(l) * nblk * 4 + 3, nblk, 4);
  if (threadIdx.x == 0) {
    const double norm = sqrt(s);
    // max(rtol |f|, atol) that keeps a NaN norm, as torch.maximum (resident.py:191): the load then never
    // activates and fails certification
    const double rt = norm * rtol;
    const double t = (isnan(rt) || rt > atol) ? rt : atol;
    st.norm[l] = norm;
    st.target2[l] = t * t;
    st.cap[l] = cap;
    st.it[l] = 0;
    st.bad[l] = 0;
    st.active[l] = 0;
    st.first[l] = 0;
    st.stopped[l] = 0;
    st.valid[l] = 0;
  }
}

__global__ void certify_finish(const double* part, int nblk, StateView st) {
  const int l = blockIdx.x;
  const double s = wave_reduce_partials(part + int64_t(l) * nblk * 4 + 3, nblk, 4);
  if (threadIdx.x == 0) st.absolute[l] = sqrt(s);
}

template <typename T>
struct ResidentBuffersT {
  ResidentLayout lay;
  double* ppq;
  double* pcert;
  cplx<T>* spec;
  double* pr;
  double* pg;
  double* rscale;
};
using ResidentBuffers = ResidentBuffersT<double>;

// Workspace regions of a resident set of L loads (spectrum in precision T);
// the per-load rows start at load l0.
template <typename T = double>
ResidentBuffersT<T> resident_buffers(const pgj_grid& g, int L, void* ws, int l0 = 0) {
  ResidentBuffersT<T> b{};
  b.lay = resident_layout(g, L, sizeof(cplx<T>));
  const FFTGeom f = fft_geom(g);
  char* w = static_cast<char*>(ws);
  b.ppq = reinterpret_cast<double*>(w + b.lay.ppq) + int64_t(l0) * hex8_blocks_bound(g);
  b.pcert = reinterpret_cast<double*>(w + b.lay.pcert) + int64_t(l0) * b.lay.nblk_cert * 4;
  b.spec = reinterpret_cast<cplx<T>*>(w + b.lay.pre.spec) + int64_t(l0) * 3 * f.nf;
  b.pr = reinterpret_cast<double*>(w + b.lay.pre.pr) + int64_t(l0) * g.nz * 6;
  b.pg = reinterpret_cast<double*>(w + b.lay.pre.pg) + int64_t(l0) * f.nkb * g.ny * 4;
  b.rscale = reinterpret_cast<double*>(w + b.lay.rscale) + l0;
  return b;
}

// Loads [l0, l0 + nl) of a resident set of L loads in designs of `group` loads.
struct Slice {
  int L, group, l0, nl;
  int64_t N, nf;
  int64_t vec() const { return int64_t(l0) * 3 * N; }          // offset into (L, 3N) vectors
  int64_t design() const { return int64_t(l0 / group); }      // first design of the slice
};

}  // namespace pgj

using namespace pgj;

extern "C" size_t pgj_resident_state_bytes(int L) { return state_layout(L).bytes; }

extern "C" int pgj_resident_state_offsets(int L, size_t* o) {
  PGJ_REQUIRE(L >= 1 && o, "invalid state layout query");
  StateLayout s = state_layout(L);
  o[0] = s.rz; o[1] = s.target2; o[2] = s.norm; o[3] = s.absolute; o[4] = s.it; o[5] = s.active; o[6] = s.bad;
  return PGJ_OK;
}

extern "C" size_t pgj_resident_ws_bytes(const pgj_grid* g, int L) { return resident_layout(*g, L).bytes; }

namespace {
int make_slice(const pgj_grid* g, int L, int group, int l0, int nl, Slice* out) {
  PGJ_REQUIRE(L >= 1 && L <= kMaxResidentLoads && group >= 1 && L % group == 0,
              "resident set needs 1 <= L <= %d loads in designs of `group` loads, got L=%d group=%d",
              kMaxResidentLoads, L, group);
  // whole designs, or loads of one design (the design offset of the slice is that design's, and every load
  // inside the slice maps to it)
  const bool whole = l0 % group == 0 && nl % group == 0;
  const bool within = l0 >= 0 && nl >= 1 && l0 / group == (l0 + nl - 1) / group;
  PGJ_REQUIRE(l0 >= 0 && nl >= 1 && l0 + nl <= L && (whole || within),
              "load slice [%d, %d) must cover whole designs of %d loads, or lie in one, within %d", l0, l0 + nl,
              group, L);
  out->L = L; out->group = group; out->l0 = l0; out->nl = nl;
  out->N = int64_t(g->nx) * g->ny * g->nz;
  out->nf = int64_t(g->nz) * g->ny * (g->nx / 2 + 1);
  return PGJ_OK;
}
}  // namespace

extern "C" int pgj_resident_setup(const pgj_grid* g, int L, int group, int l0, int nl, const double* a,
                                  const double* macros, double rtol, int64_t cap, double* rhs, void* state,
                                  void* ws, void* stream) {
  if (int e = check_grid(g)) return e;
  Slice sl{};
  if (int e = make_slice(g, L, group, l0, nl, &sl)) return e;
  PGJ_REQUIRE(a && macros && rhs && state && ws, "invalid resident setup arguments");
  PGJ_REQUIRE(cap >= 1, "iteration cap must be at least 1, got %lld", (long long)cap);
  const int64_t N = sl.N;
  hipStream_t s = as_stream(stream);
  double* f = rhs + sl.vec();
  for (int l = 0; l < nl; ++l)
    if (int e = launch_rhs(*g, a + (sl.design() + l / group) * N, macros + 6 * l, f + int64_t(l) * 3 * N, 1, s))
      return e;
  ResidentBuffers b = resident_buffers<double>(*g, L, ws, l0);
  const int nblk = b.lay.nblk_cert;
  hipLaunchKernelGGL(square_partials, dim3(nblk, nl), dim3(256), 0, s, f, N, b.pcert);
  PGJ_LAUNCHED("square_partials");
  hipLaunchKernelGGL(setup_finish, dim3(nl), dim3(64), 0, s, b.pcert, nblk, rtol, 1e-13 / sqrt(double(g->nx)),
                     cap, state_view(state, L, l0));
  PGJ_LAUNCHED("setup_finish");
  return PGJ_OK;
}

extern "C" int pgj_resident_certify(const pgj_grid* g, int L, int group, int l0, int nl, const double* a,
                                    double* u, const double* rhs, double* r, void* state, void* ws, void* stream) {
  if (int e = check_grid(g)) return e;
  Slice sl{};
  if (int e = make_slice(g, L, group, l0, nl, &sl)) return e;
  PGJ_REQUIRE(a && u && rhs && r && state && ws, "invalid certify arguments");
  hipStream_t s = as_stream(stream);
  ResidentBuffers b = resident_buffers<double>(*g, L, ws, l0);
  double* us = u + sl.vec();
  if (int e = launch_component_sums(*g, nl, us, b.pcert, 1, s)) return e;
  if (int e = launch_subtract_mean(*g, nl, us, b.pcert, 1, s)) return e;
  RefArgs A{};
  A.a = a + sl.design() * sl.N; A.u = us; A.out = r + sl.vec(); A.rhs = rhs + sl.vec(); A.pres = b.pcert;
  A.soa = 1; A.group = group;
  if (int e = launch_reference(*g, nl, A, true, s)) return e;
  hipLaunchKernelGGL(certify_finish, dim3(nl), dim3(64), 0, s, b.pcert, b.lay.nblk_cert, state_view(state, L, l0));
  PGJ_LAUNCHED("certify_finish");
  return PGJ_OK;
}

namespace pgj {
// FFT/Green stage launchers: compile-time shapes where they exist; FP64 also
// has the runtime-shaped kernels, FP32 does not.
template <typename T>
int fxy(FFTArgsT<T>& A, const pgj_grid& g, int L, int mode, hipStream_t s) {
  if (static_fft(g)) return launch_fxy_static(A, L, mode, s);
  if constexpr (sizeof(T) == 8) return launch_fxy_mode(A, L, mode, s);
  set_error("FP32 FFT kernels exist for cubic n in 16, 24, 32, 48, 64, 96, 128 only");
  return PGJ_ERR_UNSUPPORTED;
}
template <typename T>
int fz(FFTArgsT<T>& A, const pgj_grid& g, int L, int mode, hipStream_t s) {
  if (static_fft(g)) return launch_fz_static(A, L, mode, s);
  if constexpr (sizeof(T) == 8) return launch_fz_mode(A, L, mode, s);
  set_error("FP32 FFT kernels exist for cubic n in 16, 24, 32, 48, 64, 96, 128 only");
  return PGJ_ERR_UNSUPPORTED;
}
template <typename T>
int fxyi(FFTArgsT<T>& A, const pgj_grid& g, int L, hipStream_t s) {
  if (static_fft(g)) return launch_fxyi_static(A, L, s);
  if constexpr (sizeof(T) == 8) return launch_fxyi_mode(A, L, kFxyiPCG, s);
  set_error("FP32 FFT kernels exist for cubic n in 16, 24, 32, 48, 64, 96, 128 only");
  return PGJ_ERR_UNSUPPORTED;
}

// First direction p = M^-1 r on a slice with the active-mask set-up (precision T); A carries the plan.
template <typename T>
int resident_init_t(const pgj_grid* g, int L, int group, int l0, int nl, FFTArgsT<T> A, const T* sc,
                    const cplx<T>* s_hat, const T* r, T* p, void* state, void* ws, hipStream_t s) {
  Slice sl{};
  if (int e = make_slice(g, L, group, l0, nl, &sl)) return e;
  PGJ_REQUIRE(r && p && state && ws, "invalid init arguments");
  PGJ_REQUIRE((sc == nullptr) == (s_hat == nullptr), "s and s_hat must be given together");
  ResidentBuffersT<T> b = resident_buffers<T>(*g, L, ws, l0);
  A.rin = r + sl.vec();
  A.s = sc ? sc + sl.design() * 3 * sl.N : nullptr;
  A.shat = s_hat ? s_hat + sl.design() * 3 * sl.nf : nullptr;
  A.spec = b.spec; A.pr = b.pr; A.pg = b.pg; A.zout = p + sl.vec();
  A.p = nullptr;   // right after init the direction is the preconditioned residual itself
  A.ppq = b.ppq; A.ppq_count = hex8_blocks_per_load(*g, nl, int(sizeof(T))); A.ppq_stride = hex8_blocks_bound(*g);
  A.st = state_view(state, L, l0);
  A.soa = 1;
  A.group = group;
  A.zwide = zpass_variant(*g, L, int(sizeof(T)));  // from the workspace's load count: every slice agrees
  if (int e = fxy(A, *g, nl, kFxyPre, s)) return e;
  if (int e = fz(A, *g, nl, kFzInit, s)) return e;
  return fxyi(A, *g, nl, s);
}

// One of the four fused iteration stages on a slice (precision T). The
// direction of this iteration is p (read by stages 0, 1, 3); stage 3 writes
// the next direction to p_next.
// The operator of stage 0: the scalar SIMP field a (the row-march), or a general-C21 pool (the tiled apply).
struct C21Operand {
  const double* pool;
  const int32_t* index;
};

template <typename T>
int resident_stage_t(const pgj_grid* g, int L, int group, int l0, int nl, FFTArgsT<T> F, const T* a, const T* sc,
                     const cplx<T>* s_hat, T* u, T* r, T* p, T* p_next, T* q, void* state, void* ws, int stage,
                     hipStream_t s, const C21Operand* c21 = nullptr) {
  Slice sl{};
  if (int e = make_slice(g, L, group, l0, nl, &sl)) return e;
  PGJ_REQUIRE((a || c21) && u && r && p && p_next && q && state && ws, "invalid stage arguments");
  PGJ_REQUIRE(stage >= 0 && stage <= 3, "stage must be 0..3, got %d", stage);
  PGJ_REQUIRE((sc == nullptr) == (s_hat == nullptr), "s and s_hat must be given together");
  PGJ_REQUIRE(c21 || rowmarch_config(*g, nl).valid, "resident iteration needs nx <= 128 with nx %% ceil(nx/64) == 0");
  ResidentBuffersT<T> b = resident_buffers<T>(*g, L, ws, l0);
  StateView st = state_view(state, L, l0);
  const int64_t v = sl.vec();
  F.u = u + v; F.r = r + v; F.q = q + v; F.p = p + v;
  F.s = sc ? sc + sl.design() * 3 * sl.N : nullptr;
  F.shat = s_hat ? s_hat + sl.design() * 3 * sl.nf : nullptr;
  F.spec = b.spec; F.pr = b.pr; F.pg = b.pg; F.zout = p_next + v;
  F.ppq = b.ppq; F.ppq_count = hex8_blocks_per_load(*g, nl, int(sizeof(T))); F.ppq_stride = hex8_blocks_bound(*g);
  if constexpr (sizeof(T) == 8) {
    if (c21) {
      if (int e = c21_pcg_blocks(*g, nl, &F.ppq_count)) return e;
      PGJ_REQUIRE(F.ppq_count <= F.ppq_stride, "tiled apply blocks (%d) exceed the partial rows (%d)", F.ppq_count,
                  F.ppq_stride);
    }
  }
  F.st = st;
  F.soa = 1;
  F.group = group;
  F.zwide = zpass_variant(*g, L, int(sizeof(T)));
  if (stage == 0) {
    if constexpr (sizeof(T) == 8) {
      if (c21) return launch_c21_pcg(*g, nl, c21->pool, c21->index, p + v, st.active, q + v, b.ppq, F.ppq_stride, s);
    }
    RowArgsT<T> K{};
    K.a = a + sl.design() * sl.N; K.u = p + v; K.out = q + v; K.st = st;
    K.ppq = b.ppq; K.ppq_stride = F.ppq_stride;
    K.group = group;
    return launch_rowmarch(*g, nl, kModePCG, K, s);
  }
  if (stage == 1) return fxy(F, *g, nl, kFxyPCG, s);
  if (stage == 2) return fz(F, *g, nl, kFzPCG, s);
  return fxyi(F, *g, nl, s);
}
}  // namespace pgj

extern "C" int pgj_resident_scale(const pgj_grid* g, const void* plan, const double* a, double alpha, double* sc,
                                  void* s_hat, void* ws, void* stream) {
  if (int e = check_fft(g)) return e;
  PGJ_REQUIRE(plan && a && sc && s_hat && ws, "invalid resident scale arguments");
  PGJ_REQUIRE(alpha >= 0.0 && alpha <= 1.0, "alpha must be in [0, 1], got %g", alpha);
  hipStream_t s = as_stream(stream);
  ResidentBuffers b = resident_buffers<double>(*g, 1, ws);
  if (int e = launch_scale(*g, a, alpha, nullptr, sc, b.pcert, 1, s)) return e;
  FFTArgs A = fft_args(*g, plan);
  A.s = sc; A.spec = static_cast<double2*>(s_hat); A.pr = b.pr; A.pg = b.pg; A.soa = 1;
  if (int e = fxy(A, *g, 1, kFxyScale, s)) return e;
  return fz(A, *g, 1, kFzScale, s);
}

extern "C" int pgj_resident_scale_dev(const pgj_grid* g, const void* plan, const double* a, const double* alpha,
                                      double* sc, void* s_hat, void* ws, void* stream) {
  if (int e = check_fft(g)) return e;
  PGJ_REQUIRE(plan && a && alpha && sc && s_hat && ws, "invalid resident scale arguments");
  hipStream_t s = as_stream(stream);
  ResidentBuffers b = resident_buffers<double>(*g, 1, ws);
  if (int e = launch_scale(*g, a, 0.0, nullptr, sc, b.pcert, 1, s, alpha)) return e;
  FFTArgs A = fft_args(*g, plan);
  A.s = sc; A.spec = static_cast<double2*>(s_hat); A.pr = b.pr; A.pg = b.pg; A.soa = 1;
  if (int e = fxy(A, *g, 1, kFxyScale, s)) return e;
  return fz(A, *g, 1, kFzScale, s);
}

extern "C" int pgj_resident_init(const pgj_grid* g, int L, int group, int l0, int nl, const void* plan,
                                 const double* sc, const void* s_hat, const double* r, double* p, void* state,
                                 void* ws, void* stream) {
  if (int e = check_fft(g)) return e;
  PGJ_REQUIRE(plan, "invalid init arguments");
  return resident_init_t<double>(g, L, group, l0, nl, fft_args(*g, plan), sc, static_cast<const double2*>(s_hat), r,
                                 p, state, ws, as_stream(stream));
}

extern "C" int pgj_resident_stage(const pgj_grid* g, int L, int group, int l0, int nl, const void* plan,
                                  const double* a, const double* sc, const void* s_hat, double* u, double* r,
                                  double* p, double* p_next, double* q, void* state, void* ws, int stage,
                                  void* stream) {
  if (int e = check_fft(g)) return e;
  PGJ_REQUIRE(plan, "invalid stage arguments");
  return resident_stage_t<double>(g, L, group, l0, nl, fft_args(*g, plan), a, sc, static_cast<const double2*>(s_hat),
                                  u, r, p, p_next, q, state, ws, stage, as_stream(stream));
}

extern "C" int pgj_resident_iterate(const pgj_grid* g, int L, int group, int l0, int nl, const void* plan,
                                    const double* a, const double* sc, const void* s_hat, double* u, double* r,
                                    double* p0, double* p1, double* q, void* state, void* ws, int iters,
                                    void* stream) {
  PGJ_REQUIRE(iters >= 2 && iters % 2 == 0, "iteration count must be even and >= 2, got %d", iters);
  for (int i = 0; i < iters; ++i) {
    double* p = (i % 2 == 0) ? p0 : p1;
    double* p_next = (i % 2 == 0) ? p1 : p0;
    for (int stage = 0; stage < 4; ++stage)
      if (int e = pgj_resident_stage(g, L, group, l0, nl, plan, a, sc, s_hat, u, r, p, p_next, q, state, ws,
                                     stage, stream))
        return e;
  }
  return PGJ_OK;
}

// ------------------------------------------ general-C21 operator (FP64)
// The resident solver over a general-C21 material (one design: group = L): the tiled apply in PCG mode as
// stage 0, the same FFT stages, and the gather apply for set-up and certification.

extern "C" int pgj_hybrid_resident_setup(const pgj_grid* g, int L, int l0, int nl, const double* pool,
                                         const int32_t* index, const double* macros, double rtol, int64_t cap,
                                         double* rhs, void* state, void* ws, void* stream) {
  if (int e = check_grid(g)) return e;
  Slice sl{};
  if (int e = make_slice(g, L, L, l0, nl, &sl)) return e;
  PGJ_REQUIRE(pool && macros && rhs && state && ws, "invalid hybrid resident setup arguments");
  PGJ_REQUIRE(cap >= 1, "iteration cap must be at least 1, got %lld", (long long)cap);
  const int64_t N = sl.N;
  hipStream_t s = as_stream(stream);
  double* f = rhs + sl.vec();
  for (int l = 0; l < nl; ++l)
    if (int e = launch_c21_rhs(*g, pool, index, macros + 6 * l, f + int64_t(l) * 3 * N, 1, s)) return e;
  ResidentBuffers b = resident_buffers<double>(*g, L, ws, l0);
  const int nblk = b.lay.nblk_cert;
  hipLaunchKernelGGL(square_partials, dim3(nblk, nl), dim3(256), 0, s, f, N, b.pcert);
  PGJ_LAUNCHED("square_partials");
  hipLaunchKernelGGL(setup_finish, dim3(nl), dim3(64), 0, s, b.pcert, nblk, rtol, 1e-13 / sqrt(double(g->nx)),
                     cap, state_view(state, L, l0));
  PGJ_LAUNCHED("setup_finish");
  return PGJ_OK;
}

extern "C" int pgj_hybrid_resident_certify(const pgj_grid* g, int L, int l0, int nl, const double* pool,
                                           const int32_t* index, double* u, const double* rhs, double* r,
                                           void* state, void* ws, void* stream) {
  if (int e = check_grid(g)) return e;
  Slice sl{};
  if (int e = make_slice(g, L, L, l0, nl, &sl)) return e;
  PGJ_REQUIRE(pool && u && rhs && r && state && ws, "invalid hybrid certify arguments");
  hipStream_t s = as_stream(stream);
  ResidentBuffers b = resident_buffers<double>(*g, L, ws, l0);
  double* us = u + sl.vec();
  if (int e = launch_component_sums(*g, nl, us, b.pcert, 1, s)) return e;
  if (int e = launch_subtract_mean(*g, nl, us, b.pcert, 1, s)) return e;
  if (int e = launch_c21_certify(*g, nl, pool, index, us, rhs + sl.vec(), r + sl.vec(), b.pcert, s)) return e;
  hipLaunchKernelGGL(certify_finish, dim3(nl), dim3(64), 0, s, b.pcert, b.lay.nblk_cert, state_view(state, L, l0));
  PGJ_LAUNCHED("certify_finish");
  return PGJ_OK;
}

extern "C" int pgj_hybrid_resident_stage(const pgj_grid* g, int L, int l0, int nl, const void* plan,
                                         const double* pool, const int32_t* index, const double* sc,
                                         const void* s_hat, double* u, double* r, double* p, double* p_next,
                                         double* q, void* state, void* ws, int stage, void* stream) {
  if (int e = check_fft(g)) return e;
  PGJ_REQUIRE(plan && pool, "invalid hybrid stage arguments");
  const C21Operand c21{pool, index};
  return resident_stage_t<double>(g, L, L, l0, nl, fft_args(*g, plan), nullptr, sc,
                                  static_cast<const double2*>(s_hat), u, r, p, p_next, q, state, ws, stage,
                                  as_stream(stream), &c21);
}

extern "C" int pgj_hybrid_resident_iterate(const pgj_grid* g, int L, int l0, int nl, const void* plan,
                                           const double* pool, const int32_t* index, const double* sc,
                                           const void* s_hat, double* u, double* r, double* p0, double* p1,
                                           double* q, void* state, void* ws, int iters, void* stream) {
  PGJ_REQUIRE(iters >= 2 && iters % 2 == 0, "iteration count must be even and >= 2, got %d", iters);
  for (int i = 0; i < iters; ++i) {
    double* p = (i % 2 == 0) ? p0 : p1;
    double* p_next = (i % 2 == 0) ? p1 : p0;
    for (int stage = 0; stage < 4; ++stage)
      if (int e = pgj_hybrid_resident_stage(g, L, l0, nl, plan, pool, index, sc, s_hat, u, r, p, p_next, q, state,
                                            ws, stage, stream))
        return e;
  }
  return PGJ_OK;
}

// ------------------------------------------------- FP32 inner resident solver
namespace pgj {
FFTArgsT<float> fft_args32(const pgj_grid& g, const void* plan32);
}

extern "C" int pgj_inner_init(const pgj_grid* g, int L, int group, int l0, int nl, const void* plan32,
                              const float* sc, const void* s_hat, const float* r, float* p, void* state, void* ws,
                              void* stream) {
  PGJ_REQUIRE(g && plan32, "invalid inner init arguments");
  PGJ_REQUIRE(static_cube(*g), "the FP32 inner solver needs a cubic n in 16, 24, 32, 48, 64, 96, 128");
  return resident_init_t<float>(g, L, group, l0, nl, fft_args32(*g, plan32), sc, static_cast<const float2*>(s_hat), r,
                                p, state, ws, as_stream(stream));
}

extern "C" int pgj_inner_stage(const pgj_grid* g, int L, int group, int l0, int nl, const void* plan32,
                               const float* a, const float* sc, const void* s_hat, float* u, float* r, float* p,
                               float* p_next, float* q, void* state, void* ws, int stage, void* stream) {
  PGJ_REQUIRE(g && plan32, "invalid inner stage arguments");
  PGJ_REQUIRE(static_cube(*g), "the FP32 inner solver needs a cubic n in 16, 24, 32, 48, 64, 96, 128");
  return resident_stage_t<float>(g, L, group, l0, nl, fft_args32(*g, plan32), a, sc, static_cast<const float2*>(s_hat),
                                 u, r, p, p_next, q, state, ws, stage, as_stream(stream));
}

extern "C" int pgj_inner_iterate(const pgj_grid* g, int L, int group, int l0, int nl, const void* plan32,
                                 const float* a, const float* sc, const void* s_hat, float* u, float* r, float* p0,
                                 float* p1, float* q, void* state, void* ws, int iters, void* stream) {
  PGJ_REQUIRE(iters >= 2 && iters % 2 == 0, "iteration count must be even and >= 2, got %d", iters);
  for (int i = 0; i < iters; ++i) {
    float* p = (i % 2 == 0) ? p0 : p1;
    float* p_next = (i % 2 == 0) ? p1 : p0;
    for (int stage = 0; stage < 4; ++stage)
      if (int e = pgj_inner_stage(g, L, group, l0, nl, plan32, a, sc, s_hat, u, r, p, p_next, q, state, ws, stage,
                                  stream))
        return e;
  }
  return PGJ_OK;
}
