// Declarations shared by the FFT/Green kernels and the resident solver.
#pragma once

#include "layout.h"
#include "pgj_internal.h"

namespace pgj {

// ------------------------------------------------------------------ arguments
enum { kFxyScale = 0, kFxyPre = 1, kFxyPCG = 2 };
enum { kFzScale = 0, kFzPre = 1, kFzInit = 2, kFzPCG = 3 };
enum { kFxyiPre = 0, kFxyiPCG = 1 };
// FFTArgs::zwide value of the register-resident z pass of the 64^3 cube (fz64_reg).
constexpr int kZReg = 2;

// Arguments of the FFT/Green kernels; T is the vector precision (FP64 plain
// path, FP32 inner solver of the mixed-precision form). Partial sums and the
// PCG scalars stay FP64 in both.
template <typename T>
struct FFTArgsT {
  using C = cplx<T>;
  int nz, ny, nx;
  int64_t N;
  FFTGeom f;
  const C* twx;
  const C* twy;
  const C* twz;
  const T* sym;
  const T* rin;          // fxy pre: input residual
  T* u;                  // fxy pcg
  T* r;                  // fxy pcg
  const T* p;            // fxyi pcg: direction of this iteration (u += alpha p, p_next = z + beta p)
  const T* q;            // fxy pcg
  const T* s;            // scale (nullable)
  const C* shat;         // spectrum of s (nullable)
  C* spec;               // spectrum buffer (L x 3 x nf)
  T* zout;               // fxyi output: z (preconditioner), or the next PCG direction
  double* pr;            // fxy partials  [l][z][c][2] = (sum r^2, sum r)
  double* pg;            // fz partials   [l][blk][4]  = (v.w, s.w_x, s.w_y, s.w_z)
  const double* ppq;     // hex8 partials [l][blk], ppq_count blocks in rows of ppq_stride
  int ppq_count, ppq_stride;
  StateView st;
  int soa;               // vectors (and s) are SoA (L, 3, N) instead of AoS (L, N, 3)
  int group;             // loads per design (s is (designs, 3N), s_hat (designs, 3, nf)); 0: shared
  int zwide;             // z pass variant of a cube (zpass_variant): 0 default kx chunks, 1 wide chunks,
                         // kZReg register-resident; the same for every launch that writes or reads one
                         // workspace's z partials
  const double* rin64;   // FP32 preconditioner apply on AoS vectors: FP64 input residual, rounded to T on load
  double* zout64;        // ... and FP64 output z (exact widening of the T result); both null otherwise
};
using FFTArgs = FFTArgsT<double>;

// Index of component c of node `node` of load l in a vector (and of s with l = 0).
__device__ __forceinline__ int64_t vindex(const FFTArgs& A, int l, int c, int64_t node) {
  return A.soa ? (int64_t(l) * 3 + c) * A.N + node : (int64_t(l) * A.N + node) * 3 + c;
}

FFTArgs fft_args(const pgj_grid& g, const void* plan);
int check_fft(const pgj_grid* g);
int launch_fxy_mode(FFTArgs& A, int L, int mode, hipStream_t s);
int launch_fz_mode(FFTArgs& A, int L, int mode, hipStream_t s);
int launch_fxyi_mode(FFTArgs& A, int L, int mode, hipStream_t s);
// Compile-time-shaped SoA kernels for cubic grids (n in 16, 24, 32, 48, 64, 96, 128).
bool static_fft(const pgj_grid& g);
// Whether the z pass of a solve with L loads in all (its workspace's load count) and elements of es bytes
// takes the wide kx chunks: where the default chunks need more than one round of blocks on the device
// (64^3 FP64 from 6 loads on 256 CUs).
bool zpass_wide(const pgj_grid& g, int L, int es);
// The z pass variant of the compile-time kernels for such a solve: kZReg where the 64^3 z pass runs on its
// register-resident plan (pgj_fft_set_plan pass 1, plan 2), else the wide chunks where zpass_wide holds, else 0.
int zpass_variant(const pgj_grid& g, int L, int es);
template <typename T> int launch_fxy_static(FFTArgsT<T>& A, int L, int mode, hipStream_t s);
template <typename T> int launch_fz_static(FFTArgsT<T>& A, int L, int mode, hipStream_t s);
template <typename T> int launch_fxyi_static(FFTArgsT<T>& A, int L, hipStream_t s, int mode = kFxyiPCG);

}  // namespace pgj
