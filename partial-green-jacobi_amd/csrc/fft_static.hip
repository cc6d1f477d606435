// Compile-time-shaped FFT/Green kernels of the resident solver for cubic
// grids n in {16, 24, 32, 48, 64, 96, 128}. Vectors and the scale s are SoA
// (L, 3, N): one block transforms one component plane, so a plane needs only
// n (n/2+1) complex values of LDS (34 KiB at 64^3) and every index is a
// constant expression. Semantics are those of the runtime kernels in fft.hip.
#include <cmath>

#include "fft_kernels.h"
#include "stockham.h"

namespace pgj {

template <int N>
struct Cube {
  static constexpr int M = N / 2;             // complex length of the packed x transform
  static constexpr int NKX = N / 2 + 1;       // half-spectrum width
  static constexpr int RS = NKX;              // LDS row stride (complex)
  static constexpr int PLANE = N * N;         // nodes per z-plane
  static constexpr int64_t NF = int64_t(N) * N * NKX;
  static constexpr int TXY = N >= 96 ? 512 : 256;
  static constexpr int TZ = 256;
  static constexpr int KC = N == 16 ? 9 : N == 24 ? 13 : N == 32 ? 17 : N == 48 ? 5 : N == 64 ? 11 : N == 96 ? 7 : 5;
  static constexpr int NKB = NKX / KC;
  static_assert(NKX % KC == 0, "kx chunking must divide the half spectrum");
  static constexpr size_t LDS_XY = size_t(N) * RS * 16 + 4 * 8 * 8;
  static constexpr size_t LDS_Z = size_t(3) * N * KC * 16 + 4 * 8 * 8;
  using Rows = Lines<N, N, 0, RS, 1, false>;          // x transforms (length M) on the N rows
  using Cols = Lines<NKX, NKX, 0, 1, RS, true>;       // y transforms (length N) on the NKX columns
  using ZCols = Lines<3 * KC, KC, N * KC, 1, KC, true>;  // z transforms on (component, column)
};

// ------------------------------------------------------------ forward xy
template <int N, int MODE>
__global__ __launch_bounds__(Cube<N>::TXY) void fxy_static(FFTArgs A) {
  using C = Cube<N>;
  constexpr int T = C::TXY;
  extern __shared__ double2 plane[];
  double* red = reinterpret_cast<double*>(plane + N * C::RS);
  const int z = blockIdx.x, c = blockIdx.y, l = blockIdx.z;
  const int64_t NN = int64_t(N) * N * N;
  double alpha = 0.0;
  if (MODE == kFxyPCG) {
    if (!A.st.active[l]) return;
    const double pq = wave_reduce_partials(A.ppq + int64_t(l) * A.ppq_stride, A.ppq_count, 1);
    const double rz = wave_reduce_partials(A.pg + int64_t(l) * C::NKB * N * 4, C::NKB * N, 4);
    const bool valid = isfinite(pq) && pq > 0.0 && isfinite(rz) && rz > 0.0;
    alpha = valid ? rz / pq : 0.0;
    if (threadIdx.x == 0 && z == 0 && c == 0) {
      A.st.rz[l] = rz;
      A.st.alpha[l] = alpha;
      A.st.first[l] = 0;
    }
  }
  const int64_t off = (int64_t(l) * 3 + c) * NN + int64_t(z) * C::PLANE;
  const int64_t soff = (int64_t(design_of(l, A.group)) * 3 + c) * NN + int64_t(z) * C::PLANE;
  double sq = 0.0, sm = 0.0;
  // two nodes (x even, x+1) per step: one packed complex of the x transform
  for (int e = 2 * threadIdx.x; e < C::PLANE; e += 2 * T) {
    double2 v;
    if (MODE == kFxyScale) {
      v = *reinterpret_cast<const double2*>(A.s + soff + e);
    } else if (MODE == kFxyPre) {
      const double2 rv = *reinterpret_cast<const double2*>(A.rin + off + e);
      sq = fma(rv.x, rv.x, fma(rv.y, rv.y, sq));
      sm += rv.x + rv.y;
      if (A.s) {
        const double2 sv = *reinterpret_cast<const double2*>(A.s + soff + e);
        v = c2(sv.x * rv.x, sv.y * rv.y);
      } else {
        v = rv;
      }
    } else {
      const double2 uu = *reinterpret_cast<const double2*>(A.u + off + e);
      const double2 rr = *reinterpret_cast<const double2*>(A.r + off + e);
      const double2 pp = *reinterpret_cast<const double2*>(A.p + off + e);
      const double2 qq = *reinterpret_cast<const double2*>(A.q + off + e);
      const double2 un = c2(fma(alpha, pp.x, uu.x), fma(alpha, pp.y, uu.y));
      const double2 rn = c2(fma(-alpha, qq.x, rr.x), fma(-alpha, qq.y, rr.y));
      *reinterpret_cast<double2*>(A.u + off + e) = un;
      *reinterpret_cast<double2*>(A.r + off + e) = rn;
      sq = fma(rn.x, rn.x, fma(rn.y, rn.y, sq));
      sm += rn.x + rn.y;
      if (A.s) {
        const double2 sv = *reinterpret_cast<const double2*>(A.s + soff + e);
        v = c2(sv.x * rn.x, sv.y * rn.y);
      } else {
        v = rn;
      }
    }
    const int x = e % N, y = e / N;
    plane[y * C::RS + x / 2] = v;
  }
  if (MODE != kFxyScale) {
    double t[2] = {sq, sm};
    block_sums<2>(t, red);
    if (threadIdx.x == 0) {
      double* p = A.pr + ((int64_t(l) * N + z) * 3 + c) * 2;
      p[0] = t[0];
      p[1] = t[1];
    }
  }
  __syncthreads();
  sfft<false, C::M, typename C::Rows, T>(plane, A.twx, 2);
  // split the packed spectrum of every row into its half spectrum
  constexpr int HALF = C::M / 2, PAIRS = N * (HALF + 1);
  for (int b = threadIdx.x; b < PAIRS; b += T) {
    const int row = b / (HALF + 1), k = b % (HALF + 1);
    double2* L = plane + row * C::RS;
    const double2 zk = L[k % C::M], zm = L[(C::M - k) % C::M];
    const double2 E = c2(0.5 * (zk.x + zm.x), 0.5 * (zk.y - zm.y));
    const double2 O = c2(0.5 * (zk.y + zm.y), -0.5 * (zk.x - zm.x));
    const double2 wo = cmul(A.twx[k], O);
    L[k] = cadd(E, wo);
    if (C::M - k != k) L[C::M - k] = conjd(csub(E, wo));
  }
  __syncthreads();
  sfft<false, N, typename C::Cols, T>(plane, A.twy, 1);
  double2* out = A.spec + ((int64_t(l) * 3 + c) * N + z) * int64_t(N) * C::NKX;
  for (int e = threadIdx.x; e < N * C::NKX; e += T) out[e] = plane[e];
}

// ---------------------------------------------------------------- z + Green
template <int N, int MODE>
__global__ __launch_bounds__(Cube<N>::TZ) void fz_static(FFTArgs A) {
  using C = Cube<N>;
  constexpr int T = C::TZ, KC = C::KC, NKX = C::NKX;
  extern __shared__ double2 cols[];
  double* red = reinterpret_cast<double*>(cols + 3 * N * KC);
  // ky is the fastest grid index: the kx chunks of one ky row (which share
  // cache lines) get block ids N apart, a multiple of 8 for every supported N,
  // and so run on the same XCD / L2.
  const int y = blockIdx.x, kb = blockIdx.y, l = blockIdx.z;
  const int kx0 = kb * KC;
  const int blk = y * C::NKB + kb;
  const bool lead = threadIdx.x == 0 && kb == 0 && y == 0;
  const int64_t NN = int64_t(N) * N * N;
  if (MODE == kFzInit || MODE == kFzPCG) {
    const int lane = threadIdx.x & 63;
    double v = 0.0;
    for (int i = lane; i < N; i += 64) {
      const double* p = A.pr + (int64_t(l) * N + i) * 6;
      v += (p[0] + p[2]) + p[4];
    }
    const double res2 = wave_sum(v);
    bool act;
    if (MODE == kFzInit) {
      act = res2 > A.st.target2[l] && A.st.it[l] < kCap && !A.st.bad[l];
      if (lead) {
        A.st.active_next[l] = act;
        A.st.it_next[l] = A.st.it[l];
        A.st.bad_next[l] = A.st.bad[l];
        A.st.first[l] = 1;
      }
    } else {
      const bool active = A.st.active[l] != 0;
      const double pq = wave_reduce_partials(A.ppq + int64_t(l) * A.ppq_stride, A.ppq_count, 1);
      const double rz = A.st.rz[l];
      const bool valid = isfinite(pq) && pq > 0.0 && isfinite(rz) && rz > 0.0;
      const bool finite = isfinite(res2);
      const int64_t it = A.st.it[l] + (active ? 1 : 0);
      act = active && valid && finite && res2 > A.st.target2[l] && it < kCap;
      if (lead) {
        A.st.active_next[l] = act;
        A.st.it_next[l] = it;
        A.st.bad_next[l] = A.st.bad[l] || (active && !valid) || (active && !finite);
      }
    }
    if (!act) return;
  }
  double mean[3] = {0.0, 0.0, 0.0};
  if (MODE != kFzScale && A.shat) {
    for (int c = 0; c < 3; ++c)
      mean[c] = wave_reduce_partials(A.pr + (int64_t(l) * N * 3 + c) * 2 + 1, N, 6) / double(NN);
  }
  double2* S = A.spec + int64_t(l) * 3 * C::NF;
  for (int e = threadIdx.x; e < 3 * N * KC; e += T) {
    const int kcol = e % KC, cz = e / KC;  // cz = c * N + zz
    cols[e] = S[(int64_t(cz) * N + y) * NKX + kx0 + kcol];
  }
  __syncthreads();
  sfft<false, N, typename C::ZCols, T>(cols, A.twz, 1);
  if (MODE == kFzScale) {
    for (int e = threadIdx.x; e < 3 * N * KC; e += T) {
      const int kcol = e % KC, cz = e / KC;
      S[(int64_t(cz) * N + y) * NKX + kx0 + kcol] = cols[e];
    }
    return;
  }
  double rz = 0.0, sw[3] = {0.0, 0.0, 0.0};
  for (int e = threadIdx.x; e < N * KC; e += T) {
    const int kcol = e % KC, zz = e / KC;
    const int kx = kx0 + kcol;
    const int64_t fidx = (int64_t(zz) * N + y) * NKX + kx;
    double2 v[3], sh[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) v[c] = cols[(c * N + zz) * KC + kcol];
    if (A.shat) {
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        sh[c] = A.shat[(int64_t(design_of(l, A.group)) * 3 + c) * C::NF + fidx];
        v[c] = c2(fma(-mean[c], sh[c].x, v[c].x), fma(-mean[c], sh[c].y, v[c].y));
      }
    }
    const double g0 = A.sym[fidx], g1 = A.sym[C::NF + fidx], g2 = A.sym[2 * C::NF + fidx];
    const double g3 = A.sym[3 * C::NF + fidx], g4 = A.sym[4 * C::NF + fidx], g5 = A.sym[5 * C::NF + fidx];
    double2 w[3];
    w[0] = c2(g0 * v[0].x + g3 * v[1].x + g4 * v[2].x, g0 * v[0].y + g3 * v[1].y + g4 * v[2].y);
    w[1] = c2(g3 * v[0].x + g1 * v[1].x + g5 * v[2].x, g3 * v[0].y + g1 * v[1].y + g5 * v[2].y);
    w[2] = c2(g4 * v[0].x + g5 * v[1].x + g2 * v[2].x, g4 * v[0].y + g5 * v[1].y + g2 * v[2].y);
    const double wgt = (kx == 0 || kx == N / 2) ? 1.0 : 2.0;
    double t = 0.0;
#pragma unroll
    for (int c = 0; c < 3; ++c) t += v[c].x * w[c].x + v[c].y * w[c].y;
    rz = fma(wgt, t, rz);
    if (A.shat) {
#pragma unroll
      for (int c = 0; c < 3; ++c) sw[c] = fma(wgt, sh[c].x * w[c].x + sh[c].y * w[c].y, sw[c]);
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) cols[(c * N + zz) * KC + kcol] = w[c];
  }
  __syncthreads();
  sfft<true, N, typename C::ZCols, T>(cols, A.twz, 1);
  for (int e = threadIdx.x; e < 3 * N * KC; e += T) {
    const int kcol = e % KC, cz = e / KC;
    S[(int64_t(cz) * N + y) * NKX + kx0 + kcol] = cols[e];
  }
  double t[4] = {rz, sw[0], sw[1], sw[2]};
  block_sums<4>(t, red);
  if (threadIdx.x == 0) {
    double* pg = A.pg + (int64_t(l) * C::NKB * N + blk) * 4;
    pg[0] = t[0];
    pg[1] = t[1];
    pg[2] = t[2];
    pg[3] = t[3];
  }
}

// --------------------------------------------------------------- inverse xy
template <int N>
__global__ __launch_bounds__(Cube<N>::TXY) void fxyi_static(FFTArgs A) {
  using C = Cube<N>;
  constexpr int T = C::TXY;
  extern __shared__ double2 plane[];
  const int z = blockIdx.x, c = blockIdx.y, l = blockIdx.z;
  const int64_t NN = int64_t(N) * N * N;
  if (threadIdx.x == 0 && z == 0 && c == 0) {
    A.st.active[l] = A.st.active_next[l];
    A.st.it[l] = A.st.it_next[l];
    A.st.bad[l] = A.st.bad_next[l];
  }
  if (!A.st.active_next[l]) return;
  double mean = 0.0;
  if (A.shat) mean = wave_reduce_partials(A.pg + int64_t(l) * C::NKB * N * 4 + 1 + c, C::NKB * N, 4) / double(NN);
  const double2* in = A.spec + ((int64_t(l) * 3 + c) * N + z) * int64_t(N) * C::NKX;
  for (int e = threadIdx.x; e < N * C::NKX; e += T) plane[e] = in[e];
  __syncthreads();
  sfft<true, N, typename C::Cols, T>(plane, A.twy, 1);
  constexpr int HALF = C::M / 2, PAIRS = N * (HALF + 1);
  for (int b = threadIdx.x; b < PAIRS; b += T) {
    const int row = b / (HALF + 1), k = b % (HALF + 1);
    double2* L = plane + row * C::RS;
    double2 xk = L[k], xm = L[C::M - k];
    if (k == 0) {
      xk.y = 0.0;
      xm.y = 0.0;
    }
    const double2 wk = conjd(A.twx[k]);
    const double2 e1 = c2(xk.x + xm.x, xk.y - xm.y);
    const double2 o1 = cmul(c2(xk.x - xm.x, xk.y + xm.y), wk);
    const double2 zk = c2(e1.x - o1.y, e1.y + o1.x);
    if (k != 0 && C::M - k != k) {
      const double2 wm = conjd(A.twx[C::M - k]);
      const double2 e2 = c2(xm.x + xk.x, xm.y - xk.y);
      const double2 o2 = cmul(c2(xm.x - xk.x, xm.y + xk.y), wm);
      L[C::M - k] = c2(e2.x - o2.y, e2.y + o2.x);
    }
    L[k] = zk;
  }
  __syncthreads();
  sfft<true, C::M, typename C::Rows, T>(plane, A.twx, 2);
  const int64_t off = (int64_t(l) * 3 + c) * NN + int64_t(z) * C::PLANE;
  const int64_t soff = (int64_t(design_of(l, A.group)) * 3 + c) * NN + int64_t(z) * C::PLANE;
  for (int e = 2 * threadIdx.x; e < C::PLANE; e += 2 * T) {
    const int x = e % N, y = e / N;
    const double2 w = plane[y * C::RS + x / 2];
    double2 v = w;
    if (A.s) {
      const double2 sv = *reinterpret_cast<const double2*>(A.s + soff + e);
      v = c2(sv.x * w.x - mean, sv.y * w.y - mean);
    }
    *reinterpret_cast<double2*>(A.zout + off + e) = v;
  }
}

// ------------------------------------------------------------------ dispatch
template <typename K>
int allow_lds_static(K kernel, size_t bytes) {
  if (bytes <= 65536) return PGJ_OK;
  hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(kernel),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, int(bytes));
  if (e != hipSuccess) {
    set_error("cannot reserve %zu bytes of LDS: %s", bytes, hipGetErrorString(e));
    return PGJ_ERR_HIP;
  }
  return PGJ_OK;
}

template <int N>
int fxy_static_n(FFTArgs& A, int L, int mode, hipStream_t s) {
  using C = Cube<N>;
  dim3 grid(N, 3, L), block(C::TXY);
  switch (mode) {
    case kFxyScale:
      if (int e = allow_lds_static(fxy_static<N, kFxyScale>, C::LDS_XY)) return e;
      hipLaunchKernelGGL((fxy_static<N, kFxyScale>), grid, block, C::LDS_XY, s, A);
      break;
    case kFxyPre:
      if (int e = allow_lds_static(fxy_static<N, kFxyPre>, C::LDS_XY)) return e;
      hipLaunchKernelGGL((fxy_static<N, kFxyPre>), grid, block, C::LDS_XY, s, A);
      break;
    default:
      if (int e = allow_lds_static(fxy_static<N, kFxyPCG>, C::LDS_XY)) return e;
      hipLaunchKernelGGL((fxy_static<N, kFxyPCG>), grid, block, C::LDS_XY, s, A);
  }
  PGJ_LAUNCHED("fxy_static");
  return PGJ_OK;
}

template <int N>
int fz_static_n(FFTArgs& A, int L, int mode, hipStream_t s) {
  using C = Cube<N>;
  dim3 grid(N, C::NKB, L), block(C::TZ);
  switch (mode) {
    case kFzScale: hipLaunchKernelGGL((fz_static<N, kFzScale>), grid, block, C::LDS_Z, s, A); break;
    case kFzPre: hipLaunchKernelGGL((fz_static<N, kFzPre>), grid, block, C::LDS_Z, s, A); break;
    case kFzInit: hipLaunchKernelGGL((fz_static<N, kFzInit>), grid, block, C::LDS_Z, s, A); break;
    default: hipLaunchKernelGGL((fz_static<N, kFzPCG>), grid, block, C::LDS_Z, s, A);
  }
  PGJ_LAUNCHED("fz_static");
  return PGJ_OK;
}

template <int N>
int fxyi_static_n(FFTArgs& A, int L, hipStream_t s) {
  using C = Cube<N>;
  if (int e = allow_lds_static(fxyi_static<N>, C::LDS_XY)) return e;
  hipLaunchKernelGGL(fxyi_static<N>, dim3(N, 3, L), dim3(C::TXY), C::LDS_XY, s, A);
  PGJ_LAUNCHED("fxyi_static");
  return PGJ_OK;
}

bool static_fft(const pgj_grid& g) { return static_cube(g); }

template <int N> struct FxyCall { static int go(FFTArgs& A, int L, int m, hipStream_t s) { return fxy_static_n<N>(A, L, m, s); } };
template <int N> struct FzCall { static int go(FFTArgs& A, int L, int m, hipStream_t s) { return fz_static_n<N>(A, L, m, s); } };
template <int N> struct FxyiCall { static int go(FFTArgs& A, int L, hipStream_t s) { return fxyi_static_n<N>(A, L, s); } };

int launch_fxy_static(FFTArgs& A, int L, int mode, hipStream_t s) {
#define CALL(n) FxyCall<n>::go(A, L, mode, s)
  switch (A.nx) {
    case 16: return CALL(16);
    case 24: return CALL(24);
    case 32: return CALL(32);
    case 48: return CALL(48);
    case 64: return CALL(64);
    case 96: return CALL(96);
    case 128: return CALL(128);
    default: set_error("no static FFT for n=%d", A.nx); return PGJ_ERR_UNSUPPORTED;
  }
#undef CALL
}

int launch_fz_static(FFTArgs& A, int L, int mode, hipStream_t s) {
#define CALL(n) FzCall<n>::go(A, L, mode, s)
  switch (A.nx) {
    case 16: return CALL(16);
    case 24: return CALL(24);
    case 32: return CALL(32);
    case 48: return CALL(48);
    case 64: return CALL(64);
    case 96: return CALL(96);
    case 128: return CALL(128);
    default: set_error("no static FFT for n=%d", A.nx); return PGJ_ERR_UNSUPPORTED;
  }
#undef CALL
}

int launch_fxyi_static(FFTArgs& A, int L, hipStream_t s) {
#define CALL(n) FxyiCall<n>::go(A, L, s)
  switch (A.nx) {
    case 16: return CALL(16);
    case 24: return CALL(24);
    case 32: return CALL(32);
    case 48: return CALL(48);
    case 64: return CALL(64);
    case 96: return CALL(96);
    case 128: return CALL(128);
    default: set_error("no static FFT for n=%d", A.nx); return PGJ_ERR_UNSUPPORTED;
  }
#undef CALL
}

}  // namespace pgj
