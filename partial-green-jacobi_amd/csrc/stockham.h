// Compile-time LDS Stockham machinery: radix plan, small DFTs, batched
// in-place stages over lines placed in LDS. Used by the FFT kernels whose
// grid size is a template parameter, so every index is a constant expression.
#pragma once

#include "pgj_internal.h"

namespace pgj {

// Complex helpers for double2 (FP64) and float2 (FP32) values (cplx<T> in pgj_internal.h).
__device__ __forceinline__ double2 c2(double x, double y) { return make_double2(x, y); }
__device__ __forceinline__ float2 c2(float x, float y) { return make_float2(x, y); }
template <typename C> __device__ __forceinline__ C cadd(C a, C b) { return c2(a.x + b.x, a.y + b.y); }
template <typename C> __device__ __forceinline__ C csub(C a, C b) { return c2(a.x - b.x, a.y - b.y); }
template <typename C> __device__ __forceinline__ C cmul(C a, C b) {
  return c2(fma(a.x, b.x, -a.y * b.y), fma(a.x, b.y, a.y * b.x));
}
template <typename C> __device__ __forceinline__ C conjd(C a) { return c2(a.x, -a.y); }

// W_n^k from a table of length T = n * step (W_T[j] = exp(-2 pi i j / T)).
template <bool INV, typename C>
__device__ __forceinline__ C twiddle(const C* tw, int k, int step) {
  const C t = tw[k * step];
  return INV ? conjd(t) : t;
}

template <int R, bool INV, typename C>
__device__ __forceinline__ void dft(C (&v)[R], const C* tw, int n, int step) {
  using T = decltype(C{}.x);
  if constexpr (R == 2) {
    const C a = v[0];
    v[0] = cadd(a, v[1]);
    v[1] = csub(a, v[1]);
  } else if constexpr (R == 4) {
    const C t0 = cadd(v[0], v[2]), t1 = csub(v[0], v[2]), t2 = cadd(v[1], v[3]);
    C t3 = csub(v[1], v[3]);
    t3 = INV ? c2(-t3.y, t3.x) : c2(t3.y, -t3.x);
    v[0] = cadd(t0, t2);
    v[2] = csub(t0, t2);
    v[1] = cadd(t1, t3);
    v[3] = csub(t1, t3);
  } else if constexpr (R == 8) {
    // split radix-8: a_j = x_j + x_{j+4}, b_j = (x_j - x_{j+4}) W_8^j (conjugate for the inverse); the even
    // outputs are the 4-point DFT of a, the odd ones that of b
    constexpr T h = T(0.70710678118654752440);  // 1/sqrt(2)
    C a[4], b[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      a[j] = cadd(v[j], v[j + 4]);
      b[j] = csub(v[j], v[j + 4]);
    }
    // W_8^1 = (1 - i)/sqrt(2), W_8^2 = -i, W_8^3 = (-1 - i)/sqrt(2)
    b[1] = INV ? c2(h * (b[1].x - b[1].y), h * (b[1].x + b[1].y)) : c2(h * (b[1].x + b[1].y), h * (b[1].y - b[1].x));
    b[2] = INV ? c2(-b[2].y, b[2].x) : c2(b[2].y, -b[2].x);
    b[3] = INV ? c2(-h * (b[3].x + b[3].y), h * (b[3].x - b[3].y))
               : c2(h * (b[3].y - b[3].x), -h * (b[3].x + b[3].y));
    dft<4, INV>(a, tw, n, step);
    dft<4, INV>(b, tw, n, step);
#pragma unroll
    for (int m = 0; m < 4; ++m) {
      v[2 * m] = a[m];
      v[2 * m + 1] = b[m];
    }
  } else if constexpr (R == 3) {
    constexpr T h = T(0.86602540378443864676);  // sqrt(3)/2
    const C s = cadd(v[1], v[2]), d = csub(v[1], v[2]);
    const C m = c2(v[0].x - T(0.5) * s.x, v[0].y - T(0.5) * s.y);
    const C jd = INV ? c2(-h * d.y, h * d.x) : c2(h * d.y, -h * d.x);
    v[0] = cadd(v[0], s);
    v[1] = cadd(m, jd);
    v[2] = csub(m, jd);
  } else {
    C out[R];
#pragma unroll
    for (int p = 0; p < R; ++p) {
      C acc = v[0];
#pragma unroll
      for (int q = 1; q < R; ++q) acc = cadd(acc, cmul(v[q], twiddle<INV>(tw, ((p * q) % R) * (n / R), step)));
      out[p] = acc;
    }
#pragma unroll
    for (int p = 0; p < R; ++p) v[p] = out[p];
  }
}

// Radix plans of a length. Plan 0: radices 4 first, then 2, 3, 5, 7 (the order of factorize()). Plan 1:
// radix 8 first, then as plan 0 (64 = 8 x 8: two LDS passes instead of three).
constexpr int radix_count(int n, int plan = 0) {
  int c = 0;
  if (plan == 1)
    while (n % 8 == 0) { n /= 8; ++c; }
  while (n % 4 == 0) { n /= 4; ++c; }
  while (n % 2 == 0) { n /= 2; ++c; }
  for (int p = 3; p <= 7; p += 2)
    while (n % p == 0) { n /= p; ++c; }
  return c;
}

constexpr int radix_at(int n, int s, int plan = 0) {
  int c = 0;
  if (plan == 1)
    while (n % 8 == 0) { if (c++ == s) return 8; n /= 8; }
  while (n % 4 == 0) { if (c++ == s) return 4; n /= 4; }
  while (n % 2 == 0) { if (c++ == s) return 2; n /= 2; }
  for (int p = 3; p <= 7; p += 2)
    while (n % p == 0) { if (c++ == s) return p; n /= p; }
  return 1;
}

constexpr int radix_ns(int n, int s, int plan = 0) {
  int prod = 1;
  for (int k = 0; k < s; ++k) prod *= radix_at(n, k, plan);
  return prod;
}

// Lines of a batched transform in LDS: line l starts at (l / PER) * OUTER +
// (l % PER) * INNER, elements ES apart; LINE_FAST maps consecutive threads to
// consecutive lines (else to consecutive butterflies of one line).
template <int NLINES, int PER, int OUTER, int INNER, int ES, bool LINE_FAST>
struct Lines {
  static constexpr int nlines = NLINES;
  static constexpr int es = ES;
  static constexpr bool line_fast = LINE_FAST;
  __device__ static __forceinline__ int base(int l) { return (l / PER) * OUTER + (l % PER) * INNER; }
};

template <int R, bool INV, int N, int NS, class LN, int THREADS, typename C>
__device__ __forceinline__ void sstage(C* buf, const C* __restrict__ tw, int step) {
  constexpr int NB = N / R, COUNT = LN::nlines * NB, PT = (COUNT + THREADS - 1) / THREADS, M = N / (NS * R);
  C v[PT][R];
#pragma unroll
  for (int i = 0; i < PT; ++i) {
    const int b = threadIdx.x + i * THREADS;
    if (COUNT % THREADS == 0 || b < COUNT) {
      const int line = LN::line_fast ? b % LN::nlines : b / NB;
      const int j = LN::line_fast ? b / LN::nlines : b % NB;
      const int k = j % NS;
      const C* L = buf + LN::base(line);
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const C x = L[(j + r * NB) * LN::es];
        v[i][r] = (NS == 1 || r == 0) ? x : cmul(x, twiddle<INV>(tw, k * r * M, step));
      }
    }
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < PT; ++i) {
    const int b = threadIdx.x + i * THREADS;
    if (COUNT % THREADS == 0 || b < COUNT) {
      const int line = LN::line_fast ? b % LN::nlines : b / NB;
      const int j = LN::line_fast ? b / LN::nlines : b % NB;
      const int k = j % NS;
      dft<R, INV>(v[i], tw, N, step);
      C* L = buf + LN::base(line);
      const int idx = (j / NS) * NS * R + k;
#pragma unroll
      for (int r = 0; r < R; ++r) L[(idx + r * NS) * LN::es] = v[i][r];
    }
  }
  __syncthreads();
}

template <bool INV, int N, class LN, int THREADS, int PLAN = 0, int S = 0, typename C>
__device__ __forceinline__ void sfft(C* buf, const C* __restrict__ tw, int step) {
  if constexpr (S < radix_count(N, PLAN)) {
    sstage<radix_at(N, S, PLAN), INV, N, radix_ns(N, S, PLAN), LN, THREADS>(buf, tw, step);
    sfft<INV, N, LN, THREADS, PLAN, S + 1>(buf, tw, step);
  }
}

}  // namespace pgj
