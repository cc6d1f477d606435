// Exact 2x2x2-Gauss Q1/HEX8 element action for an isotropic material a*C0(lambda, mu).
//
// Derivation (independent of any particular code):
//   Node a of the element sits at local coordinates xi_i = 2*bit_i(a) - 1. The
//   trilinear field of component d is u_d(xi) = 1/8 sum_m U_d[m] xi^m over the
//   eight monomials m in {0,1}^3, with U_d[m] = sum_a prod_{i in m}(2 a_i - 1) u_d[a]
//   (three 2-point butterflies). Its gradient is
//     du_d/dx_j = g_j sum_{m : m_j = 1} U_d[m] xi^(m - e_j),  g_j = 1/(4 h_j),
//   so every strain component is a combination of seven monomials t (t != 111).
//   Distinct monomials are orthogonal under the element integral, and
//   (1/8) int xi^(2t) = 3^-|t|, which the 2-point rule reproduces exactly. Hence
//     Ke u = a V sum_t 3^-|t| (d eps_t / du)^T C0 eps_t,
//   evaluated as: forward butterflies -> per-monomial isotropic stress
//   sigma = k_t (lambda tr(eps) I + 2 mu eps) -> modal forces
//   F_d[m] = sum_{j: m_j=1} g_j sigma_dj(m - e_j) -> transposed butterflies.
// Element node a has offset (a&1, (a>>1)&1, (a>>2)&1) in (x, y, z), dof 3a+d.
#pragma once

#ifndef PGJ_HD
#if defined(__HIPCC__)
#define PGJ_HD __host__ __device__
#else
#define PGJ_HD
#endif
#endif

namespace pgj {

struct ElementCoef {
  double g[3];      // 1/(4 h_j)
  double lam, mu;   // isotropic C0
  double vol;       // hx*hy*hz
};

// In-plane forward butterflies of one component on a 2x2 node block
// (n = x + 2y): x then y, (lo, hi) -> (lo + hi, hi - lo). P[m] for m = mx + 2 my.
template <typename T>
PGJ_HD inline void plane_forward(T n0, T n1, T n2, T n3, T (&P)[4]) {
  const T a0 = n0 + n1, a1 = n1 - n0, b0 = n2 + n3, b1 = n3 - n2;
  P[0] = a0 + b0;
  P[2] = b0 - a0;
  P[1] = a1 + b1;
  P[3] = b1 - a1;
}

// In-plane transposed butterflies of one component: y then x, (lo, hi) -> (lo - hi, lo + hi).
template <typename T>
PGJ_HD inline void plane_transposed(const T (&M)[4], T (&f)[4]) {
  const T t0 = M[0] - M[2], t2 = M[0] + M[2], t1 = M[1] - M[3], t3 = M[1] + M[3];
  f[0] = t0 - t1;
  f[1] = t0 + t1;
  f[2] = t2 - t3;
  f[3] = t2 + t3;
}

// Modal core: forward-transformed nodal values U[d][m] (m = mx + 2 my + 4 mz)
// -> modal forces F[d][m] of a * Ke(C0), before the transposed butterflies.
template <bool kCubic, typename T = double>
PGJ_HD inline void hex8_modal_core(const T (&U)[3][8], T a, const ElementCoef& c, T (&F)[3][8]) {
  // Gradient scale: with a cubic voxel g^2 is folded into k_t.
  const T gs0 = kCubic ? T(1) : T(c.g[0]);
  const T gs1 = kCubic ? T(1) : T(c.g[1]);
  const T gs2 = kCubic ? T(1) : T(c.g[2]);
  const T gs[3] = {gs0, gs1, gs2};
  const T base = kCubic ? T(double(a) * c.vol * c.g[0] * c.g[0]) : T(double(a) * c.vol);
#pragma unroll
  for (int d = 0; d < 3; ++d)
#pragma unroll
    for (int n = 0; n < 8; ++n) F[d][n] = T(0);
#pragma unroll
  for (int t = 0; t < 7; ++t) {
    const int bits = ((t & 1) ? 1 : 0) + ((t & 2) ? 1 : 0) + ((t & 4) ? 1 : 0);
    const T w = bits == 0 ? T(1) : (bits == 1 ? T(1.0 / 3.0) : T(1.0 / 9.0));
    const T k = base * w;
    // Gradient entries G[d][j] = g_j U_d[t | e_j], present iff bit j of t is clear.
    T G[3][3];
#pragma unroll
    for (int d = 0; d < 3; ++d)
#pragma unroll
      for (int j = 0; j < 3; ++j) G[d][j] = (t & (1 << j)) ? T(0) : gs[j] * U[d][t | (1 << j)];
    T tr = T(0);
#pragma unroll
    for (int j = 0; j < 3; ++j)
      if (!(t & (1 << j))) tr += G[j][j];
    const T ltr = T(c.lam) * k * tr;
    const T two_mu_k = T(2.0 * c.mu) * k;
    const T mu_k = T(c.mu) * k;
#pragma unroll
    for (int j = 0; j < 3; ++j)
      if (!(t & (1 << j))) F[j][t | (1 << j)] += gs[j] * (ltr + two_mu_k * G[j][j]);
#pragma unroll
    for (int d = 0; d < 3; ++d)
#pragma unroll
      for (int j = d + 1; j < 3; ++j) {
        const bool jd = !(t & (1 << j));  // G[d][j] present
        const bool dj = !(t & (1 << d));  // G[j][d] present
        if (jd || dj) {
          const T s = mu_k * ((jd ? G[d][j] : T(0)) + (dj ? G[j][d] : T(0)));
          if (jd) F[d][t | (1 << j)] += gs[j] * s;
          if (dj) F[j][t | (1 << d)] += gs[d] * s;
        }
      }
  }
}

// f = a * Ke(C0) * u for one element. u and f are 24 values, dof 3a+d; T is
// the arithmetic precision (the coefficients are rounded to it).
template <bool kCubic, typename T = double>
PGJ_HD inline void hex8_element(const T (&u)[24], T a, const ElementCoef& c, T (&f)[24]) {
  T U[3][8];
#pragma unroll
  for (int d = 0; d < 3; ++d)
#pragma unroll
    for (int n = 0; n < 8; ++n) U[d][n] = u[3 * n + d];
  // Forward butterflies: (lo, hi) -> (lo + hi, hi - lo) along x, y, z.
#pragma unroll
  for (int b = 1; b < 8; b <<= 1)
#pragma unroll
    for (int d = 0; d < 3; ++d)
#pragma unroll
      for (int n = 0; n < 8; ++n)
        if (!(n & b)) {
          const T lo = U[d][n], hi = U[d][n | b];
          U[d][n] = lo + hi;
          U[d][n | b] = hi - lo;
        }
  T F[3][8];
  hex8_modal_core<kCubic, T>(U, a, c, F);
  // Transposed butterflies: (lo, hi) -> (lo - hi, lo + hi) along z, y, x.
#pragma unroll
  for (int b = 4; b >= 1; b >>= 1)
#pragma unroll
    for (int d = 0; d < 3; ++d)
#pragma unroll
      for (int n = 0; n < 8; ++n)
        if (!(n & b)) {
          const T lo = F[d][n], hi = F[d][n | b];
          F[d][n] = lo - hi;
          F[d][n | b] = lo + hi;
        }
#pragma unroll
  for (int n = 0; n < 8; ++n)
#pragma unroll
    for (int d = 0; d < 3; ++d) f[3 * n + d] = F[d][n];
}

// General anisotropic material: the element constant C (6x6 Voigt matrix,
// order 11, 22, 33, 23, 13, 12, engineering shear) in place of a C0. The same
// monomial decomposition holds for any constant C; the stress of strain
// monomial t is k_t C eps_t with eps_t = (G00, G11, G22, G12 + G21, G02 + G20,
// G01 + G10), and the modal force of component d along x_j is g_j sigma_dj.
// c.g and c.vol give the voxel (c.lam, c.mu are not used).
template <typename T = double>
PGJ_HD inline void hex8_modal_core_c(const T (&U)[3][8], const T (&C)[6][6], const ElementCoef& c, T (&F)[3][8]) {
  // tensor index (d, j) -> Voigt row
  constexpr int kVoigt[3][3] = {{0, 5, 4}, {5, 1, 3}, {4, 3, 2}};
#pragma unroll
  for (int d = 0; d < 3; ++d)
#pragma unroll
    for (int n = 0; n < 8; ++n) F[d][n] = T(0);
#pragma unroll
  for (int t = 0; t < 7; ++t) {
    const int bits = ((t & 1) ? 1 : 0) + ((t & 2) ? 1 : 0) + ((t & 4) ? 1 : 0);
    const T k = T(c.vol) * (bits == 0 ? T(1) : (bits == 1 ? T(1.0 / 3.0) : T(1.0 / 9.0)));
    T G[3][3];
#pragma unroll
    for (int d = 0; d < 3; ++d)
#pragma unroll
      for (int j = 0; j < 3; ++j) G[d][j] = (t & (1 << j)) ? T(0) : T(c.g[j]) * U[d][t | (1 << j)];
    const T eps[6] = {G[0][0], G[1][1], G[2][2], G[1][2] + G[2][1], G[0][2] + G[2][0], G[0][1] + G[1][0]};
    T sig[6];
#pragma unroll
    for (int r = 0; r < 6; ++r) {
      T acc = T(0);
#pragma unroll
      for (int q = 0; q < 6; ++q) acc += C[r][q] * eps[q];
      sig[r] = k * acc;
    }
#pragma unroll
    for (int d = 0; d < 3; ++d)
#pragma unroll
      for (int j = 0; j < 3; ++j)
        if (!(t & (1 << j))) F[d][t | (1 << j)] += T(c.g[j]) * sig[kVoigt[d][j]];
  }
}

// f = Ke(C) u for one element (dof 3a+d), C as in hex8_modal_core_c.
template <typename T = double>
PGJ_HD inline void hex8_element_c(const T (&u)[24], const T (&C)[6][6], const ElementCoef& c, T (&f)[24]) {
  T U[3][8];
#pragma unroll
  for (int d = 0; d < 3; ++d)
#pragma unroll
    for (int n = 0; n < 8; ++n) U[d][n] = u[3 * n + d];
#pragma unroll
  for (int b = 1; b < 8; b <<= 1)
#pragma unroll
    for (int d = 0; d < 3; ++d)
#pragma unroll
      for (int n = 0; n < 8; ++n)
        if (!(n & b)) {
          const T lo = U[d][n], hi = U[d][n | b];
          U[d][n] = lo + hi;
          U[d][n | b] = hi - lo;
        }
  T F[3][8];
  hex8_modal_core_c<T>(U, C, c, F);
#pragma unroll
  for (int b = 4; b >= 1; b >>= 1)
#pragma unroll
    for (int d = 0; d < 3; ++d)
#pragma unroll
      for (int n = 0; n < 8; ++n)
        if (!(n & b)) {
          const T lo = F[d][n], hi = F[d][n | b];
          F[d][n] = lo - hi;
          F[d][n | b] = lo + hi;
        }
#pragma unroll
  for (int n = 0; n < 8; ++n)
#pragma unroll
    for (int d = 0; d < 3; ++d) f[3 * n + d] = F[d][n];
}

// Unpack a 21-entry upper triangle (rows of the 6x6 Voigt matrix: C00..C05,
// C11..C15, C22..C25, C33..C35, C44, C45, C55) into the full symmetric matrix.
template <typename T = double>
PGJ_HD inline void unpack_c21(const double* p, T (&C)[6][6]) {
  int k = 0;
#pragma unroll
  for (int r = 0; r < 6; ++r)
#pragma unroll
    for (int q = r; q < 6; ++q) {
      C[r][q] = C[q][r] = T(p[k]);
      ++k;
    }
}

inline ElementCoef make_coef(double hx, double hy, double hz, double lam, double mu) {
  ElementCoef c;
  c.g[0] = 1.0 / (4.0 * hx);
  c.g[1] = 1.0 / (4.0 * hy);
  c.g[2] = 1.0 / (4.0 * hz);
  c.lam = lam;
  c.mu = mu;
  c.vol = hx * hy * hz;
  return c;
}

}  // namespace pgj
