// Declarations shared by the HEX8 kernels and the resident solver.
#pragma once

#include "hex8_element.h"
#include "layout.h"
#include "pgj_internal.h"

namespace pgj {

enum { kModeApply = 0, kModePCG = 1 };

template <typename T>
struct RowArgsT {
  int nz, ny, nx;
  int64_t N;
  RowMarch cfg;
  ElementCoef coef;
  const T* a;
  const T* u;            // apply: input vectors; PCG: the direction p
  T* out;                // apply: K u; PCG: q = K p
  const uint8_t* mask;   // apply: optional active mask
  StateView st;          // PCG: active mask
  double* ppq;           // per-block p.q partials, per load stride ppq_stride
  int ppq_stride;
  int group;             // loads per design (a is (designs, N)); 0: one shared a
};
using RowArgs = RowArgsT<double>;

// Independent reference apply: one thread per node gathers the rows of its
// eight elements; optionally fused with the certification residual.
struct RefArgs {
  int nz, ny, nx;
  int64_t N;
  ElementCoef coef;
  bool cubic;
  const double* a;
  const double* u;
  double* out;
  // certification residual: out = rhs - K u, per-block |r|^2 into slot 3 of
  // 4-wide partial rows [l][block][4]
  const double* rhs;
  double* pres;
  int soa;               // vectors are SoA (L, 3, N) instead of AoS (L, N, 3)
  int group;             // loads per design (a is (designs, N)); 0: one shared a
};


bool cubic(const pgj_grid& g);
ElementCoef coef_of(const pgj_grid& g);
int check_grid(const pgj_grid* g);
template <typename T>
int launch_rowmarch(const pgj_grid& g, int L, int mode, RowArgsT<T>& A, hipStream_t s);
int launch_reference(const pgj_grid& g, int L, RefArgs& A, bool cert, hipStream_t s);
int launch_component_sums(const pgj_grid& g, int L, const double* v, double* part, int soa, hipStream_t s);
int launch_subtract_mean(const pgj_grid& g, int L, double* v, const double* part, int soa, hipStream_t s);
int launch_rhs(const pgj_grid& g, const double* a, const double* macro, double* out, int soa, hipStream_t s);
// alpha_dev (nullable): read alpha from device memory instead of `alpha`
int launch_scale(const pgj_grid& g, const double* a, double alpha, double* d, double* s_out, double* ws, int soa,
                 hipStream_t s, const double* alpha_dev = nullptr);

// General-C21 operator (hybrid.hip) in the resident solver: SoA vectors throughout.
// Blocks per load of the PCG-mode tiled apply over L loads (its p.q partial count).
int c21_pcg_blocks(const pgj_grid& g, int L, int* blocks);
// q = K p for the active loads of `active`, p.q partials per block into ppq[l * ppq_stride + block].
int launch_c21_pcg(const pgj_grid& g, int L, const double* pool, const int32_t* index, const double* p,
                   const uint8_t* active, double* q, double* ppq, int ppq_stride, hipStream_t s);
// r = rhs - K u, per-block |r|^2 into slot 3 of the 4-wide partial rows [l][block][4] (256-node blocks).
int launch_c21_certify(const pgj_grid& g, int L, const double* pool, const int32_t* index, const double* u,
                       const double* rhs, double* r, double* pres, hipStream_t s);
// rhs of one macro strain (host 6 doubles), AoS or SoA (soa != 0).
int launch_c21_rhs(const pgj_grid& g, const double* pool, const int32_t* index, const double* macro, double* out,
                   int soa, hipStream_t s);

}  // namespace pgj
