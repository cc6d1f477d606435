// FP32 HEX8 apply as a dense 24x24 Ke contraction on the matrix cores.
//
// For every element e and load l the element forces are F = a_e Ke u_e, with
// u_e the 24 nodal DOFs (DOF k = 3 o + c, corner o = x + 2 y + 4 z bits) and
// Ke the element stiffness of the trilinear hexahedron (2x2x2 Gauss), shared
// by all elements. Sixteen consecutive elements of one element row form the
// M side of v_mfma_f32_16x16x4_f32: A (16 x 4) holds four DOFs of the sixteen
// elements, B (4 x 16) the matching four rows of Ke, and six K steps over the
// 24 DOFs give the forces on output DOFs 0..15; a second accumulator takes
// DOFs 16..23 (its other eight columns are zero). The forces are scaled by
// a_e, written to LDS DOF-major and summed into the nodes.
//
// A block owns TY node rows of all nx nodes of one load and marches ZC element
// layers in z; it recomputes the element row below its first node row and the
// element layer below its first plane (the halo), so every node it owns is
// summed complete in registers and written once. Node planes stream through a
// three-slot LDS ring, the next plane loaded into registers while the current
// layer is on the matrix cores.
#include <cmath>

#include "hex8_kernels.h"

namespace pgj {
namespace {

typedef float v4f __attribute__((ext_vector_type(4)));

constexpr int kMfmaTY = 4;

struct MfmaArgs {
  int nz, ny, zc;
  int64_t N;
  const float* a;
  const float* u;
  float* out;
  const uint8_t* mask;
  float ke[576];  // Ke[k][j]
};

template <int NX>
struct MfmaTile {
  static constexpr int TY = kMfmaTY;
  static constexpr int RS = NX + 1;             // node row stride in LDS; column NX repeats x = 0
  static constexpr int PROWS = TY + 2;          // node rows y0 - 1 .. y0 + TY
  static constexpr int PLANE = 3 * PROWS * RS;  // one node plane, three components
  static constexpr int EROWS = TY + 1;          // element rows y0 - 1 .. y0 + TY - 1
  static constexpr int AMAT = EROWS * NX;
  static constexpr int DS = EROWS * NX + 1;     // stride between force DOFs in LDS
  static constexpr int XT = NX / 64;            // 16-element x tiles per wave and element row
  static constexpr int PLOADS = (PLANE + 255) / 256;
  static constexpr int ALOADS = (AMAT + 255) / 256;
  static constexpr int NODES = TY * NX / 256;   // owned nodes per thread and plane
  static constexpr int LDS_FLOATS = 3 * PLANE + 2 * AMAT + 24 * DS;
  static_assert(NX % 64 == 0 && (TY * NX) % 256 == 0, "tile shape");
};

template <bool SOA, int NX>
__global__ __launch_bounds__(256) void hex8_mfma(MfmaArgs A) {
  using P = MfmaTile<NX>;
  extern __shared__ __align__(16) unsigned char smem[];
  float* planes = reinterpret_cast<float*>(smem);
  float* amat = planes + 3 * P::PLANE;
  float* F = amat + 2 * P::AMAT;
  const int tid = int(threadIdx.x), lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int l = blockIdx.z;
  const int ny = A.ny, nz = A.nz, zc = A.zc;
  const int64_t N = A.N;
  const int y0 = int(blockIdx.x) * P::TY, z0 = int(blockIdx.y) * zc;
  const float* __restrict__ in = A.u + int64_t(l) * 3 * N;
  float* __restrict__ out = A.out + int64_t(l) * 3 * N;
  auto at = [&](int64_t node, int c) { return SOA ? int64_t(c) * N + node : 3 * node + c; };

  if (A.mask && !A.mask[l]) {
    for (int zi = 0; zi < zc; ++zi)
      for (int e = tid; e < 3 * P::TY * NX; e += 256) {
        const int c = e % 3, x = (e / 3) % NX, r = e / (3 * NX);
        out[at((int64_t(z0 + zi) * ny + y0 + r) * NX + x, c)] = 0.f;
      }
    return;
  }

  // LDS slot of plane element e: component c, node row r (y = y0 - 1 + r), column x
  auto plane_coords = [](int e, int& c, int& r, int& x) {
    if (SOA) {
      c = e / (P::PROWS * P::RS);
      r = (e / P::RS) % P::PROWS;
      x = e % P::RS;
    } else {
      r = e / (3 * P::RS);
      x = (e / 3) % P::RS;
      c = e % 3;
    }
  };
  auto plane_load = [&](int zp, float (&v)[P::PLOADS]) {
#pragma unroll
    for (int k = 0; k < P::PLOADS; ++k) {
      const int e = tid + k * 256;
      if (e < P::PLANE) {
        int c, r, x;
        plane_coords(e, c, r, x);
        const int64_t node = (int64_t(zp) * ny + wrapi(y0 - 1 + r, ny)) * NX + (x == NX ? 0 : x);
        v[k] = in[at(node, c)];
      }
    }
  };
  auto plane_store = [&](int slot, const float (&v)[P::PLOADS]) {
    float* dst = planes + slot * P::PLANE;
#pragma unroll
    for (int k = 0; k < P::PLOADS; ++k) {
      const int e = tid + k * 256;
      if (e < P::PLANE) {
        int c, r, x;
        plane_coords(e, c, r, x);
        dst[(c * P::PROWS + r) * P::RS + x] = v[k];
      }
    }
  };
  auto a_load = [&](int ze, float (&v)[P::ALOADS]) {
#pragma unroll
    for (int k = 0; k < P::ALOADS; ++k) {
      const int e = tid + k * 256;
      if (e < P::AMAT) v[k] = A.a[(int64_t(ze) * ny + wrapi(y0 - 1 + e / NX, ny)) * NX + e % NX];
    }
  };
  auto a_store = [&](int slot, const float (&v)[P::ALOADS]) {
#pragma unroll
    for (int k = 0; k < P::ALOADS; ++k) {
      const int e = tid + k * 256;
      if (e < P::AMAT) amat[slot * P::AMAT + e] = v[k];
    }
  };

  // Per-lane operands: A[i][k] = u_{x0+i}[4 s + k] and B[k][j] = Ke[4 s + k][j] with i, j = lane % 16 and
  // k = lane / 16; the accumulator holds D[4 (lane / 16) + r][lane % 16] in its component r.
  const int i16 = lane & 15, kq = lane >> 4;
  float b0[6], b1[6];
  int off[6];
  bool up[6];
#pragma unroll
  for (int s = 0; s < 6; ++s) {
    const int k = 4 * s + kq, o = k / 3, c = k % 3;
    b0[s] = A.ke[k * 24 + i16];
    b1[s] = i16 < 8 ? A.ke[k * 24 + 16 + i16] : 0.f;
    off[s] = (c * P::PROWS + ((o >> 1) & 1)) * P::RS + (o & 1) + i16;
    up[s] = (o >> 2) & 1;
  }

  float pv[P::PLOADS], av[P::ALOADS];
  plane_load(wrapi(z0 - 1, nz), pv);
  plane_store(0, pv);
  plane_load(z0, pv);
  plane_store(1, pv);
  a_load(wrapi(z0 - 1, nz), av);
  a_store(0, av);
  float carry[P::NODES][3];
#pragma unroll
  for (int m = 0; m < P::NODES; ++m)
#pragma unroll
    for (int c = 0; c < 3; ++c) carry[m][c] = 0.f;

  for (int li = 0; li <= zc; ++li) {
    const int z = wrapi(z0 - 1 + li, nz);  // element layer between node planes z and z + 1
    __syncthreads();                       // its planes and materials are in LDS, the last forces consumed
    const bool more = li < zc;
    if (more) {
      plane_load(wrapi(z + 2, nz), pv);
      a_load(wrapi(z + 1, nz), av);
    }
    const int bot = (li % 3) * P::PLANE, top = ((li + 1) % 3) * P::PLANE;
    const float* am = amat + (li & 1) * P::AMAT;
    int base[6];
#pragma unroll
    for (int s = 0; s < 6; ++s) base[s] = (up[s] ? top : bot) + off[s];
#pragma unroll
    for (int t = 0; t < P::XT; ++t) {
      const int x0 = (wave + 4 * t) * 16;
      v4f acc0[P::EROWS], acc1[P::EROWS];
#pragma unroll
      for (int j = 0; j < P::EROWS; ++j) {
        acc0[j] = v4f{0.f, 0.f, 0.f, 0.f};
        acc1[j] = v4f{0.f, 0.f, 0.f, 0.f};
      }
#pragma unroll
      for (int s = 0; s < 6; ++s)
#pragma unroll
        for (int j = 0; j < P::EROWS; ++j) {
          const float v = planes[base[s] + j * P::RS + x0];
          acc0[j] = __builtin_amdgcn_mfma_f32_16x16x4f32(v, b0[s], acc0[j], 0, 0, 0);
          acc1[j] = __builtin_amdgcn_mfma_f32_16x16x4f32(v, b1[s], acc1[j], 0, 0, 0);
        }
#pragma unroll
      for (int j = 0; j < P::EROWS; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int xe = x0 + 4 * kq + r;
          const float ae = am[j * NX + xe];
          F[i16 * P::DS + j * NX + xe] = acc0[j][r] * ae;
          if (i16 < 8) F[(16 + i16) * P::DS + j * NX + xe] = acc1[j][r] * ae;
        }
    }
    if (more) {
      plane_store((li + 2) % 3, pv);
      a_store((li + 1) & 1, av);
    }
    __syncthreads();
    // node (y0 + jj, x) of plane z gets the corners oz = 0 of layer z (elements (x - ox, y - oy)) on top of the
    // corners oz = 1 of layer z - 1 carried from the previous step
#pragma unroll
    for (int m = 0; m < P::NODES; ++m) {
      const int nidx = tid + m * 256, jj = nidx / NX, x = nidx % NX;
      const int xm = x == 0 ? NX - 1 : x - 1;
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        float lo = 0.f, hi = 0.f;
#pragma unroll
        for (int o = 0; o < 4; ++o) {
          const int e = (jj + 1 - (o >> 1)) * NX + ((o & 1) ? xm : x);
          lo += F[(3 * o + c) * P::DS + e];
          hi += F[(3 * (o + 4) + c) * P::DS + e];
        }
        const float q = carry[m][c] + lo;
        carry[m][c] = hi;
        if (li >= 1) out[at((int64_t(z) * ny + y0 + jj) * NX + x, c)] = q;
      }
    }
  }
}

// Dense element stiffness Ke = sum_q B_q^T C0 B_q |J| over the 2x2x2 Gauss points (weights 1), engineering
// shear in Voigt order 11, 22, 33, 23, 13, 12; DOF k = 3 o + c.
void element_stiffness(const pgj_grid& g, double ke[576]) {
  const double h[3] = {g.hx, g.hy, g.hz};
  const double gp = 1.0 / std::sqrt(3.0), vol = g.hx * g.hy * g.hz / 8.0;
  double C[6][6] = {};
  for (int i = 0; i < 3; ++i) {
    for (int j = 0; j < 3; ++j) C[i][j] = g.lam;
    C[i][i] += 2.0 * g.mu;
    C[i + 3][i + 3] = g.mu;
  }
  for (int k = 0; k < 576; ++k) ke[k] = 0.0;
  for (int q = 0; q < 8; ++q) {
    double xi[3], B[6][24] = {};
    for (int i = 0; i < 3; ++i) xi[i] = ((q >> i) & 1) ? gp : -gp;
    for (int o = 0; o < 8; ++o) {
      double s[3], d[3];
      for (int i = 0; i < 3; ++i) s[i] = ((o >> i) & 1) ? 1.0 : -1.0;
      for (int i = 0; i < 3; ++i) {
        d[i] = s[i] / h[i];
        for (int j = 0; j < 3; ++j)
          if (j != i) d[i] *= 0.5 * (1.0 + s[j] * xi[j]);
      }
      B[0][3 * o] = d[0];
      B[1][3 * o + 1] = d[1];
      B[2][3 * o + 2] = d[2];
      B[3][3 * o + 1] = d[2];
      B[3][3 * o + 2] = d[1];
      B[4][3 * o] = d[2];
      B[4][3 * o + 2] = d[0];
      B[5][3 * o] = d[1];
      B[5][3 * o + 1] = d[0];
    }
    double CB[6][24];
    for (int i = 0; i < 6; ++i)
      for (int k = 0; k < 24; ++k) {
        double t = 0.0;
        for (int m = 0; m < 6; ++m) t += C[i][m] * B[m][k];
        CB[i][k] = t;
      }
    for (int k = 0; k < 24; ++k)
      for (int j = 0; j < 24; ++j) {
        double t = 0.0;
        for (int i = 0; i < 6; ++i) t += B[i][k] * CB[i][j];
        ke[k * 24 + j] += vol * t;
      }
  }
}

template <bool SOA, int NX>
int launch_mfma(MfmaArgs& A, int L, hipStream_t s) {
  using P = MfmaTile<NX>;
  const size_t lds = size_t(P::LDS_FLOATS) * sizeof(float);
  if (lds > 64 * 1024) {
    const hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&hex8_mfma<SOA, NX>),
                                             hipFuncAttributeMaxDynamicSharedMemorySize, int(lds));
    PGJ_REQUIRE(e == hipSuccess, "hex8_mfma: LDS attribute: %s", hipGetErrorString(e));
  }
  const dim3 grid(unsigned(A.ny / P::TY), unsigned(A.nz / A.zc), unsigned(L));
  hipLaunchKernelGGL((hex8_mfma<SOA, NX>), grid, dim3(256), lds, s, A);
  PGJ_LAUNCHED("hex8_mfma");
  return PGJ_OK;
}

}  // namespace
}  // namespace pgj

using namespace pgj;

extern "C" int pgj_hex8_apply32_mfma(const pgj_grid* g, int L, const float* a, const float* u,
                                     const uint8_t* active, float* out, int soa, void* stream) {
  if (int e = check_grid(g)) return e;
  PGJ_REQUIRE(L >= 1 && L <= 65535 && a && u && out, "invalid apply arguments");
  PGJ_REQUIRE(g->nx == 64 || g->nx == 128, "the MFMA apply needs nx = 64 or 128 (got %d)", g->nx);
  PGJ_REQUIRE(g->ny % kMfmaTY == 0, "the MFMA apply needs ny divisible by %d (got %d)", kMfmaTY, g->ny);
  MfmaArgs A{};
  A.nz = g->nz;
  A.ny = g->ny;
  A.N = int64_t(g->nx) * g->ny * g->nz;
  A.a = a;
  A.u = u;
  A.out = out;
  A.mask = active;
  // the longest z march (fewest halo layers) that still gives three blocks per CU
  int zc = 1;
  for (int c : {8, 4, 2})
    if (g->nz % c == 0) {
      zc = c;
      break;
    }
  while (zc > 1 && int64_t(g->ny / kMfmaTY) * (g->nz / zc) * L < 768) zc /= 2;
  A.zc = zc;
  double ke[576];
  element_stiffness(*g, ke);
  for (int k = 0; k < 576; ++k) A.ke[k] = float(ke[k]);
  const hipStream_t s = as_stream(stream);
  if (g->nx == 64) return soa ? launch_mfma<true, 64>(A, L, s) : launch_mfma<false, 64>(A, L, s);
  return soa ? launch_mfma<true, 128>(A, L, s) : launch_mfma<false, 128>(A, L, s);
}
