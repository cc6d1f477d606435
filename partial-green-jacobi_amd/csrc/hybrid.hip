// General-C21 material (the reference's hybrid operator): element e carries a
// constant anisotropic C, the 21-entry upper triangle of its 6x6 Voigt matrix,
// either as row e of a coefficient array (direct) or as row index[e] of a
// shared pool (indirect: phases share one row, multi-phase composites). The
// apply gathers per node over its eight elements (fixed order, no atomics, so
// results are independent of the launch); RHS and mean stress use the exact
// centroid gradient of the trilinear field.
#include <algorithm>
#include <climits>

#include "hex8_element.h"
#include "hex8_kernels.h"
#include "layout.h"
#include "pgj_internal.h"

namespace pgj {

// One Voigt macro strain (engineering shears), passed by value.
struct Strain6 {
  double v[6];
};

__device__ __forceinline__ const double* c21_row(const double* pool, const int32_t* index, int64_t e) {
  return pool + (index ? int64_t(index[e]) : e) * 21;
}

// q_l = sum_e P_e^T Ke(C_e) P_e u_l; AoS (L, 3N), or SoA (L, 3, N) with SOA; inactive loads give zeros.
// CERT: out = rhs - K u and per-block |out|^2 into slot 3 of the 4-wide partial rows [l][block][4].
template <bool SOA, bool CERT>
__global__ __launch_bounds__(256) void c21_apply_kernel(int nz, int ny, int nx, ElementCoef coef, const double* pool,
                                                        const int32_t* index, const double* u, const uint8_t* active,
                                                        double* out, const double* rhs, double* pres) {
  __shared__ double red[4];
  const int l = blockIdx.y;
  const int64_t N = int64_t(nx) * ny * nz;
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  const int64_t base = int64_t(l) * 3 * N;
  auto at = [&](int64_t node, int d) { return SOA ? base + int64_t(d) * N + node : base + 3 * node + d; };
  double res2 = 0.0;
  if (i < N) {
  double acc[3] = {0.0, 0.0, 0.0};
  if (!active || active[l]) {
    const int x = int(i % nx), y = int((i / nx) % ny), z = int(i / (int64_t(nx) * ny));
#pragma unroll 1
    for (int o = 0; o < 8; ++o) {
      const int ex = wrapi(x - (o & 1), nx), ey = wrapi(y - ((o >> 1) & 1), ny), ez = wrapi(z - ((o >> 2) & 1), nz);
      const int64_t e = (int64_t(ez) * ny + ey) * nx + ex;
      double ue[24], fe[24], C[6][6];
#pragma unroll
      for (int n = 0; n < 8; ++n) {
        const int64_t node = (int64_t(wrapi(ez + ((n >> 2) & 1), nz)) * ny + wrapi(ey + ((n >> 1) & 1), ny)) * nx +
                             wrapi(ex + (n & 1), nx);
#pragma unroll
        for (int d = 0; d < 3; ++d) ue[3 * n + d] = u[at(node, d)];
      }
      unpack_c21(c21_row(pool, index, e), C);
      hex8_element_c(ue, C, coef, fe);
      // this node is local node o of element e (static register indices)
#pragma unroll
      for (int k = 0; k < 8; ++k)
        if (k == o) {
#pragma unroll
          for (int d = 0; d < 3; ++d) acc[d] += fe[3 * k + d];
        }
    }
  }
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    if (CERT) {
      const double r = rhs[at(i, d)] - acc[d];
      out[at(i, d)] = r;
      res2 = fma(r, r, res2);
    } else {
      out[at(i, d)] = acc[d];
    }
  }
  }
  if (CERT) {
    const double t = block_sum(res2, red);
    if (threadIdx.x == 0) pres[(int64_t(l) * gridDim.x + blockIdx.x) * 4 + 3] = t;
  }
}

// Tiled apply: a block owns a 16 x 16 column of nodes over a z-chunk and marches
// the element layers below each owned node plane, computing every element
// once (plus a one-element halo in -x, -y and -z). Threads 0..255 take the
// elements under the owned nodes, threads 256..288 the halo elements. Per
// layer: the next node plane is staged in LDS (two planes, by parity), each element writes the forces
// on its three in-plane neighbour nodes (bottom and top) to LDS, and the
// owner of a node adds them in a fixed order; the top-plane sums carry to the
// next layer in registers. Sums are independent of the launch shape.
constexpr int kTile = 16;
constexpr int kTileThreads = 320;
constexpr int kSide = kTile + 2;  // staged node plane: x0-1 .. x0+16
constexpr int kElem = kTile + 1;  // element columns: x0-1 .. x0+15

__device__ __forceinline__ int modi(int v, int n) {
  const int r = v % n;
  return r < 0 ? r + n : r;
}

// SOA: vectors (L, 3, N) instead of (L, N, 3). PCG: the resident iteration's q = K p, where an inactive load
// writes nothing and each block leaves the p.q partial of its owned nodes in ppq[l * ppq_stride + block].
template <bool SOA, bool PCG>
__global__ __launch_bounds__(kTileThreads) void c21_tile_kernel(int nz, int ny, int nx, int zc, ElementCoef coef,
                                                                const double* pool, const int32_t* index,
                                                                const double* u, const uint8_t* active,
                                                                double* out, double* ppq, int ppq_stride) {
  __shared__ double np[2][kSide * kSide * 3];  // node planes z and z+1 (by parity)
  __shared__ double fx[2][kElem * kElem * 9];
  __shared__ double red[kTileThreads / 64];
  const int l = blockIdx.z;
  const int tiles_x = (nx + kTile - 1) / kTile;
  const int x0 = int(blockIdx.x % tiles_x) * kTile, y0 = int(blockIdx.x / tiles_x) * kTile;
  const int z0 = int(blockIdx.y) * zc, zend = min(z0 + zc, nz);
  const int64_t N = int64_t(nx) * ny * nz;
  const double* ul = u + int64_t(l) * 3 * N;
  double* ol = out + int64_t(l) * 3 * N;
  auto at = [&](int64_t node, int d) { return SOA ? int64_t(d) * N + node : 3 * node + d; };
  const int t = threadIdx.x;
  int i = 0, j = 0;
  bool has = true;
  if (t < kTile * kTile) {
    i = t % kTile;
    j = t / kTile;
  } else if (t < kTile * kTile + kElem) {
    i = -1;
    j = t - kTile * kTile - 1;
  } else if (t < kTile * kTile + kElem + kTile) {
    i = t - kTile * kTile - kElem;
    j = -1;
  } else {
    has = false;
  }
  has = has && x0 + i < nx && y0 + j < ny;
  const bool owner = t < kTile * kTile && has;
  const int gx = modi(x0 + i, nx), gy = modi(y0 + j, ny);

  if (active && !active[l]) {
    if (owner && !PCG)
      for (int z = z0; z < zend; ++z) {
        const int64_t node = (int64_t(z) * ny + gy) * nx + gx;
        for (int d = 0; d < 3; ++d) ol[at(node, d)] = 0.0;
      }
    return;
  }

  // staged plane entries of this thread: k = t and k = t + 320 (< 324)
  constexpr int kPlane = kSide * kSide;
  auto fetch = [&](int plane, double (&v)[2][3]) {
    const int pz = modi(plane, nz);
#pragma unroll
    for (int s = 0; s < 2; ++s) {
      const int k = t + s * kTileThreads;
      if (k < kPlane) {
        const int64_t node = (int64_t(pz) * ny + modi(y0 - 1 + k / kSide, ny)) * nx + modi(x0 - 1 + k % kSide, nx);
#pragma unroll
        for (int d = 0; d < 3; ++d) v[s][d] = ul[at(node, d)];
      }
    }
  };
  auto stage = [&](int plane, const double (&v)[2][3]) {
    double* b = np[plane & 1];
#pragma unroll
    for (int s = 0; s < 2; ++s) {
      const int k = t + s * kTileThreads;
      if (k < kPlane)
#pragma unroll
        for (int d = 0; d < 3; ++d) b[3 * k + d] = v[s][d];
    }
  };
  // element nodes of one plane, local order x + 2y
  auto read_nodes = [&](int plane, double* w) {
    const double* b = np[plane & 1];
#pragma unroll
    for (int n = 0; n < 4; ++n) {
      const int k = (j + 1 + (n >> 1)) * kSide + (i + 1 + (n & 1));
#pragma unroll
      for (int d = 0; d < 3; ++d) w[3 * n + d] = b[3 * k + d];
    }
  };

  double pre[2][3] = {};
  fetch(z0 - 1, pre);
  stage(z0 - 1, pre);
  fetch(z0, pre);
  stage(z0, pre);
  __syncthreads();

  const int slot = (j + 1) * kElem + (i + 1);
  double carry[3] = {0.0, 0.0, 0.0};
  double pq = 0.0;
#pragma unroll 1
  for (int z = z0 - 1; z < zend; ++z) {
    const bool more = z + 1 < zend;
    if (more) fetch(z + 2, pre);
    double fe[24], pown[3] = {0.0, 0.0, 0.0};
    if (has) {
      double ue[24], C[6][6];
      read_nodes(z, ue);
      read_nodes(z + 1, ue + 12);
      // local node 0 of the element is the node this thread owns (its input value at plane z)
#pragma unroll
      for (int d = 0; d < 3; ++d) pown[d] = ue[d];
      const int64_t e = (int64_t(modi(z, nz)) * ny + gy) * nx + gx;
      unpack_c21(c21_row(pool, index, e), C);
      hex8_element_c(ue, C, coef, fe);
#pragma unroll
      for (int n = 1; n < 4; ++n)
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          fx[0][slot * 9 + 3 * (n - 1) + d] = fe[3 * n + d];
          fx[1][slot * 9 + 3 * (n - 1) + d] = fe[3 * (n + 4) + d];
        }
    }
    __syncthreads();
    if (owner) {
      const int sx = slot - 1, sy = slot - kElem, sxy = slot - kElem - 1;
      double bottom[3];
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        bottom[d] = carry[d] + fe[d] + fx[0][sx * 9 + d] + fx[0][sy * 9 + 3 + d] + fx[0][sxy * 9 + 6 + d];
        carry[d] = fe[12 + d] + fx[1][sx * 9 + d] + fx[1][sy * 9 + 3 + d] + fx[1][sxy * 9 + 6 + d];
      }
      if (z >= z0) {
        const int64_t node = (int64_t(z) * ny + gy) * nx + gx;
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          ol[at(node, d)] = bottom[d];
          if (PCG) pq = fma(pown[d], bottom[d], pq);
        }
      }
    }
    if (more) stage(z + 2, pre);  // over plane z, which every thread has read before the barrier
    __syncthreads();
  }
  if (PCG) {
    const double total = block_sum(pq, red);
    if (t == 0) ppq[int64_t(l) * ppq_stride + blockIdx.y * gridDim.x + blockIdx.x] = total;
  }
}

// f = -sum_e V Bbar_e^T C_e E (the element force of the affine field of E), AoS (1, 3N).
__global__ __launch_bounds__(256) void c21_rhs_kernel(int nz, int ny, int nx, ElementCoef coef, const double* pool,
                                                      const int32_t* index, Strain6 E, double* out, int soa) {
  const int64_t N = int64_t(nx) * ny * nz;
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const int x = int(i % nx), y = int((i / nx) % ny), z = int(i / (int64_t(nx) * ny));
  double f[3] = {0.0, 0.0, 0.0};
#pragma unroll 1
  for (int o = 0; o < 8; ++o) {
    const int ex = wrapi(x - (o & 1), nx), ey = wrapi(y - ((o >> 1) & 1), ny), ez = wrapi(z - ((o >> 2) & 1), nz);
    double C[6][6];
    unpack_c21(c21_row(pool, index, (int64_t(ez) * ny + ey) * nx + ex), C);
    double s[6];
    for (int r = 0; r < 6; ++r) {
      double acc = 0.0;
      for (int q = 0; q < 6; ++q) acc += C[r][q] * E.v[q];
      s[r] = coef.vol * acc;
    }
    const double dx = (o & 1) ? coef.g[0] : -coef.g[0];
    const double dy = (o & 2) ? coef.g[1] : -coef.g[1];
    const double dz = (o & 4) ? coef.g[2] : -coef.g[2];
    f[0] -= dx * s[0] + dy * s[5] + dz * s[4];
    f[1] -= dx * s[5] + dy * s[1] + dz * s[3];
    f[2] -= dx * s[4] + dy * s[3] + dz * s[2];
  }
  for (int d = 0; d < 3; ++d) out[soa ? int64_t(d) * N + i : 3 * i + d] = f[d];
}

// Per-block sums of C_e (E + Bbar u_e) for the mean stress.
__global__ __launch_bounds__(256) void c21_stress_partials(int nz, int ny, int nx, ElementCoef coef, const double* pool,
                                                           const int32_t* index, const double* u, Strain6 E,
                                                           double* part) {
  __shared__ double red[8];
  const int64_t N = int64_t(nx) * ny * nz;
  const int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  double sig[6] = {0, 0, 0, 0, 0, 0};
  if (e < N) {
    const int x = int(e % nx), y = int((e / nx) % ny), z = int(e / (int64_t(nx) * ny));
    double eps[6];
    for (int k = 0; k < 6; ++k) eps[k] = E.v[k];
    for (int n = 0; n < 8; ++n) {
      const int64_t node = (int64_t(wrapi(z + ((n >> 2) & 1), nz)) * ny + wrapi(y + ((n >> 1) & 1), ny)) * nx +
                           wrapi(x + (n & 1), nx);
      const double ux = u[3 * node], uy = u[3 * node + 1], uz = u[3 * node + 2];
      const double dx = (n & 1) ? coef.g[0] : -coef.g[0];
      const double dy = (n & 2) ? coef.g[1] : -coef.g[1];
      const double dz = (n & 4) ? coef.g[2] : -coef.g[2];
      eps[0] += dx * ux;
      eps[1] += dy * uy;
      eps[2] += dz * uz;
      eps[3] += dz * uy + dy * uz;
      eps[4] += dz * ux + dx * uz;
      eps[5] += dy * ux + dx * uy;
    }
    double C[6][6];
    unpack_c21(c21_row(pool, index, e), C);
    for (int r = 0; r < 6; ++r) {
      double acc = 0.0;
      for (int q = 0; q < 6; ++q) acc += C[r][q] * eps[q];
      sig[r] = acc;
    }
  }
  for (int k = 0; k < 6; ++k) {
    const double t = block_sum(sig[k], red);
    if (threadIdx.x == 0) part[int64_t(blockIdx.x) * 6 + k] = t;
  }
}

__global__ void c21_stress_finish(const double* part, int nblk, int64_t N, double* out) {
  const int k = blockIdx.x;
  const double s = wave_reduce_partials(part + k, nblk, 6);
  if (threadIdx.x == 0) out[k] = s / double(N);
}

}  // namespace pgj

using namespace pgj;

namespace pgj {
namespace {
// Blocks of the tiled apply resident at once on the current device (CUs x blocks per CU), cached per device and
// kernel instance.
template <bool SOA, bool PCG>
int tile_slots(int* out) {
  static int cache[64] = {};
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e == hipSuccess && (dev < 0 || dev >= 64)) e = hipErrorInvalidDevice;
  if (e == hipSuccess && cache[dev] == 0) {
    int cus = 0, per_cu = 0;
    e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (e == hipSuccess)
      e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, c21_tile_kernel<SOA, PCG>, kTileThreads, 0);
    if (e == hipSuccess) cache[dev] = std::max(1, cus * std::max(1, per_cu));
  }
  if (e != hipSuccess) {
    set_error("device query for the tiled apply failed: %s", hipGetErrorString(e));
    return PGJ_ERR_HIP;
  }
  *out = cache[dev];
  return PGJ_OK;
}

int tiles_of(const pgj_grid& g) { return ((g.nx + kTile - 1) / kTile) * ((g.ny + kTile - 1) / kTile); }

// z-chunk with the fewest layer steps on the critical path: rounds of resident blocks x (chunk + halo layer)
template <bool SOA, bool PCG>
int tile_zc(const pgj_grid& g, int L, int* zc_out) {
  int slots = 0;
  if (int e = tile_slots<SOA, PCG>(&slots)) return e;
  int zc = 2;
  int64_t best = INT64_MAX;
  for (int c = 2; c <= 32; c *= 2) {
    const int64_t blocks = int64_t(tiles_of(g)) * L * ((g.nz + c - 1) / c);
    const int64_t cost = ((blocks + slots - 1) / slots) * (std::min(c, g.nz) + 1);
    if (cost < best) {
      best = cost;
      zc = c;
    }
  }
  *zc_out = zc;
  return PGJ_OK;
}

template <bool SOA, bool PCG>
int launch_tile(const pgj_grid& g, int L, const double* pool, const int32_t* index, const double* u,
                const uint8_t* active, double* out, double* ppq, int ppq_stride, hipStream_t s) {
  int zc = 0;
  if (int e = tile_zc<SOA, PCG>(g, L, &zc)) return e;
  hipLaunchKernelGGL((c21_tile_kernel<SOA, PCG>), dim3(tiles_of(g), (g.nz + zc - 1) / zc, L), dim3(kTileThreads), 0,
                     s, g.nz, g.ny, g.nx, zc, coef_of(g), pool, index, u, active, out, ppq, ppq_stride);
  PGJ_LAUNCHED("c21_tile_kernel");
  return PGJ_OK;
}

int c21_blocks(const pgj_grid& g) { return int((int64_t(g.nx) * g.ny * g.nz + 255) / 256); }
Strain6 macro_of(const double* m) {
  Strain6 E{};
  for (int k = 0; k < 6; ++k) E.v[k] = m[k];
  return E;
}
}  // namespace

int c21_pcg_blocks(const pgj_grid& g, int L, int* blocks) {
  int zc = 0;
  if (int e = tile_zc<true, true>(g, L, &zc)) return e;
  *blocks = tiles_of(g) * ((g.nz + zc - 1) / zc);
  return PGJ_OK;
}

int launch_c21_pcg(const pgj_grid& g, int L, const double* pool, const int32_t* index, const double* p,
                   const uint8_t* active, double* q, double* ppq, int ppq_stride, hipStream_t s) {
  return launch_tile<true, true>(g, L, pool, index, p, active, q, ppq, ppq_stride, s);
}

int launch_c21_certify(const pgj_grid& g, int L, const double* pool, const int32_t* index, const double* u,
                       const double* rhs, double* r, double* pres, hipStream_t s) {
  hipLaunchKernelGGL((c21_apply_kernel<true, true>), dim3(c21_blocks(g), L), dim3(256), 0, s, g.nz, g.ny, g.nx,
                     coef_of(g), pool, index, u, nullptr, r, rhs, pres);
  PGJ_LAUNCHED("c21_apply_kernel");
  return PGJ_OK;
}

int launch_c21_rhs(const pgj_grid& g, const double* pool, const int32_t* index, const double* macro, double* out,
                   int soa, hipStream_t s) {
  hipLaunchKernelGGL(c21_rhs_kernel, dim3(c21_blocks(g)), dim3(256), 0, s, g.nz, g.ny, g.nx, coef_of(g), pool, index,
                     macro_of(macro), out, soa);
  PGJ_LAUNCHED("c21_rhs_kernel");
  return PGJ_OK;
}
}  // namespace pgj

extern "C" int pgj_hybrid_apply(const pgj_grid* g, int L, const double* pool, const int32_t* index, const double* u,
                                const uint8_t* active, double* out, void* stream) {
  if (int e = check_grid(g)) return e;
  PGJ_REQUIRE(L >= 1 && L <= 65535 && pool && u && out, "invalid hybrid apply arguments");
  return launch_tile<false, false>(*g, L, pool, index, u, active, out, nullptr, 0, as_stream(stream));
}

extern "C" int pgj_hybrid_apply_gather(const pgj_grid* g, int L, const double* pool, const int32_t* index,
                                       const double* u, const uint8_t* active, double* out, void* stream) {
  if (int e = check_grid(g)) return e;
  PGJ_REQUIRE(L >= 1 && L <= 65535 && pool && u && out, "invalid hybrid apply arguments");
  hipLaunchKernelGGL((c21_apply_kernel<false, false>), dim3(c21_blocks(*g), L), dim3(256), 0, as_stream(stream),
                     g->nz, g->ny, g->nx, coef_of(*g), pool, index, u, active, out, nullptr, nullptr);
  PGJ_LAUNCHED("c21_apply_kernel");
  return PGJ_OK;
}

extern "C" int pgj_hybrid_rhs(const pgj_grid* g, const double* pool, const int32_t* index, const double macro[6],
                              double* out, void* stream) {
  if (int e = check_grid(g)) return e;
  PGJ_REQUIRE(pool && macro && out, "invalid hybrid rhs arguments");
  return launch_c21_rhs(*g, pool, index, macro, out, 0, as_stream(stream));
}

extern "C" size_t pgj_hybrid_stress_ws_bytes(const pgj_grid* g) {
  if (!g) return 0;
  return align256(size_t(pgj::c21_blocks(*g)) * 6 * 8);
}

extern "C" int pgj_hybrid_avg_stress(const pgj_grid* g, const double* pool, const int32_t* index, const double* u,
                                     const double macro[6], double* ws, double* out6, void* stream) {
  if (int e = check_grid(g)) return e;
  PGJ_REQUIRE(pool && u && macro && ws && out6, "invalid hybrid stress arguments");
  const int nblk = pgj::c21_blocks(*g);
  hipStream_t s = as_stream(stream);
  hipLaunchKernelGGL(c21_stress_partials, dim3(nblk), dim3(256), 0, s, g->nz, g->ny, g->nx, coef_of(*g), pool, index, u,
                     pgj::macro_of(macro), ws);
  PGJ_LAUNCHED("c21_stress_partials");
  hipLaunchKernelGGL(c21_stress_finish, dim3(6), dim3(64), 0, s, ws, nblk, int64_t(g->nx) * g->ny * g->nz, out6);
  PGJ_LAUNCHED("c21_stress_finish");
  return PGJ_OK;
}
