// HEX8 operator kernels: the row-march apply (drop-in and fused PCG form),
// the independent thread-per-node reference apply, rhs, energies, C^H,
// mean stress, nodal diagonal / partial Green-Jacobi scale and projection.
#include <cmath>
#include <mutex>

#include "hex8_kernels.h"

namespace pgj {

// Raw inputs of one node row segment (the input vector: u, or the PCG direction p).
template <int MODE, int R, int XPL, typename T>
struct Rows {
  T v[R + 1][XPL][3];
};

// Row-march HEX8 apply. Lanes map to x (XPL consecutive x per lane), each of
// the W row-waves marches R element rows over the z-layers of the block, a
// halo wave computes the element row below the block, and contributions to a
// neighbour wave's first node row go through a double-buffered LDS slot. The
// next plane's inputs are fetched before the current layer is computed.
// PLANE selects the plane-carry march: the in-plane butterflies of each node
// plane are computed once and shared by the two element layers that touch it
// (forward: the top plane of layer z is the bottom plane of layer z+1;
// transposed: the modal forces of both layers on a plane are summed before
// one in-plane transform), which halves the in-plane butterflies, the x
// neighbour exchanges and the LDS traffic per element.
template <int MODE, int XPL, int R, bool CUBIC, int MAXT, typename T, bool PLANE>
__global__ __launch_bounds__(MAXT) void hex8_rowmarch(RowArgsT<T> A) {
  // the drop-in apply works on AoS vectors, the resident iteration on SoA
  constexpr bool SOA = MODE == kModePCG;
  constexpr unsigned ES = sizeof(T);
  extern __shared__ __align__(16) unsigned char smem[];
  T* lds = reinterpret_cast<T*>(smem);
  const int lane = threadIdx.x & 63;
  // wave index and everything derived from it are wave-uniform: keep them in SGPRs
  const int wave = __builtin_amdgcn_readfirstlane(int(threadIdx.x >> 6));
  const int l = blockIdx.z;
  const T* __restrict__ mat = A.a + size_t(design_of(l, A.group)) * A.N;
  const RowMarch& c = A.cfg;
  const int W = c.W;
  const int nx = A.nx, ny = A.ny, nz = A.nz;
  const unsigned N = unsigned(A.N);
  // XCD-aware row groups: workgroups are dealt to the 8 XCDs round robin by
  // linear id, so when the row-group count is a multiple of 8 each XCD gets a
  // contiguous range of row groups, and the halo row a block reads has just
  // been fetched into the same L2 by its neighbour.
  const int nrg = int(gridDim.x);
  const int rg = nrg % 8 == 0 ? int(blockIdx.x % 8) * (nrg / 8) + int(blockIdx.x / 8) : int(blockIdx.x);
  const int y0 = rg * W * R, z0 = blockIdx.y * c.ZC;
  const int xl = lane % c.nl;
  const bool lane_on = lane < c.nl;
  const unsigned xo = unsigned(xl * XPL);

  if (MODE == kModePCG) {
    if (!A.st.active[l]) return;
  } else if (A.mask && !A.mask[l]) {
    // Inactive load: zeros, as the reference's pre-zeroed output.
    T* o = A.out + size_t(l) * 3 * N;
    if (wave < W && lane_on) {
      for (int j = 0; j < R; ++j)
        for (int zi = 0; zi < c.ZC; ++zi)
          for (int xi = 0; xi < XPL; ++xi) {
            const unsigned node = unsigned((z0 + zi) * ny + (y0 + wave * R + j)) * unsigned(nx) + xo + xi;
            for (int d = 0; d < 3; ++d) o[3 * node + d] = T(0);
          }
    }
    return;
  }

  const bool halo = wave == W;
  const int ybase = halo ? y0 - 1 : y0 + wave * R;
  const int rows = halo ? 1 : R;  // element rows computed by this wave
  // uniform base pointers per component (SoA) or per load (AoS)
  const T* __restrict__ in = A.u + size_t(l) * 3 * N;
  T* __restrict__ out = A.out + size_t(l) * 3 * N;
  // element offsets in bytes (32-bit): uniform base + 32-bit offset selects the SGPR-base addressing form
  auto at = [&](unsigned node, int d) -> unsigned { return ES * (SOA ? unsigned(d) * N + node : 3u * node + unsigned(d)); };
  auto ld = [](const T* base, unsigned off) -> T {
    return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + off);
  };
  auto st = [](T* base, unsigned off, T v) { *reinterpret_cast<T*>(reinterpret_cast<char*>(base) + off) = v; };
  auto row_of = [&](int plane, int j) -> unsigned {
    return unsigned(plane * ny + wrapi(ybase + j, ny)) * unsigned(nx);
  };

  auto fetch = [&](int plane, Rows<MODE, R, XPL, T>& dst, T (&arow)[R][XPL]) {
#pragma unroll
    for (int j = 0; j <= R; ++j) {
      if (j > rows) break;
      const unsigned rb = row_of(plane, j);
#pragma unroll
      for (int xi = 0; xi < XPL; ++xi) {
        const unsigned node = rb + xo + xi;
#pragma unroll
        for (int d = 0; d < 3; ++d) dst.v[j][xi][d] = ld(in, at(node, d));
        if (j < R) arow[j][xi] = ld(mat, ES * node);
      }
    }
  };
  // Node values of a fetched plane.
  auto settle = [&](const Rows<MODE, R, XPL, T>& src, T (&dst)[R + 1][XPL][3]) {
#pragma unroll
    for (int j = 0; j <= R; ++j) {
      if (j > rows) break;
#pragma unroll
      for (int xi = 0; xi < XPL; ++xi)
#pragma unroll
        for (int d = 0; d < 3; ++d) dst[j][xi][d] = src.v[j][xi][d];
    }
  };

  double pq = 0.0;
  const int src_right = (xl + 1) % c.nl;       // lane holding x + XPL
  const int src_left = (xl + c.nl - 1) % c.nl;  // lane holding x - 1
  if constexpr (PLANE) {
    T nodeU[R + 1][XPL][3];
    T nodeLo[R][XPL][3];             // bottom-plane nodes of the owned rows (p.q)
    T PL[R][XPL][3][4];              // in-plane modes of the bottom plane, per element column
    T CM[R][XPL][3][4];              // modal forces on the bottom plane from the layer below
    T accL[R + 1][XPL][3];
    T aL[R][XPL], aU[R][XPL], aN[R][XPL];
    const int slot_size = (c.W + 1) * 64 * 3 * XPL;
    // in-plane modes of every element column of the rows this wave computes
    auto plane_modes = [&](const T (&nd)[R + 1][XPL][3], T (&P)[R][XPL][3][4]) {
#pragma unroll
      for (int j = 0; j < R; ++j) {
        if (j >= rows) break;
        T nb[2][3];
#pragma unroll
        for (int oy = 0; oy < 2; ++oy)
#pragma unroll
          for (int d = 0; d < 3; ++d) nb[oy][d] = __shfl(nd[j + oy][0][d], src_right, 64);
#pragma unroll
        for (int xi = 0; xi < XPL; ++xi)
#pragma unroll
          for (int d = 0; d < 3; ++d)
            plane_forward(nd[j][xi][d], xi + 1 < XPL ? nd[j][xi + 1][d] : nb[0][d], nd[j + 1][xi][d],
                          xi + 1 < XPL ? nd[j + 1][xi + 1][d] : nb[1][d], P[j][xi][d]);
      }
    };
    Rows<MODE, R, XPL, T> raw;
    fetch(wrapi(z0 - 1, nz), raw, aL);
    settle(raw, nodeU);
    plane_modes(nodeU, PL);
#pragma unroll
    for (int j = 0; j < R; ++j)
#pragma unroll
      for (int xi = 0; xi < XPL; ++xi)
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          nodeLo[j][xi][d] = nodeU[j][xi][d];
#pragma unroll
          for (int m = 0; m < 4; ++m) CM[j][xi][d][m] = T(0);
        }
    fetch(z0, raw, aU);
    for (int li = 0; li <= c.ZC; ++li) {
      const int zl = wrapi(z0 - 1 + li, nz), zu = wrapi(z0 + li, nz);
      settle(raw, nodeU);
      if (li < c.ZC) fetch(wrapi(z0 + li + 1, nz), raw, aN);  // prefetch the next plane
#pragma unroll
      for (int j = 0; j <= R; ++j)
#pragma unroll
        for (int xi = 0; xi < XPL; ++xi)
#pragma unroll
          for (int d = 0; d < 3; ++d) accL[j][xi][d] = T(0);
#pragma unroll
      for (int j = 0; j < R; ++j) {
        if (j >= rows) break;
        T nb[2][3];
#pragma unroll
        for (int oy = 0; oy < 2; ++oy)
#pragma unroll
          for (int d = 0; d < 3; ++d) nb[oy][d] = __shfl(nodeU[j + oy][0][d], src_right, 64);
        T spill[2][3];  // contributions of this lane's last element column to x + 1 (next lane)
#pragma unroll
        for (int xi = 0; xi < XPL; ++xi) {
          T PU[3][4], Um[3][8], F[3][8];
#pragma unroll
          for (int d = 0; d < 3; ++d) {
            plane_forward(nodeU[j][xi][d], xi + 1 < XPL ? nodeU[j][xi + 1][d] : nb[0][d], nodeU[j + 1][xi][d],
                          xi + 1 < XPL ? nodeU[j + 1][xi + 1][d] : nb[1][d], PU[d]);
#pragma unroll
            for (int m = 0; m < 4; ++m) {
              Um[d][m] = PL[j][xi][d][m] + PU[d][m];
              Um[d][m | 4] = PU[d][m] - PL[j][xi][d][m];
            }
          }
          hex8_modal_core<CUBIC, T>(Um, aL[j][xi], A.coef, F);
#pragma unroll
          for (int d = 0; d < 3; ++d) {
            T M[4], f[4];
#pragma unroll
            for (int m = 0; m < 4; ++m) {
              M[m] = CM[j][xi][d][m] + (F[d][m] - F[d][m | 4]);
              CM[j][xi][d][m] = F[d][m] + F[d][m | 4];
              PL[j][xi][d][m] = PU[d][m];
            }
            plane_transposed(M, f);
            accL[j][xi][d] += f[0];
            accL[j + 1][xi][d] += f[2];
            if (xi + 1 < XPL) {
              accL[j][xi + 1][d] += f[1];
              accL[j + 1][xi + 1][d] += f[3];
            } else {
              spill[0][d] = f[1];
              spill[1][d] = f[3];
            }
          }
        }
#pragma unroll
        for (int oy = 0; oy < 2; ++oy)
#pragma unroll
          for (int d = 0; d < 3; ++d) accL[j + oy][0][d] += __shfl(spill[oy][d], src_left, 64);
      }
      // Export the contributions to the first node row of the next wave.
      T* slots = lds + (li & 1) * slot_size;
      T* slot = slots + wave * 64 * 3 * XPL;
#pragma unroll
      for (int xi = 0; xi < XPL; ++xi)
#pragma unroll
        for (int d = 0; d < 3; ++d) slot[(xi * 3 + d) * 64 + lane] = halo ? accL[1][xi][d] : accL[R][xi][d];
      __syncthreads();
      if (!halo) {
        const T* from = slots + (wave == 0 ? W : wave - 1) * 64 * 3 * XPL;
#pragma unroll
        for (int xi = 0; xi < XPL; ++xi)
#pragma unroll
          for (int d = 0; d < 3; ++d) accL[0][xi][d] += from[(xi * 3 + d) * 64 + lane];
        if (li >= 1 && lane_on) {
#pragma unroll
          for (int j = 0; j < R; ++j) {
            const unsigned rb = unsigned(zl * ny + ybase + j) * unsigned(nx);
#pragma unroll
            for (int xi = 0; xi < XPL; ++xi)
#pragma unroll
              for (int d = 0; d < 3; ++d) {
                st(out, at(rb + xo + xi, d), accL[j][xi][d]);
                if (MODE == kModePCG) pq = fma(double(nodeLo[j][xi][d]), double(accL[j][xi][d]), pq);
              }
          }
        }
      }
#pragma unroll
      for (int j = 0; j < R; ++j)
#pragma unroll
        for (int xi = 0; xi < XPL; ++xi) {
#pragma unroll
          for (int d = 0; d < 3; ++d) nodeLo[j][xi][d] = nodeU[j][xi][d];
          aL[j][xi] = aU[j][xi];
          aU[j][xi] = aN[j][xi];
        }
    }
    if (MODE == kModePCG) {
      const double total = block_sum(pq, reinterpret_cast<double*>(lds + 2 * slot_size));
      if (threadIdx.x == 0) A.ppq[int64_t(l) * A.ppq_stride + blockIdx.y * gridDim.x + rg] = total;
    }
    return;
  } else {
  T nodeL[R + 1][XPL][3], nodeU[R + 1][XPL][3];
  T carry[R][XPL][3];
  T accL[R + 1][XPL][3], accU[R + 1][XPL][3];
  T aL[R][XPL], aU[R][XPL], aN[R][XPL];
#pragma unroll
  for (int j = 0; j < R; ++j)
#pragma unroll
    for (int xi = 0; xi < XPL; ++xi)
#pragma unroll
      for (int d = 0; d < 3; ++d) carry[j][xi][d] = T(0);

  const int slot_size = (c.W + 1) * 64 * 6 * XPL;

  Rows<MODE, R, XPL, T> raw;
  fetch(wrapi(z0 - 1, nz), raw, aL);
  settle(raw, nodeL);
  fetch(z0, raw, aU);
  for (int li = 0; li <= c.ZC; ++li) {
    const int zl = wrapi(z0 - 1 + li, nz), zu = wrapi(z0 + li, nz);
    settle(raw, nodeU);
    if (li < c.ZC) fetch(wrapi(z0 + li + 1, nz), raw, aN);  // prefetch the next plane
#pragma unroll
    for (int j = 0; j <= R; ++j)
#pragma unroll
      for (int xi = 0; xi < XPL; ++xi)
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          accL[j][xi][d] = (j < R) ? carry[j][xi][d] : T(0);
          accU[j][xi][d] = T(0);
        }
#pragma unroll
    for (int j = 0; j < R; ++j) {
      if (j >= rows) break;
      // Right neighbours (x + XPL) of the four node rows this element row touches.
      T nb[2][2][3];
#pragma unroll
      for (int oy = 0; oy < 2; ++oy)
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          nb[oy][0][d] = __shfl(nodeL[j + oy][0][d], src_right, 64);
          nb[oy][1][d] = __shfl(nodeU[j + oy][0][d], src_right, 64);
        }
      T spill[2][2][3];  // contributions of this lane's last element to x + 1 (next lane)
#pragma unroll
      for (int xi = 0; xi < XPL; ++xi) {
        T ue[24], fe[24];
#pragma unroll
        for (int n = 0; n < 8; ++n) {
          const int ox = n & 1, oy = (n >> 1) & 1, oz = (n >> 2) & 1;
#pragma unroll
          for (int d = 0; d < 3; ++d) {
            T v;
            if (ox == 0 || xi + 1 < XPL) {
              const int xs = xi + ox;
              v = oz ? nodeU[j + oy][xs][d] : nodeL[j + oy][xs][d];
            } else {
              v = nb[oy][oz][d];
            }
            ue[3 * n + d] = v;
          }
        }
        hex8_element<CUBIC, T>(ue, aL[j][xi], A.coef, fe);
#pragma unroll
        for (int n = 0; n < 8; ++n) {
          const int ox = n & 1, oy = (n >> 1) & 1, oz = (n >> 2) & 1;
#pragma unroll
          for (int d = 0; d < 3; ++d) {
            const T v = fe[3 * n + d];
            if (ox == 0) {
              if (oz) accU[j + oy][xi][d] += v; else accL[j + oy][xi][d] += v;
            } else if (xi + 1 < XPL) {
              if (oz) accU[j + oy][xi + 1][d] += v; else accL[j + oy][xi + 1][d] += v;
            } else {
              spill[oy][oz][d] = v;
            }
          }
        }
      }
#pragma unroll
      for (int oy = 0; oy < 2; ++oy)
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          accL[j + oy][0][d] += __shfl(spill[oy][0][d], src_left, 64);
          accU[j + oy][0][d] += __shfl(spill[oy][1][d], src_left, 64);
        }
    }
    // Export the contributions to the first node row of the next wave.
    T* slots = lds + (li & 1) * slot_size;
    T* slot = slots + wave * 64 * 6 * XPL;
#pragma unroll
    for (int xi = 0; xi < XPL; ++xi)
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        slot[(xi * 6 + d) * 64 + lane] = halo ? accL[1][xi][d] : accL[R][xi][d];
        slot[(xi * 6 + 3 + d) * 64 + lane] = halo ? accU[1][xi][d] : accU[R][xi][d];
      }
    __syncthreads();
    if (!halo) {
      const T* from = slots + (wave == 0 ? W : wave - 1) * 64 * 6 * XPL;
#pragma unroll
      for (int xi = 0; xi < XPL; ++xi)
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          accL[0][xi][d] += from[(xi * 6 + d) * 64 + lane];
          accU[0][xi][d] += from[(xi * 6 + 3 + d) * 64 + lane];
        }
      if (li >= 1 && lane_on) {
#pragma unroll
        for (int j = 0; j < R; ++j) {
          const unsigned rb = unsigned(zl * ny + ybase + j) * unsigned(nx);
#pragma unroll
          for (int xi = 0; xi < XPL; ++xi) {
#pragma unroll
            for (int d = 0; d < 3; ++d) {
              st(out, at(rb + xo + xi, d), accL[j][xi][d]);
              if (MODE == kModePCG) pq = fma(double(nodeL[j][xi][d]), double(accL[j][xi][d]), pq);
            }
          }
        }
      }
#pragma unroll
      for (int j = 0; j < R; ++j)
#pragma unroll
        for (int xi = 0; xi < XPL; ++xi)
#pragma unroll
          for (int d = 0; d < 3; ++d) carry[j][xi][d] = accU[j][xi][d];
    }
#pragma unroll
    for (int j = 0; j <= R; ++j)
#pragma unroll
      for (int xi = 0; xi < XPL; ++xi) {
#pragma unroll
        for (int d = 0; d < 3; ++d) nodeL[j][xi][d] = nodeU[j][xi][d];
        if (j < R) {
          aL[j][xi] = aU[j][xi];
          aU[j][xi] = aN[j][xi];
        }
      }
  }
  if (MODE == kModePCG) {
    const double total = block_sum(pq, reinterpret_cast<double*>(lds + 2 * slot_size));
    if (threadIdx.x == 0) A.ppq[int64_t(l) * A.ppq_stride + blockIdx.y * gridDim.x + rg] = total;
  }
  }
}

template <bool CUBIC, bool CERT>
__global__ __launch_bounds__(256) void hex8_reference(RefArgs A) {
  __shared__ double red[8];
  const int l = blockIdx.y;
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  const int64_t base = int64_t(l) * 3 * A.N;
  double res2 = 0.0;
  if (i < A.N) {
    const int x = int(i % A.nx), y = int((i / A.nx) % A.ny), z = int(i / (int64_t(A.nx) * A.ny));
    double acc[3] = {0.0, 0.0, 0.0};
    // one element at a time (unrolling the eight elements spills); the node's
    // local index in element o is o, selected with static register indices
#pragma unroll 1
    for (int o = 0; o < 8; ++o) {
      const int ex = wrapi(x - (o & 1), A.nx), ey = wrapi(y - ((o >> 1) & 1), A.ny),
                ez = wrapi(z - ((o >> 2) & 1), A.nz);
      double ue[24], fe[24];
#pragma unroll
      for (int n = 0; n < 8; ++n) {
        const int nxp = wrapi(ex + (n & 1), A.nx), nyp = wrapi(ey + ((n >> 1) & 1), A.ny),
                  nzp = wrapi(ez + ((n >> 2) & 1), A.nz);
        const int64_t node = (int64_t(nzp) * A.ny + nyp) * A.nx + nxp;
#pragma unroll
        for (int d = 0; d < 3; ++d) ue[3 * n + d] = A.u[A.soa ? (int64_t(l) * 3 + d) * A.N + node : base + 3 * node + d];
      }
      const double ae = A.a[int64_t(design_of(l, A.group)) * A.N + (int64_t(ez) * A.ny + ey) * A.nx + ex];
      hex8_element<CUBIC>(ue, ae, A.coef, fe);
#pragma unroll
      for (int k = 0; k < 8; ++k)
        if (k == o) {
#pragma unroll
          for (int d = 0; d < 3; ++d) acc[d] += fe[3 * k + d];
        }
    }
    for (int d = 0; d < 3; ++d) {
      const int64_t idx = A.soa ? (int64_t(l) * 3 + d) * A.N + i : base + 3 * i + d;
      if (CERT) {
        const double r = A.rhs[idx] - acc[d];
        A.out[idx] = r;
        res2 = fma(r, r, res2);
      } else {
        A.out[idx] = acc[d];
      }
    }
  }
  if (CERT) {
    const double t = block_sum(res2, red);
    if (threadIdx.x == 0) A.pres[(int64_t(l) * gridDim.x + blockIdx.x) * 4 + 3] = t;
  }
}

// Per-block component sums of a (L, 3N) vector into partial rows of 4 doubles.
__device__ __forceinline__ int64_t lidx(int soa, int l, int d, int64_t i, int64_t N) {
  return soa ? (int64_t(l) * 3 + d) * N + i : (int64_t(l) * N + i) * 3 + d;
}

__global__ __launch_bounds__(256) void component_sums(const double* v, int64_t N, double* part, int soa) {
  __shared__ double red[8];
  const int l = blockIdx.y;
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  double s[3] = {0.0, 0.0, 0.0};
  if (i < N)
    for (int d = 0; d < 3; ++d) s[d] = v[lidx(soa, l, d, i, N)];
  for (int d = 0; d < 3; ++d) {
    const double t = block_sum(s[d], red);
    if (threadIdx.x == 0) part[(int64_t(l) * gridDim.x + blockIdx.x) * 4 + d] = t;
  }
}

// v -= mean per load/component (mean from component_sums partials; the first
// wave reduces them in a fixed order and shares the result through LDS).
__global__ __launch_bounds__(256) void subtract_mean(double* v, int64_t N, const double* part, int nblk, int soa) {
  __shared__ double mean[3];
  const int l = blockIdx.y;
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (threadIdx.x < 64) {
    const double* p = part + int64_t(l) * nblk * 4;
    for (int d = 0; d < 3; ++d) {
      const double s = wave_reduce_partials(p + d, nblk, 4);
      if (threadIdx.x == 0) mean[d] = s / double(N);
    }
  }
  __syncthreads();
  if (i < N)
    for (int d = 0; d < 3; ++d) v[lidx(soa, l, d, i, N)] -= mean[d];
}

// AoS (L, N, 3) <-> SoA (L, 3, N)
__global__ __launch_bounds__(256) void transpose_kernel(const double* in, double* out, int64_t N, int to_soa) {
  const int l = blockIdx.y;
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= N) return;
  for (int d = 0; d < 3; ++d) {
    if (to_soa) out[lidx(1, l, d, i, N)] = in[lidx(0, l, d, i, N)];
    else out[lidx(0, l, d, i, N)] = in[lidx(1, l, d, i, N)];
  }
}

// ------------------------------------------------------------------------ rhs
// f_node = -sum_{8 elements} a_e V Bbar_a^T C0 eps with centroid gradients +-1/(4h).
__global__ __launch_bounds__(256) void hex8_rhs_kernel(int nz, int ny, int nx, ElementCoef c, const double* a,
                                                       double m0, double m1, double m2, double m3, double m4,
                                                       double m5, double* out, int soa) {
  const int64_t N = int64_t(nx) * ny * nz;
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const int x = int(i % nx), y = int((i / nx) % ny), z = int(i / (int64_t(nx) * ny));
  // C0 eps (Voigt, engineering shear)
  const double tr = m0 + m1 + m2;
  const double s0 = c.lam * tr + 2 * c.mu * m0, s1 = c.lam * tr + 2 * c.mu * m1, s2 = c.lam * tr + 2 * c.mu * m2;
  const double s3 = c.mu * m3, s4 = c.mu * m4, s5 = c.mu * m5;
  double f[3] = {0.0, 0.0, 0.0};
  for (int o = 0; o < 8; ++o) {
    const int ex = wrapi(x - (o & 1), nx), ey = wrapi(y - ((o >> 1) & 1), ny), ez = wrapi(z - ((o >> 2) & 1), nz);
    const double ae = a[(int64_t(ez) * ny + ey) * nx + ex] * c.vol;
    const double dx = (o & 1) ? c.g[0] : -c.g[0];
    const double dy = (o & 2) ? c.g[1] : -c.g[1];
    const double dz = (o & 4) ? c.g[2] : -c.g[2];
    f[0] -= ae * (dx * s0 + dy * s5 + dz * s4);
    f[1] -= ae * (dx * s5 + dy * s1 + dz * s3);
    f[2] -= ae * (dx * s4 + dy * s3 + dz * s2);
  }
  for (int d = 0; d < 3; ++d) out[lidx(soa, 0, d, i, N)] = f[d];
}

// Affine element displacement chi_n,d = (E * (offset_n * h))_d for a macro strain.
__device__ __forceinline__ void affine_local(const double* m, const ElementCoef& c, double (&chi)[24]) {
  const double h[3] = {1.0 / (4.0 * c.g[0]), 1.0 / (4.0 * c.g[1]), 1.0 / (4.0 * c.g[2])};
  const double E[3][3] = {{m[0], 0.5 * m[5], 0.5 * m[4]}, {0.5 * m[5], m[1], 0.5 * m[3]}, {0.5 * m[4], 0.5 * m[3], m[2]}};
#pragma unroll
  for (int n = 0; n < 8; ++n) {
    const double p[3] = {(n & 1) * h[0], ((n >> 1) & 1) * h[1], ((n >> 2) & 1) * h[2]};
#pragma unroll
    for (int d = 0; d < 3; ++d) chi[3 * n + d] = E[d][0] * p[0] + E[d][1] * p[1] + E[d][2] * p[2];
  }
}

struct MacroSet {
  double m[6][6];
};

struct DiagTable {
  double v[8][3];
};

__device__ __forceinline__ void gather_element(const double* u, int nz, int ny, int nx, int x, int y, int z,
                                               double (&ue)[24]) {
#pragma unroll
  for (int n = 0; n < 8; ++n) {
    const int xx = wrapi(x + (n & 1), nx), yy = wrapi(y + ((n >> 1) & 1), ny), zz = wrapi(z + ((n >> 2) & 1), nz);
    const int64_t node = (int64_t(zz) * ny + yy) * nx + xx;
#pragma unroll
    for (int d = 0; d < 3; ++d) ue[3 * n + d] = u[3 * node + d];
  }
}

template <bool CUBIC>
__global__ __launch_bounds__(256) void hex8_energy_kernel(int nz, int ny, int nx, ElementCoef c, const double* u,
                                                          MacroSet ms, double* out) {
  const int64_t N = int64_t(nx) * ny * nz;
  const int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  const int l = blockIdx.y;
  if (e >= N) return;
  const int x = int(e % nx), y = int((e / nx) % ny), z = int(e / (int64_t(nx) * ny));
  double ue[24], fe[24], chi[24];
  gather_element(u + int64_t(l) * 3 * N, nz, ny, nx, x, y, z, ue);
  affine_local(ms.m[l], c, chi);
#pragma unroll
  for (int k = 0; k < 24; ++k) ue[k] += chi[k];
  hex8_element<CUBIC>(ue, 1.0, c, fe);
  double en = 0.0;
#pragma unroll
  for (int k = 0; k < 24; ++k) en = fma(ue[k], fe[k], en);
  out[int64_t(l) * N + e] = en / c.vol;
}

// C^H partials: per block the 21 upper-triangle sums of a_e (Ke v_i) . v_j / V.
template <bool CUBIC>
__global__ __launch_bounds__(256) void effective_partials(int nz, int ny, int nx, ElementCoef c, const double* a,
                                                          const double* u6, double* part) {
  __shared__ double red[21][4];
  const int64_t N = int64_t(nx) * ny * nz;
  const int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  double s[21];
#pragma unroll
  for (int k = 0; k < 21; ++k) s[k] = 0.0;
  if (e < N) {
    const int x = int(e % nx), y = int((e / nx) % ny), z = int(e / (int64_t(nx) * ny));
    const double ae = a[e];
    int idx = 0;
    for (int i = 0; i < 6; ++i) {
      double m[6] = {0, 0, 0, 0, 0, 0};
      m[i] = 1.0;
      double vi[24], fi[24], chi[24];
      gather_element(u6 + int64_t(i) * 3 * N, nz, ny, nx, x, y, z, vi);
      affine_local(m, c, chi);
#pragma unroll
      for (int k = 0; k < 24; ++k) vi[k] += chi[k];
      hex8_element<CUBIC>(vi, 1.0, c, fi);
      for (int j = i; j < 6; ++j, ++idx) {
        double vj[24];
        double mj[6] = {0, 0, 0, 0, 0, 0};
        mj[j] = 1.0;
        gather_element(u6 + int64_t(j) * 3 * N, nz, ny, nx, x, y, z, vj);
        affine_local(mj, c, chi);
        double t = 0.0;
#pragma unroll
        for (int k = 0; k < 24; ++k) t = fma(fi[k], vj[k] + chi[k], t);
        s[idx] = ae * t / c.vol;
      }
    }
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < 21; ++k) {
    const double t = wave_sum(s[k]);
    if (lane == 0) red[k][wave] = t;
  }
  __syncthreads();
  if (threadIdx.x < 21) {
    double t = 0.0;
    for (int w = 0; w < 4; ++w) t += red[threadIdx.x][w];
    part[int64_t(blockIdx.x) * 21 + threadIdx.x] = t;
  }
}

__global__ void effective_finish(const double* part, int nblk, int64_t N, double* out) {
  const int idx = blockIdx.x;  // one wavefront per upper-triangle entry
  const double s = wave_reduce_partials(part + idx, nblk, 21);
  if (threadIdx.x != 0) return;
  int k = 0;
  for (int i = 0; i < 6; ++i)
    for (int j = i; j < 6; ++j, ++k)
      if (k == idx) {
        out[i * 6 + j] = s / double(N);
        out[j * 6 + i] = s / double(N);
      }
}

__global__ __launch_bounds__(256) void avg_stress_partials(int nz, int ny, int nx, ElementCoef c, const double* a,
                                                           const double* u, MacroSet ms, double* part) {
  __shared__ double red[8];
  const int64_t N = int64_t(nx) * ny * nz;
  const int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  double sig[6] = {0, 0, 0, 0, 0, 0};
  if (e < N) {
    const int x = int(e % nx), y = int((e / nx) % ny), z = int(e / (int64_t(nx) * ny));
    double eps[6];
    for (int k = 0; k < 6; ++k) eps[k] = ms.m[0][k];
    for (int n = 0; n < 8; ++n) {
      const int xx = wrapi(x + (n & 1), nx), yy = wrapi(y + ((n >> 1) & 1), ny), zz = wrapi(z + ((n >> 2) & 1), nz);
      const int64_t node = (int64_t(zz) * ny + yy) * nx + xx;
      const double ux = u[3 * node], uy = u[3 * node + 1], uz = u[3 * node + 2];
      const double dx = (n & 1) ? c.g[0] : -c.g[0], dy = (n & 2) ? c.g[1] : -c.g[1], dz = (n & 4) ? c.g[2] : -c.g[2];
      eps[0] += dx * ux;
      eps[1] += dy * uy;
      eps[2] += dz * uz;
      eps[3] += dz * uy + dy * uz;
      eps[4] += dz * ux + dx * uz;
      eps[5] += dy * ux + dx * uy;
    }
    const double ae = a[e];
    const double tr = eps[0] + eps[1] + eps[2];
    for (int k = 0; k < 3; ++k) sig[k] = ae * (c.lam * tr + 2 * c.mu * eps[k]);
    for (int k = 3; k < 6; ++k) sig[k] = ae * c.mu * eps[k];
  }
  for (int k = 0; k < 6; ++k) {
    const double t = block_sum(sig[k], red);
    if (threadIdx.x == 0) part[int64_t(blockIdx.x) * 6 + k] = t;
  }
}

__global__ void avg_stress_finish(const double* part, int nblk, int64_t N, double* out) {
  const int k = blockIdx.x;
  const double s = wave_reduce_partials(part + k, nblk, 6);
  if (threadIdx.x == 0) out[k] = s / double(N);
}

// Nodal diagonal: d_i = sum_o a_{i-o} diag(Ke)[o]; log partials for the scale.
__global__ __launch_bounds__(256) void diagonal_kernel(int nz, int ny, int nx, const double* a, DiagTable ed,
                                                       double* d, double* logpart) {
  __shared__ double red[8];
  const int64_t N = int64_t(nx) * ny * nz;
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  double lsum = 0.0;
  if (i < N) {
    const int x = int(i % nx), y = int((i / nx) % ny), z = int(i / (int64_t(nx) * ny));
    double acc[3] = {0.0, 0.0, 0.0};
    for (int o = 0; o < 8; ++o) {
      const int ex = wrapi(x - (o & 1), nx), ey = wrapi(y - ((o >> 1) & 1), ny), ez = wrapi(z - ((o >> 2) & 1), nz);
      const double ae = a[(int64_t(ez) * ny + ey) * nx + ex];
      for (int k = 0; k < 3; ++k) acc[k] += ae * ed.v[o][k];
    }
    for (int k = 0; k < 3; ++k) {
      if (d) d[3 * i + k] = acc[k];
      lsum += log(acc[k]);
    }
  }
  const double t = block_sum(lsum, red);
  if (threadIdx.x == 0) logpart[blockIdx.x] = t;
}

__global__ __launch_bounds__(256) void scale_kernel(int nz, int ny, int nx, const double* a, DiagTable ed,
                                                    double alpha_value, const double* alpha_dev, const double* logpart,
                                                    int nblk, double* s, int soa) {
  // alpha from device memory (an on-device selection) when given
  const double alpha = alpha_dev ? *alpha_dev : alpha_value;
  __shared__ double smean;
  const int64_t N = int64_t(nx) * ny * nz;
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (threadIdx.x < 64) {
    const double tot = wave_reduce_partials(logpart, nblk, 1);
    if (threadIdx.x == 0) smean = tot / double(3 * N);
  }
  __syncthreads();
  const double mean = smean;
  if (i >= N) return;
  const int x = int(i % nx), y = int((i / nx) % ny), z = int(i / (int64_t(nx) * ny));
  double acc[3] = {0.0, 0.0, 0.0};
  for (int o = 0; o < 8; ++o) {
    const int ex = wrapi(x - (o & 1), nx), ey = wrapi(y - ((o >> 1) & 1), ny), ez = wrapi(z - ((o >> 2) & 1), nz);
    const double ae = a[(int64_t(ez) * ny + ey) * nx + ex];
    for (int k = 0; k < 3; ++k) acc[k] += ae * ed.v[o][k];
  }
  for (int k = 0; k < 3; ++k) s[lidx(soa, 0, k, i, N)] = exp(-0.5 * alpha * (log(acc[k]) - mean));
}


// ------------------------------------------------------------------ launchers

bool cubic(const pgj_grid& g) { return g.hx == g.hy && g.hy == g.hz; }

ElementCoef coef_of(const pgj_grid& g) { return make_coef(g.hx, g.hy, g.hz, g.lam, g.mu); }

DiagTable diag_table(const pgj_grid& g) {
  DiagTable t{};
  ElementCoef c = coef_of(g);
  for (int k = 0; k < 24; ++k) {
    double u[24] = {0}, f[24];
    u[k] = 1.0;
    hex8_element<false>(u, 1.0, c, f);
    t.v[k / 3][k % 3] = f[k];
  }
  return t;
}

int check_grid(const pgj_grid* g) {
  PGJ_REQUIRE(g != nullptr, "null grid");
  PGJ_REQUIRE(g->nx >= 2 && g->ny >= 2 && g->nz >= 2, "grid dimensions must be >= 2, got (%d,%d,%d)", g->nz, g->ny,
              g->nx);
  PGJ_REQUIRE(g->hx > 0 && g->hy > 0 && g->hz > 0, "positive voxel sizes required");
  PGJ_REQUIRE(int64_t(g->nx) * g->ny * g->nz < (int64_t(1) << 31) / 3, "grid too large");
  return PGJ_OK;
}

template <int MODE, int XPL, int R, int MAXT, bool PLANE, typename T>
int launch_rowmarch_t(const pgj_grid& g, int L, RowArgsT<T>& A, hipStream_t s) {
  const RowMarch& c = A.cfg;
  dim3 grid(c.by, c.bz, L), block(64 * (c.W + 1));
  const size_t lds = 2 * size_t(c.W + 1) * 64 * (PLANE ? 3 : 6) * XPL * sizeof(T) + 64 * 8;
  if (cubic(g))
    hipLaunchKernelGGL((hex8_rowmarch<MODE, XPL, R, true, MAXT, T, PLANE>), grid, block, lds, s, A);
  else
    hipLaunchKernelGGL((hex8_rowmarch<MODE, XPL, R, false, MAXT, T, PLANE>), grid, block, lds, s, A);
  PGJ_LAUNCHED("hex8_rowmarch");
  return PGJ_OK;
}

template <int MODE, typename T>
int launch_rowmarch_m(const pgj_grid& g, int L, RowArgsT<T>& A, hipStream_t s) {
  const RowMarch& c = A.cfg;
  constexpr int kPlaneThreads = sizeof(T) == 8 ? 320 : 576;
  if (c.plane) {
    if (c.xpl == 1) return launch_rowmarch_t<MODE, 1, 1, kPlaneThreads, true>(g, L, A, s);
    return launch_rowmarch_t<MODE, 2, 1, 320, true>(g, L, A, s);
  }
  if (c.R == 2) return launch_rowmarch_t<MODE, 1, 2, 320, false>(g, L, A, s);
  if (c.xpl == 1) return launch_rowmarch_t<MODE, 1, 1, 576, false>(g, L, A, s);
  return launch_rowmarch_t<MODE, 2, 1, 320, false>(g, L, A, s);
}

template <typename T>
int launch_rowmarch(const pgj_grid& g, int L, int mode, RowArgsT<T>& A, hipStream_t s) {
  A.nz = g.nz; A.ny = g.ny; A.nx = g.nx;
  A.N = int64_t(g.nx) * g.ny * g.nz;
  A.cfg = rowmarch_config(g, L, int(sizeof(T)));
  A.coef = coef_of(g);
  return mode == kModeApply ? launch_rowmarch_m<kModeApply>(g, L, A, s) : launch_rowmarch_m<kModePCG>(g, L, A, s);
}
template int launch_rowmarch<double>(const pgj_grid&, int, int, RowArgsT<double>&, hipStream_t);
template int launch_rowmarch<float>(const pgj_grid&, int, int, RowArgsT<float>&, hipStream_t);

int launch_reference(const pgj_grid& g, int L, RefArgs& A, bool cert, hipStream_t s) {
  A.nz = g.nz; A.ny = g.ny; A.nx = g.nx;
  A.N = int64_t(g.nx) * g.ny * g.nz;
  A.coef = coef_of(g);
  dim3 grid(unsigned((A.N + 255) / 256), L), block(256);
  const bool cb = cubic(g);
  if (cert) {
    if (cb) hipLaunchKernelGGL((hex8_reference<true, true>), grid, block, 0, s, A);
    else hipLaunchKernelGGL((hex8_reference<false, true>), grid, block, 0, s, A);
  } else {
    if (cb) hipLaunchKernelGGL((hex8_reference<true, false>), grid, block, 0, s, A);
    else hipLaunchKernelGGL((hex8_reference<false, false>), grid, block, 0, s, A);
  }
  PGJ_LAUNCHED("hex8_reference");
  return PGJ_OK;
}

int launch_subtract_mean(const pgj_grid& g, int L, double* v, const double* part, int soa, hipStream_t s) {
  const int64_t N = int64_t(g.nx) * g.ny * g.nz;
  const int nblk = int((N + 255) / 256);
  hipLaunchKernelGGL(subtract_mean, dim3(nblk, L), dim3(256), 0, s, v, N, part, nblk, soa);
  PGJ_LAUNCHED("subtract_mean");
  return PGJ_OK;
}

int launch_component_sums(const pgj_grid& g, int L, const double* v, double* part, int soa, hipStream_t s) {
  const int64_t N = int64_t(g.nx) * g.ny * g.nz;
  dim3 grid(unsigned((N + 255) / 256), L);
  hipLaunchKernelGGL(component_sums, grid, dim3(256), 0, s, v, N, part, soa);
  PGJ_LAUNCHED("component_sums");
  return PGJ_OK;
}

}  // namespace pgj

using namespace pgj;

extern "C" int pgj_hex8_apply(const pgj_grid* g, int L, const double* a, const double* u, const uint8_t* active,
                              double* out, void* stream) {
  if (int e = check_grid(g)) return e;
  PGJ_REQUIRE(L >= 1 && L <= 65535 && a && u && out, "invalid apply arguments");
  if (!rowmarch_config(*g, L).valid) {
    PGJ_REQUIRE(active == nullptr, "active mask needs the row-march kernel (nx <= 128 and nx %% ceil(nx/64) == 0)");
    return pgj_hex8_apply_reference(g, L, a, u, out, stream);
  }
  RowArgs A{};
  A.a = a; A.u = u; A.out = out; A.mask = active;
  return launch_rowmarch(*g, L, kModeApply, A, as_stream(stream));
}

extern "C" int pgj_hex8_apply32(const pgj_grid* g, int L, const float* a, const float* u, const uint8_t* active,
                                float* out, void* stream) {
  if (int e = check_grid(g)) return e;
  PGJ_REQUIRE(L >= 1 && L <= 65535 && a && u && out, "invalid apply arguments");
  PGJ_REQUIRE(rowmarch_config(*g, L, 4).valid, "the FP32 apply needs the row-march kernel (nx <= 128 and nx %% ceil(nx/64) == 0)");
  RowArgsT<float> A{};
  A.a = a; A.u = u; A.out = out; A.mask = active;
  return launch_rowmarch(*g, L, kModeApply, A, as_stream(stream));
}

extern "C" int pgj_hex8_apply_reference(const pgj_grid* g, int L, const double* a, const double* u, double* out,
                                        void* stream) {
  if (int e = check_grid(g)) return e;
  PGJ_REQUIRE(L >= 1 && L <= 65535 && a && u && out, "invalid apply arguments");
  RefArgs A{};
  A.a = a; A.u = u; A.out = out;
  return launch_reference(*g, L, A, false, as_stream(stream));
}

extern "C" int pgj_hex8_rhs(const pgj_grid* g, const double* a, const double macro[6], double* out, void* stream) {
  if (int e = check_grid(g)) return e;
  PGJ_REQUIRE(a && macro && out, "invalid rhs arguments");
  const int64_t N = int64_t(g->nx) * g->ny * g->nz;
  hipLaunchKernelGGL(hex8_rhs_kernel, dim3(unsigned((N + 255) / 256)), dim3(256), 0, as_stream(stream), g->nz, g->ny,
                     g->nx, coef_of(*g), a, macro[0], macro[1], macro[2], macro[3], macro[4], macro[5], out, 0);
  PGJ_LAUNCHED("hex8_rhs");
  return PGJ_OK;
}

namespace pgj {

int launch_rhs(const pgj_grid& g, const double* a, const double* macro, double* out, int soa, hipStream_t s) {
  const int64_t N = int64_t(g.nx) * g.ny * g.nz;
  hipLaunchKernelGGL(hex8_rhs_kernel, dim3(unsigned((N + 255) / 256)), dim3(256), 0, s, g.nz, g.ny, g.nx, coef_of(g),
                     a, macro[0], macro[1], macro[2], macro[3], macro[4], macro[5], out, soa);
  PGJ_LAUNCHED("hex8_rhs");
  return PGJ_OK;
}

int launch_scale(const pgj_grid& g, const double* a, double alpha, double* d, double* s_out, double* ws, int soa,
                 hipStream_t s, const double* alpha_dev) {
  const int64_t N = int64_t(g.nx) * g.ny * g.nz;
  const int nblk = int((N + 255) / 256);
  DiagTable t = diag_table(g);
  hipLaunchKernelGGL(diagonal_kernel, dim3(nblk), dim3(256), 0, s, g.nz, g.ny, g.nx, a, t, d, ws);
  PGJ_LAUNCHED("diagonal_kernel");
  hipLaunchKernelGGL(scale_kernel, dim3(nblk), dim3(256), 0, s, g.nz, g.ny, g.nx, a, t, alpha, alpha_dev, ws, nblk,
                     s_out, soa);
  PGJ_LAUNCHED("scale_kernel");
  return PGJ_OK;
}

}  // namespace pgj

extern "C" int pgj_hex8_energy(const pgj_grid* g, int L, const double* u, const double* macros, double* out,
                               void* stream) {
  if (int e = check_grid(g)) return e;
  PGJ_REQUIRE(L >= 1 && L <= 6 && u && macros && out, "energy needs 1 <= L <= 6 loads");
  MacroSet ms{};
  for (int l = 0; l < L; ++l)
    for (int k = 0; k < 6; ++k) ms.m[l][k] = macros[6 * l + k];
  const int64_t N = int64_t(g->nx) * g->ny * g->nz;
  dim3 grid(unsigned((N + 255) / 256), L);
  if (cubic(*g))
    hipLaunchKernelGGL(hex8_energy_kernel<true>, grid, dim3(256), 0, as_stream(stream), g->nz, g->ny, g->nx,
                       coef_of(*g), u, ms, out);
  else
    hipLaunchKernelGGL(hex8_energy_kernel<false>, grid, dim3(256), 0, as_stream(stream), g->nz, g->ny, g->nx,
                       coef_of(*g), u, ms, out);
  PGJ_LAUNCHED("hex8_energy");
  return PGJ_OK;
}

extern "C" size_t pgj_effective_tensor_ws_bytes(const pgj_grid* g) {
  const int64_t N = int64_t(g->nx) * g->ny * g->nz;
  return align256(size_t((N + 255) / 256) * 21 * 8);
}

extern "C" int pgj_effective_tensor(const pgj_grid* g, const double* a, const double* u6, double* ws, double* out36,
                                    void* stream) {
  if (int e = check_grid(g)) return e;
  PGJ_REQUIRE(a && u6 && ws && out36, "invalid effective tensor arguments");
  const int64_t N = int64_t(g->nx) * g->ny * g->nz;
  const int nblk = int((N + 255) / 256);
  hipStream_t s = as_stream(stream);
  if (cubic(*g))
    hipLaunchKernelGGL(effective_partials<true>, dim3(nblk), dim3(256), 0, s, g->nz, g->ny, g->nx, coef_of(*g), a, u6, ws);
  else
    hipLaunchKernelGGL(effective_partials<false>, dim3(nblk), dim3(256), 0, s, g->nz, g->ny, g->nx, coef_of(*g), a, u6,
                       ws);
  PGJ_LAUNCHED("effective_partials");
  hipLaunchKernelGGL(effective_finish, dim3(21), dim3(64), 0, s, ws, nblk, N, out36);
  PGJ_LAUNCHED("effective_finish");
  return PGJ_OK;
}

// ---- C^H without a caller workspace -----------------------------------------------------------------
// The partial sums live in a file-scope device array, one copy per GPU. Host code never passes the array's
// name to a kernel: in host code it is only a stand-in for the symbol. Its device address is looked up once
// per device (hipGetSymbolAddress, with that device current) and kept in `g_scratch_ptr`.
static constexpr int64_t kNoAllocMaxElements = int64_t(1) << 22;                 // 4,194,304 elements
static constexpr size_t kNoAllocScratchDoubles = size_t(kNoAllocMaxElements / 256) * 21;  // 2,752,512 bytes
static __device__ __align__(256) double scratch[kNoAllocScratchDoubles];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices] = {};

int scratch_for_current_device(double** out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) {
    set_error("hipGetDevice failed: %s", hipGetErrorString(e));
    return PGJ_ERR_HIP;
  }
  if (dev < 0 || dev >= kMaxDevices) {
    set_error("device %d beyond the %d devices the scratch table holds", dev, kMaxDevices);
    return PGJ_ERR_UNSUPPORTED;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      set_error("hipGetSymbolAddress(scratch) failed: %s", hipGetErrorString(e));
      return PGJ_ERR_HIP;
    }
    g_scratch_ptr[dev] = static_cast<double*>(p);
  }
  *out = g_scratch_ptr[dev];
  return PGJ_OK;
}
}  // namespace

extern "C" int64_t pgj_effective_tensor_noalloc_max_elements(void) { return kNoAllocMaxElements; }

extern "C" int pgj_effective_tensor_noalloc(const pgj_grid* g, const double* a, const double* u6, double* out36,
                                            void* stream) {
  if (int e = check_grid(g)) return e;
  PGJ_REQUIRE(a && u6 && out36, "invalid effective tensor arguments");
  const int64_t N = int64_t(g->nx) * g->ny * g->nz;
  PGJ_REQUIRE(N <= kNoAllocMaxElements,
              "pgj_effective_tensor_noalloc: grid of %lld elements exceeds the %lld-element limit of its "
              "built-in scratch; use pgj_effective_tensor with a workspace", (long long)N,
              (long long)kNoAllocMaxElements);
  double* part = nullptr;
  if (int e = scratch_for_current_device(&part)) return e;
  // every block overwrites its own 21 partials before effective_finish reads them: nothing to clear
  const int nblk = int((N + 255) / 256);
  hipStream_t s = as_stream(stream);
  if (cubic(*g))
    hipLaunchKernelGGL(effective_partials<true>, dim3(nblk), dim3(256), 0, s, g->nz, g->ny, g->nx, coef_of(*g), a,
                       u6, part);
  else
    hipLaunchKernelGGL(effective_partials<false>, dim3(nblk), dim3(256), 0, s, g->nz, g->ny, g->nx, coef_of(*g), a,
                       u6, part);
  PGJ_LAUNCHED("effective_partials");
  hipLaunchKernelGGL(effective_finish, dim3(21), dim3(64), 0, s, part, nblk, N, out36);
  PGJ_LAUNCHED("effective_finish");
  return PGJ_OK;
}

extern "C" size_t pgj_avg_stress_ws_bytes(const pgj_grid* g) {
  const int64_t N = int64_t(g->nx) * g->ny * g->nz;
  return align256(size_t((N + 255) / 256) * 6 * 8);
}

extern "C" int pgj_hex8_avg_stress(const pgj_grid* g, const double* a, const double* u, const double macro[6],
                                   double* ws, double* out6, void* stream) {
  if (int e = check_grid(g)) return e;
  PGJ_REQUIRE(a && u && macro && ws && out6, "invalid stress arguments");
  const int64_t N = int64_t(g->nx) * g->ny * g->nz;
  const int nblk = int((N + 255) / 256);
  MacroSet ms{};
  for (int k = 0; k < 6; ++k) ms.m[0][k] = macro[k];
  hipStream_t s = as_stream(stream);
  hipLaunchKernelGGL(avg_stress_partials, dim3(nblk), dim3(256), 0, s, g->nz, g->ny, g->nx, coef_of(*g), a, u, ms, ws);
  PGJ_LAUNCHED("avg_stress_partials");
  hipLaunchKernelGGL(avg_stress_finish, dim3(6), dim3(64), 0, s, ws, nblk, N, out6);
  PGJ_LAUNCHED("avg_stress_finish");
  return PGJ_OK;
}

extern "C" size_t pgj_scale_ws_bytes(const pgj_grid* g) {
  const int64_t N = int64_t(g->nx) * g->ny * g->nz;
  return align256(size_t((N + 255) / 256) * 8);
}

extern "C" int pgj_diagonal_scale(const pgj_grid* g, const double* a, double alpha, double* d, double* s_out,
                                  double* ws, void* stream) {
  if (int e = check_grid(g)) return e;
  PGJ_REQUIRE(a && s_out && ws, "invalid scale arguments");
  return launch_scale(*g, a, alpha, d, s_out, ws, 0, as_stream(stream));
}

extern "C" size_t pgj_project_ws_bytes(const pgj_grid* g, int L) {
  const int64_t N = int64_t(g->nx) * g->ny * g->nz;
  return align256(size_t(L) * size_t((N + 255) / 256) * 4 * 8);
}

extern "C" int pgj_project(const pgj_grid* g, int L, double* v, double* ws, void* stream) {
  if (int e = check_grid(g)) return e;
  PGJ_REQUIRE(L >= 1 && L <= 65535 && v && ws, "invalid projection arguments");
  const int64_t N = int64_t(g->nx) * g->ny * g->nz;
  const int nblk = int((N + 255) / 256);
  hipStream_t s = as_stream(stream);
  (void)nblk;
  if (int e = launch_component_sums(*g, L, v, ws, 0, s)) return e;
  return launch_subtract_mean(*g, L, v, ws, 0, s);
}

extern "C" int pgj_aos_to_soa(const pgj_grid* g, int L, const double* in, double* out, void* stream) {
  if (int e = check_grid(g)) return e;
  PGJ_REQUIRE(L >= 1 && L <= 65535 && in && out && in != out, "invalid transpose arguments");
  const int64_t N = int64_t(g->nx) * g->ny * g->nz;
  hipLaunchKernelGGL(transpose_kernel, dim3(unsigned((N + 255) / 256), L), dim3(256), 0, as_stream(stream), in, out, N, 1);
  PGJ_LAUNCHED("transpose");
  return PGJ_OK;
}

extern "C" int pgj_soa_to_aos(const pgj_grid* g, int L, const double* in, double* out, void* stream) {
  if (int e = check_grid(g)) return e;
  PGJ_REQUIRE(L >= 1 && L <= 65535 && in && out && in != out, "invalid transpose arguments");
  const int64_t N = int64_t(g->nx) * g->ny * g->nz;
  hipLaunchKernelGGL(transpose_kernel, dim3(unsigned((N + 255) / 256), L), dim3(256), 0, as_stream(stream), in, out, N, 0);
  PGJ_LAUNCHED("transpose");
  return PGJ_OK;
}
