// Hand-written FP64 3-D R2C/C2R FFT (LDS Stockham, radices 4/2/3/5/7) with
// the FI Green symbol fused into the z pass, and the pre/post scaling,
// projection and PCG vector updates fused into the xy passes.
//
// Data path per load (spectrum layout S[l][c][kz][ky][kx], kx < nx/2+1):
//   fxy  : per z-plane, v = s*r (PCG: r -= alpha q first);
//          x-R2C (two reals per complex line + split), y-FFT; write S.
//   fz   : per (ky, kx-chunk), z-FFT, v^ -= mean(r) s^ (projection P),
//          w^ = G v^, Parseval partials of v.w and s.w, inverse z-FFT.
//   fxyi : per z-plane, inverse y-FFT, x-C2R, z = s*w - mean(s*w)
//          (PCG: written as the next direction z + beta p; u += alpha p).
#include <cmath>

#include "fft_kernels.h"
#include "stockham.h"

namespace pgj {

constexpr int kMaxV = kStageValues;

// Placement of the lines of a batched transform in LDS: line l starts at
// (l / per) * outer + (l % per) * inner and its elements are es apart.
struct LineMap {
  int nlines, per, outer, inner, es;
  bool line_fast;  // consecutive threads take consecutive lines (else consecutive butterflies)
  __device__ __forceinline__ int base(int l) const { return (l / per) * outer + (l % per) * inner; }
};

// One in-place Stockham stage over the lines of `map`, each of length n.
// All threads of the block take part; two barriers separate the read and
// write phases, so every value of the stage is held in registers between them.
template <int R, bool INV>
__device__ void lds_stage(double2* buf, const LineMap& map, int n, int Ns, const double2* tw, int step) {
  const int nlines = map.nlines, es = map.es;
  const bool line_fast = map.line_fast;
  constexpr int NB = kMaxV / R;
  double2 v[NB][R];
  const int nb = n / R, count = nlines * nb, m = n / (Ns * R);
#pragma unroll
  for (int i = 0; i < NB; ++i) {
    const int b = threadIdx.x + i * blockDim.x;
    if (b < count) {
      const int line = line_fast ? b % nlines : b / nb;
      const int j = line_fast ? b / nlines : b % nb;
      const int k = j % Ns;
      const double2* L = buf + map.base(line);
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const double2 x = L[(j + r * nb) * es];
        v[i][r] = r == 0 ? x : cmul(x, twiddle<INV>(tw, (k * r * m) % n, step));
      }
    }
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < NB; ++i) {
    const int b = threadIdx.x + i * blockDim.x;
    if (b < count) {
      const int line = line_fast ? b % nlines : b / nb;
      const int j = line_fast ? b / nlines : b % nb;
      const int k = j % Ns;
      dft<R, INV>(v[i], tw, n, step);
      double2* L = buf + map.base(line);
      const int idx = (j / Ns) * Ns * R + k;
#pragma unroll
      for (int r = 0; r < R; ++r) L[(idx + r * Ns) * es] = v[i][r];
    }
  }
  __syncthreads();
}

template <bool INV>
__device__ void lds_fft(double2* buf, const LineMap& map, const Factors& f, const double2* tw, int step) {
  int Ns = 1;
  for (int s = 0; s < f.count; ++s) {
    const int R = f.r[s];
    switch (R) {
      case 4: lds_stage<4, INV>(buf, map, f.n, Ns, tw, step); break;
      case 2: lds_stage<2, INV>(buf, map, f.n, Ns, tw, step); break;
      case 3: lds_stage<3, INV>(buf, map, f.n, Ns, tw, step); break;
      case 5: lds_stage<5, INV>(buf, map, f.n, Ns, tw, step); break;
      default: lds_stage<7, INV>(buf, map, f.n, Ns, tw, step); break;
    }
    Ns *= R;
  }
}

__device__ __forceinline__ LineMap rows_map(int rows, int rs) { return LineMap{rows, rows, 0, rs, 1, false}; }
__device__ __forceinline__ LineMap cols_map(int comps, int ny, int nkx, int rs) {
  return LineMap{comps * nkx, nkx, ny * rs, 1, rs, true};
}

__device__ __forceinline__ double reduce_pr_sum(const FFTArgs& A, int l, int c) {
  // sum over planes of the per-plane component sums (fixed order, every lane identical)
  return wave_reduce_partials(A.pr + (int64_t(l) * A.nz * 3 + c) * 2 + 1, A.nz, 6);
}

// --------------------------------------------------------------- forward xy
template <int MODE>
__global__ __launch_bounds__(512) void fxy_kernel(FFTArgs A) {
  extern __shared__ double2 plane[];
  const FFTGeom& f = A.f;
  const int z = blockIdx.x;
  const int comps = f.comps;
  const int c0 = comps == 3 ? 0 : blockIdx.y;
  const int l = blockIdx.z;
  const int nx = A.nx, ny = A.ny;
  const int64_t plane_nodes = int64_t(nx) * ny;
  double* red = reinterpret_cast<double*>(plane + size_t(comps) * ny * f.rs);

  double alpha = 0.0;
  if (MODE == kFxyPCG) {
    if (!A.st.active[l]) return;
    double pq, rz;
    wave_reduce_partials_2(A.ppq + int64_t(l) * A.ppq_stride, A.ppq_count, 1, A.pg + int64_t(l) * f.nkb * ny * 4,
                           f.nkb * ny, 4, pq, rz);
    const bool valid = isfinite(pq) && pq > 0.0 && isfinite(rz) && rz > 0.0;
    alpha = valid ? rz / pq : 0.0;
    if (threadIdx.x == 0 && z == 0 && blockIdx.y == 0) {
      A.st.rz[l] = rz;
      A.st.alpha[l] = alpha;
      A.st.valid[l] = valid;
      A.st.first[l] = 0;
    }
  }

  // 1. load (and update) the plane; pack reals two per complex along x.
  double sq[3] = {0.0, 0.0, 0.0}, sm[3] = {0.0, 0.0, 0.0};
  const bool even = (nx % 2) == 0;
  const int64_t count = comps == 3 ? 3 * plane_nodes : plane_nodes;
  for (int64_t e = threadIdx.x; e < count; e += blockDim.x) {
    const int64_t node = comps == 3 ? e / 3 : e;
    const int c = comps == 3 ? int(e % 3) : c0;
    const int64_t gnode = int64_t(z) * plane_nodes + node;
    const int64_t idx = vindex(A, l, c, gnode);
    const int64_t sidx = vindex(A, design_of(l, A.group), c, gnode);
    double val;
    if (MODE == kFxyScale) {
      val = A.s[sidx];
    } else if (MODE == kFxyPre) {
      const double rv = A.rin[idx];
      sq[c] = fma(rv, rv, sq[c]);
      sm[c] += rv;
      val = A.s ? A.s[sidx] * rv : rv;
    } else {
      const double rn = fma(-alpha, A.q[idx], A.r[idx]);
      A.r[idx] = rn;
      sq[c] = fma(rn, rn, sq[c]);
      sm[c] += rn;
      val = A.s ? A.s[sidx] * rn : rn;
    }
    const int x = int(node % nx), y = int(node / nx);
    double2* row = plane + (size_t(comps == 3 ? c : 0) * ny + y) * f.rs;
    if (even) reinterpret_cast<double*>(row)[x] = val;
    else row[x] = c2(val, 0.0);
  }
  if (MODE != kFxyScale) {
    for (int c = 0; c < 3; ++c) {
      if (comps == 1 && c != c0) continue;
      const double tq = block_sum(sq[c], red), tm = block_sum(sm[c], red);
      if (threadIdx.x == 0) {
        double* p = A.pr + ((int64_t(l) * A.nz + z) * 3 + c) * 2;
        p[0] = tq;
        p[1] = tm;
      }
    }
  }
  __syncthreads();
  // 2. x transform on every row (complex length M).
  const int rows = comps * ny;
  lds_fft<false>(plane, rows_map(rows, f.rs), f.fx, A.twx, nx / f.M);
  // 3. split the packed spectrum into the half spectrum of each real row.
  if (even) {
    const int M = f.M, half = M / 2;
    const int pairs = rows * (half + 1);
    for (int b = threadIdx.x; b < pairs; b += blockDim.x) {
      const int row = b / (half + 1), k = b % (half + 1);
      double2* L = plane + size_t(row) * f.rs;
      const double2 zk = L[k % M], zm = L[(M - k) % M];
      const double2 E = c2(0.5 * (zk.x + zm.x), 0.5 * (zk.y - zm.y));
      const double2 O = c2(0.5 * (zk.y + zm.y), -0.5 * (zk.x - zm.x));
      const double2 w = A.twx[k];
      const double2 wo = cmul(w, O);
      L[k] = cadd(E, wo);
      if (M - k != k) L[M - k] = conjd(csub(E, wo));
    }
    __syncthreads();
  }
  // 4. y transform on every kx column.
  lds_fft<false>(plane, cols_map(comps, ny, f.nkx, f.rs), f.fy, A.twy, 1);
  // 5. write the half spectrum.
  double2* out = A.spec + ((int64_t(l) * 3 + c0) * A.nz + z) * int64_t(ny) * f.nkx;
  const int64_t total = int64_t(comps) * ny * f.nkx;
  for (int64_t e = threadIdx.x; e < total; e += blockDim.x) {
    const int c = int(e / (int64_t(ny) * f.nkx));
    const int rem = int(e % (int64_t(ny) * f.nkx));
    const int y = rem / f.nkx, kx = rem % f.nkx;
    out[int64_t(c) * A.nz * ny * f.nkx + rem] = plane[(size_t(c) * ny + y) * f.rs + kx];
  }
}

// ---------------------------------------------------------------- z + Green
template <int MODE>
__global__ __launch_bounds__(256) void fz_kernel(FFTArgs A) {
  extern __shared__ double2 cols[];
  const FFTGeom& f = A.f;
  const int kb = blockIdx.x, y = blockIdx.y, l = blockIdx.z;
  const int nz = A.nz, ny = A.ny, nkx = f.nkx, kc = f.kc;
  const int kx0 = kb * kc;
  const int kcn = min(kc, nkx - kx0);
  const int blk = y * f.nkb + kb;
  const bool lead = threadIdx.x == 0 && kb == 0 && y == 0;
  double* red = reinterpret_cast<double*>(cols + size_t(kc) * 3 * nz);
  double* pg = A.pg + (int64_t(l) * f.nkb * ny + blk) * 4;

  double mean[3] = {0.0, 0.0, 0.0};
  if (MODE == kFzInit || MODE == kFzPCG) {
    // one pass over the per-plane partials (|r|^2 and the component sums of r)
    const int lane = threadIdx.x & 63;
    double v = 0.0, vm[3] = {0.0, 0.0, 0.0};
    for (int i = lane; i < A.nz; i += 64) {
      const double* p = A.pr + (int64_t(l) * A.nz + i) * 6;
      v += (p[0] + p[2]) + p[4];
      vm[0] += p[1];
      vm[1] += p[3];
      vm[2] += p[5];
    }
    const double res2 = wave_sum(v);
    if (A.shat)
      for (int c = 0; c < 3; ++c) mean[c] = wave_sum(vm[c]) / double(A.N);
    bool act;
    if (MODE == kFzInit) {
      act = res2 > A.st.target2[l] && A.st.it[l] < A.st.cap[l] && !A.st.bad[l];
      if (lead) {
        A.st.active_next[l] = act;
        A.st.it_next[l] = A.st.it[l];
        A.st.bad_next[l] = A.st.bad[l];
        A.st.first[l] = 1;
        A.st.stopped[l] = 0;
      }
    } else {
      const bool active = A.st.active[l] != 0;
      const bool valid = A.st.valid[l] != 0;  // set by the forward pass of this iteration
      const bool finite = isfinite(res2);
      const int64_t it = A.st.it[l] + (active ? 1 : 0);
      act = active && valid && finite && res2 > A.st.target2[l] && it < A.st.cap[l];
      if (lead) {
        A.st.active_next[l] = act;
        A.st.it_next[l] = it;
        A.st.bad_next[l] = A.st.bad[l] || (active && !valid) || (active && !finite);
        A.st.stopped[l] = active && !act;
      }
    }
    if (!act) return;
  } else if (MODE != kFzScale && A.shat) {
    for (int c = 0; c < 3; ++c) mean[c] = reduce_pr_sum(A, l, c) / double(A.N);
  }
  // load the columns (c, z, kcol) of this (ky, kx-chunk)
  double2* S = A.spec + int64_t(l) * 3 * f.nf;
  const int total = 3 * nz * kcn;
  for (int e = threadIdx.x; e < total; e += blockDim.x) {
    const int kcol = e % kcn, cz = e / kcn;
    const int c = cz / nz, zz = cz % nz;
    cols[(size_t(c) * nz + zz) * kc + kcol] = S[((int64_t(c) * nz + zz) * ny + y) * nkx + kx0 + kcol];
  }
  __syncthreads();
  const LineMap zmap{3 * kcn, kcn, nz * kc, 1, kc, true};
  lds_fft<false>(cols, zmap, f.fz, A.twz, 1);
  if (MODE == kFzScale) {
    for (int e = threadIdx.x; e < total; e += blockDim.x) {
      const int kcol = e % kcn, cz = e / kcn;
      const int c = cz / nz, zz = cz % nz;
      S[((int64_t(c) * nz + zz) * ny + y) * nkx + kx0 + kcol] = cols[(size_t(c) * nz + zz) * kc + kcol];
    }
    return;
  }
  // per frequency: projection correction, Green multiply, Parseval partials
  double rz = 0.0, sw[3] = {0.0, 0.0, 0.0};
  for (int e = threadIdx.x; e < nz * kcn; e += blockDim.x) {
    const int kcol = e % kcn, zz = e / kcn;
    const int kx = kx0 + kcol;
    const int64_t fidx = (int64_t(zz) * ny + y) * nkx + kx;
    double2 v[3];
    for (int c = 0; c < 3; ++c) v[c] = cols[(size_t(c) * nz + zz) * kc + kcol];
    double2 sh[3];
    if (A.shat) {
      for (int c = 0; c < 3; ++c) {
        sh[c] = A.shat[(int64_t(design_of(l, A.group)) * 3 + c) * f.nf + fidx];
        v[c] = c2(fma(-mean[c], sh[c].x, v[c].x), fma(-mean[c], sh[c].y, v[c].y));
      }
    }
    const double g0 = A.sym[fidx], g1 = A.sym[f.nf + fidx], g2 = A.sym[2 * f.nf + fidx];
    const double g3 = A.sym[3 * f.nf + fidx], g4 = A.sym[4 * f.nf + fidx], g5 = A.sym[5 * f.nf + fidx];
    double2 w[3];
    w[0] = c2(g0 * v[0].x + g3 * v[1].x + g4 * v[2].x, g0 * v[0].y + g3 * v[1].y + g4 * v[2].y);
    w[1] = c2(g3 * v[0].x + g1 * v[1].x + g5 * v[2].x, g3 * v[0].y + g1 * v[1].y + g5 * v[2].y);
    w[2] = c2(g4 * v[0].x + g5 * v[1].x + g2 * v[2].x, g4 * v[0].y + g5 * v[1].y + g2 * v[2].y);
    const double wgt = (kx == 0 || (A.nx % 2 == 0 && kx == A.nx / 2)) ? 1.0 : 2.0;
    double t = 0.0;
    for (int c = 0; c < 3; ++c) t += v[c].x * w[c].x + v[c].y * w[c].y;
    rz = fma(wgt, t, rz);
    if (A.shat)
      for (int c = 0; c < 3; ++c) sw[c] = fma(wgt, sh[c].x * w[c].x + sh[c].y * w[c].y, sw[c]);
    for (int c = 0; c < 3; ++c) cols[(size_t(c) * nz + zz) * kc + kcol] = w[c];
  }
  __syncthreads();
  lds_fft<true>(cols, zmap, f.fz, A.twz, 1);
  for (int e = threadIdx.x; e < total; e += blockDim.x) {
    const int kcol = e % kcn, cz = e / kcn;
    const int c = cz / nz, zz = cz % nz;
    S[((int64_t(c) * nz + zz) * ny + y) * nkx + kx0 + kcol] = cols[(size_t(c) * nz + zz) * kc + kcol];
  }
  const double trz = block_sum(rz, red);
  double tsw[3];
  for (int c = 0; c < 3; ++c) tsw[c] = block_sum(sw[c], red);
  if (threadIdx.x == 0) {
    pg[0] = trz;
    pg[1] = tsw[0];
    pg[2] = tsw[1];
    pg[3] = tsw[2];
  }
}

// --------------------------------------------------------------- inverse xy
template <int MODE>
__global__ __launch_bounds__(512) void fxyi_kernel(FFTArgs A) {
  extern __shared__ double2 plane[];
  const FFTGeom& f = A.f;
  const int z = blockIdx.x;
  const int comps = f.comps;
  const int c0 = comps == 3 ? 0 : blockIdx.y;
  const int l = blockIdx.z;
  const int nx = A.nx, ny = A.ny;
  const int64_t plane_nodes = int64_t(nx) * ny;
  const int64_t count = comps == 3 ? 3 * plane_nodes : plane_nodes;
  // PCG: the displacement update u += alpha p of this iteration (alpha from the forward pass)
  const bool move = MODE == kFxyiPCG && A.p != nullptr && A.u != nullptr;
  const double alpha = move ? A.st.alpha[l] : 0.0;
  if (MODE == kFxyiPCG) {
    if (threadIdx.x == 0 && z == 0 && blockIdx.y == 0) {
      A.st.active[l] = A.st.active_next[l];
      A.st.it[l] = A.st.it_next[l];
      A.st.bad[l] = A.st.bad_next[l];
    }
    if (!A.st.active_next[l]) {
      // a load stopped by this iteration's test still takes its last displacement update
      if (move && A.st.stopped[l]) {
        for (int64_t e = threadIdx.x; e < count; e += blockDim.x) {
          const int64_t node = comps == 3 ? e / 3 : e;
          const int c = comps == 3 ? int(e % 3) : c0;
          const int64_t oi = vindex(A, l, c, int64_t(z) * plane_nodes + node);
          A.u[oi] = fma(alpha, A.p[oi], A.u[oi]);
        }
      }
      return;
    }
  }
  // PCG: the output is the next direction p = z + beta p (z alone right after init)
  double mean[3] = {0.0, 0.0, 0.0};
  double beta = 0.0;
  const bool update = MODE == kFxyiPCG && A.p != nullptr && !A.st.first[l];
  if (A.shat || update) {
    // one pass: r.z (offset 0) and the s.w partials of the three components (offsets 1..3)
    const int cnt = f.nkb * ny;
    const int off[4] = {0, 1, 2, 3};
    double sums[4];
    wave_reduce_partials_n<4>(A.pg + int64_t(l) * cnt * 4, off, cnt, 4, sums);
    if (A.shat)
      for (int c = 0; c < 3; ++c) mean[c] = sums[1 + c] / double(A.N);
    if (update) {
      const double rz = A.st.rz[l];
      beta = sums[0] / (rz != 0.0 ? rz : 1.0);
    }
  }
  // 1. load the half spectrum of the plane
  const double2* in = A.spec + ((int64_t(l) * 3 + c0) * A.nz + z) * int64_t(ny) * f.nkx;
  const int64_t total = int64_t(comps) * ny * f.nkx;
  for (int64_t e = threadIdx.x; e < total; e += blockDim.x) {
    const int c = int(e / (int64_t(ny) * f.nkx));
    const int rem = int(e % (int64_t(ny) * f.nkx));
    const int y = rem / f.nkx, kx = rem % f.nkx;
    plane[(size_t(c) * ny + y) * f.rs + kx] = in[int64_t(c) * A.nz * ny * f.nkx + rem];
  }
  __syncthreads();
  // 2. inverse y transform
  lds_fft<true>(plane, cols_map(comps, ny, f.nkx, f.rs), f.fy, A.twy, 1);
  const int rows = comps * ny;
  const bool even = (nx % 2) == 0;
  // 3. rebuild the complex line of each real row
  if (even) {
    const int M = f.M, half = M / 2;
    const int pairs = rows * (half + 1);
    for (int b = threadIdx.x; b < pairs; b += blockDim.x) {
      const int row = b / (half + 1), k = b % (half + 1);
      double2* L = plane + size_t(row) * f.rs;
      double2 xk = L[k], xm = L[M - k];
      if (k == 0) {
        xk.y = 0.0;
        xm.y = 0.0;
      }
      // Z[k] = (X[k] + conj(X[M-k])) + i (X[k] - conj(X[M-k])) conj(W^k)
      const double2 wk = conjd(A.twx[k]);
      const double2 e1 = c2(xk.x + xm.x, xk.y - xm.y);
      const double2 o1 = cmul(c2(xk.x - xm.x, xk.y + xm.y), wk);
      const double2 zk = c2(e1.x - o1.y, e1.y + o1.x);
      if (k != 0 && M - k != k) {
        const double2 wm = conjd(A.twx[M - k]);
        const double2 e2 = c2(xm.x + xk.x, xm.y - xk.y);
        const double2 o2 = cmul(c2(xm.x - xk.x, xm.y + xk.y), wm);
        L[M - k] = c2(e2.x - o2.y, e2.y + o2.x);
      }
      L[k] = zk;
    }
  } else {
    // odd nx: Hermitian completion of every row
    const int fill = nx - f.nkx;
    for (int b = threadIdx.x; b < rows * fill; b += blockDim.x) {
      const int row = b / fill, k = f.nkx + b % fill;
      double2* L = plane + size_t(row) * f.rs;
      L[k] = conjd(L[nx - k]);
    }
    for (int b = threadIdx.x; b < rows; b += blockDim.x) plane[size_t(b) * f.rs].y = 0.0;
  }
  __syncthreads();
  // 4. inverse x transform
  lds_fft<true>(plane, rows_map(rows, f.rs), f.fx, A.twx, nx / f.M);
  // 5. unpack, scale and project; PCG: next direction and displacement update
  for (int64_t e = threadIdx.x; e < count; e += blockDim.x) {
    const int64_t node = comps == 3 ? e / 3 : e;
    const int c = comps == 3 ? int(e % 3) : c0;
    const int x = int(node % nx), y = int(node / nx);
    const double2* row = plane + (size_t(comps == 3 ? c : 0) * ny + y) * f.rs;
    const double w = even ? reinterpret_cast<const double*>(row)[x] : row[x].x;
    const int64_t gnode = int64_t(z) * plane_nodes + node;
    const double v = A.s ? A.s[vindex(A, design_of(l, A.group), c, gnode)] * w - mean[c] : w;
    const int64_t oi = vindex(A, l, c, gnode);
    A.zout[oi] = update ? fma(beta, A.p[oi], v) : v;
    if (move) A.u[oi] = fma(alpha, A.p[oi], A.u[oi]);
  }
}

// ----------------------------------------------------------------- plan init
__global__ void plan_twiddles(int nx, int ny, int nz, double2* twx, double2* twy, double2* twz) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  auto w = [](int k, int n) {
    double s, c;
    sincospi(2.0 * double(k) / double(n), &s, &c);
    return c2(c, -s);
  };
  if (i < nx) twx[i] = w(i, nx);
  if (i < ny) twy[i] = w(i, ny);
  if (i < nz) twz[i] = w(i, nz);
}

// FI Green factors at half-spectrum index k: D_ii = 1/(mu t sum(d+v^2) + (lam+mu) t d_i),
// beta = (lam+mu) t / (1 + (lam+mu) t sum v_i^2 D_ii), per-axis v_i; zero at k = 0.
struct FiFactors {
  double inv[3], v[3], beta;
};

__device__ FiFactors fi_factors(int64_t k, int nz, int ny, int nx, double hx, double hy, double hz, double lam,
                                double mu) {
  const int nkx = nx / 2 + 1;
  const int kx = int(k % nkx), ky = int((k / nkx) % ny), kz = int(k / (int64_t(nkx) * ny));
  const int idx[3] = {kx, ky, kz};
  const int n[3] = {nx, ny, nz};
  const double h[3] = {hx, hy, hz};
  FiFactors F;
  double m[3], d[3];
  double t = hx * hy * hz, trace = 0.0;
  for (int i = 0; i < 3; ++i) {
    double s, c;
    sincospi(double(idx[i]) / double(n[i]), &s, &c);
    m[i] = c * c + s * s / 3.0;
    t *= m[i];
    F.v[i] = 2.0 * s * c / (h[i] * m[i]);
    d[i] = 4.0 * (s * s) * (s * s) / (3.0 * h[i] * h[i] * m[i] * m[i]);
    trace += d[i] + F.v[i] * F.v[i];
  }
  trace *= t;
  const double b = (lam + mu) * t;
  double den = 1.0;
  for (int i = 0; i < 3; ++i) {
    F.inv[i] = k == 0 ? 0.0 : 1.0 / (mu * trace + (lam + mu) * t * d[i]);
    den += b * F.v[i] * F.v[i] * F.inv[i];
  }
  F.beta = k == 0 ? 0.0 : b / den;
  return F;
}

// Dense FI Green symbol: G = (D - beta (Dv)(Dv)^T)/N.
__global__ void plan_symbol(int nz, int ny, int nx, double hx, double hy, double hz, double lam, double mu,
                            double* sym) {
  const int64_t nf = int64_t(nz) * ny * (nx / 2 + 1);
  const int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (k >= nf) return;
  const FiFactors F = fi_factors(k, nz, ny, nx, hx, hy, hz, lam, mu);
  const double beta = F.beta;
  const double invN = 1.0 / (double(nx) * ny * nz);
  double w[3];
  for (int i = 0; i < 3; ++i) w[i] = F.inv[i] * F.v[i];
  sym[k] = (F.inv[0] - beta * w[0] * w[0]) * invN;
  sym[nf + k] = (F.inv[1] - beta * w[1] * w[1]) * invN;
  sym[2 * nf + k] = (F.inv[2] - beta * w[2] * w[2]) * invN;
  sym[3 * nf + k] = -beta * w[0] * w[1] * invN;
  sym[4 * nf + k] = -beta * w[0] * w[2] * invN;
  sym[5 * nf + k] = -beta * w[1] * w[2] * invN;
}

// The factored form (D_00, D_11, D_22, beta) per k, unnormalised, and the per-axis
// v tables (x over nx/2+1, then y, then z).
__global__ void fi_coeff_kernel(int nz, int ny, int nx, double hx, double hy, double hz, double lam, double mu,
                                double* out4, double* vtab) {
  const int nkx = nx / 2 + 1;
  const int64_t nf = int64_t(nz) * ny * nkx;
  const int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (k >= nf) return;
  const FiFactors F = fi_factors(k, nz, ny, nx, hx, hy, hz, lam, mu);
  for (int i = 0; i < 3; ++i) out4[i * nf + k] = F.inv[i];
  out4[3 * nf + k] = F.beta;
  if (vtab) {
    const int kx = int(k % nkx), ky = int((k / nkx) % ny), kz = int(k / (int64_t(nkx) * ny));
    if (ky == 0 && kz == 0) vtab[kx] = F.v[0];
    if (kx == 0 && kz == 0) vtab[nkx + ky] = F.v[1];
    if (kx == 0 && ky == 0) vtab[nkx + ny + kz] = F.v[2];
  }
}

// ------------------------------------------------------------------ launchers
FFTArgs fft_args(const pgj_grid& g, const void* plan) {
  FFTArgs A{};
  A.nz = g.nz; A.ny = g.ny; A.nx = g.nx;
  A.N = int64_t(g.nx) * g.ny * g.nz;
  A.f = fft_geom(g);
  PlanLayout pl = plan_layout(g);
  const char* b = static_cast<const char*>(plan);
  A.twx = reinterpret_cast<const double2*>(b + pl.tw_x);
  A.twy = reinterpret_cast<const double2*>(b + pl.tw_y);
  A.twz = reinterpret_cast<const double2*>(b + pl.tw_z);
  A.sym = reinterpret_cast<const double*>(b + pl.sym);
  return A;
}

int check_fft(const pgj_grid* g) {
  PGJ_REQUIRE(g != nullptr, "null grid");
  FFTGeom f = fft_geom(*g);
  PGJ_REQUIRE(f.valid, "unsupported FFT grid (%d,%d,%d): axis lengths must be 2..256 with factors 2,3,5,7",
              g->nz, g->ny, g->nx);
  return PGJ_OK;
}

// Kernels using more than 64 KiB of dynamic LDS must opt in once.
template <typename K>
int allow_lds(K kernel, size_t bytes) {
  if (bytes <= 65536) return PGJ_OK;
  hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(kernel),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, int(bytes));
  if (e != hipSuccess) {
    set_error("cannot reserve %zu bytes of LDS: %s", bytes, hipGetErrorString(e));
    return PGJ_ERR_HIP;
  }
  return PGJ_OK;
}

#define PGJ_RUNTIME_FFT(A)                                                                              \
  PGJ_REQUIRE((A).f.runtime_ok, "grid (%d,%d,%d) has only the compile-time resident FFT path", (A).nz, (A).ny, \
              (A).nx)

// With fewer plane blocks than CUs (few loads), the xy kernels take one
// component per block, which triples the grid; the spectrum layout is the same.
FFTArgs xy_args(const FFTArgs& A, int L) {
  FFTArgs B = A;
  if (A.f.comps == 3 && int64_t(A.nz) * L < 256) {
    B.f.comps = 1;
    B.f.lds_xy = (A.f.lds_xy - 64 * 8) / 3 + 64 * 8;
  }
  return B;
}

template <int MODE>
int launch_fxy(FFTArgs& A0, int L, hipStream_t s) {
  PGJ_RUNTIME_FFT(A0);
  FFTArgs A = xy_args(A0, L);
  const FFTGeom& f = A.f;
  dim3 grid(A.nz, f.comps == 3 ? 1 : 3, L);
  if (int e = allow_lds(fxy_kernel<MODE>, f.lds_xy)) return e;
  hipLaunchKernelGGL(fxy_kernel<MODE>, grid, dim3(f.threads_xy), f.lds_xy, s, A);
  PGJ_LAUNCHED("fxy_kernel");
  return PGJ_OK;
}

template <int MODE>
int launch_fz(FFTArgs& A, int L, hipStream_t s) {
  PGJ_RUNTIME_FFT(A);
  const FFTGeom& f = A.f;
  dim3 grid(f.nkb, A.ny, L);
  hipLaunchKernelGGL(fz_kernel<MODE>, grid, dim3(f.threads_z), f.lds_z, s, A);
  PGJ_LAUNCHED("fz_kernel");
  return PGJ_OK;
}

template <int MODE>
int launch_fxyi(FFTArgs& A0, int L, hipStream_t s) {
  PGJ_RUNTIME_FFT(A0);
  FFTArgs A = xy_args(A0, L);
  const FFTGeom& f = A.f;
  dim3 grid(A.nz, f.comps == 3 ? 1 : 3, L);
  if (int e = allow_lds(fxyi_kernel<MODE>, f.lds_xy)) return e;
  hipLaunchKernelGGL(fxyi_kernel<MODE>, grid, dim3(f.threads_xy), f.lds_xy, s, A);
  PGJ_LAUNCHED("fxyi_kernel");
  return PGJ_OK;
}

int launch_fxy_mode(FFTArgs& A, int L, int mode, hipStream_t s) {
  switch (mode) {
    case kFxyScale: return launch_fxy<kFxyScale>(A, L, s);
    case kFxyPre: return launch_fxy<kFxyPre>(A, L, s);
    default: return launch_fxy<kFxyPCG>(A, L, s);
  }
}

int launch_fz_mode(FFTArgs& A, int L, int mode, hipStream_t s) {
  switch (mode) {
    case kFzScale: return launch_fz<kFzScale>(A, L, s);
    case kFzPre: return launch_fz<kFzPre>(A, L, s);
    case kFzInit: return launch_fz<kFzInit>(A, L, s);
    default: return launch_fz<kFzPCG>(A, L, s);
  }
}

int launch_fxyi_mode(FFTArgs& A, int L, int mode, hipStream_t s) {
  return mode == kFxyiPre ? launch_fxyi<kFxyiPre>(A, L, s) : launch_fxyi<kFxyiPCG>(A, L, s);
}

}  // namespace pgj

using namespace pgj;

extern "C" size_t pgj_plan_bytes(const pgj_grid* g) { return plan_layout(*g).bytes; }

extern "C" int pgj_plan_init(const pgj_grid* g, void* plan, void* stream) {
  if (int e = check_fft(g)) return e;
  PGJ_REQUIRE(plan != nullptr, "null plan");
  PGJ_REQUIRE(g->hx > 0 && g->hy > 0 && g->hz > 0 && g->mu > 0 && g->lam + g->mu > 0, "invalid voxel or medium");
  PlanLayout pl = plan_layout(*g);
  char* b = static_cast<char*>(plan);
  hipStream_t s = as_stream(stream);
  const int nmax = max(g->nx, max(g->ny, g->nz));
  hipLaunchKernelGGL(plan_twiddles, dim3((nmax + 255) / 256), dim3(256), 0, s, g->nx, g->ny, g->nz,
                     reinterpret_cast<double2*>(b + pl.tw_x), reinterpret_cast<double2*>(b + pl.tw_y),
                     reinterpret_cast<double2*>(b + pl.tw_z));
  PGJ_LAUNCHED("plan_twiddles");
  const FFTGeom f = fft_geom(*g);
  hipLaunchKernelGGL(plan_symbol, dim3(unsigned((f.nf + 255) / 256)), dim3(256), 0, s, g->nz, g->ny, g->nx, g->hx,
                     g->hy, g->hz, g->lam, g->mu, reinterpret_cast<double*>(b + pl.sym));
  PGJ_LAUNCHED("plan_symbol");
  return PGJ_OK;
}

extern "C" int pgj_fi_coeff(const pgj_grid* g, double* out4, double* vtab, void* stream) {
  if (int e = check_fft(g)) return e;
  PGJ_REQUIRE(out4, "null output");
  const FFTGeom f = fft_geom(*g);
  hipLaunchKernelGGL(fi_coeff_kernel, dim3(unsigned((f.nf + 255) / 256)), dim3(256), 0, as_stream(stream), g->nz,
                     g->ny, g->nx, g->hx, g->hy, g->hz, g->lam, g->mu, out4, vtab);
  PGJ_LAUNCHED("fi_coeff_kernel");
  return PGJ_OK;
}

extern "C" int pgj_plan_symbol(const pgj_grid* g, const void* plan, double* out, void* stream) {
  if (int e = check_fft(g)) return e;
  PGJ_REQUIRE(plan && out, "null plan or output");
  const FFTGeom f = fft_geom(*g);
  PlanLayout pl = plan_layout(*g);
  hipError_t e = hipMemcpyAsync(out, static_cast<const char*>(plan) + pl.sym, size_t(6) * f.nf * 8,
                                hipMemcpyDeviceToDevice, as_stream(stream));
  if (e != hipSuccess) {
    set_error("symbol copy failed: %s", hipGetErrorString(e));
    return PGJ_ERR_HIP;
  }
  return PGJ_OK;
}

extern "C" size_t pgj_precond_ws_bytes(const pgj_grid* g, int L) {
  if (!g) return 0;
  return precond_layout(*g, L).bytes;
}

extern "C" int pgj_scale_spectrum(const pgj_grid* g, const void* plan, const double* s, void* s_hat, void* ws,
                                  void* stream) {
  if (int e = check_fft(g)) return e;
  PGJ_REQUIRE(plan && s && s_hat && ws, "invalid scale spectrum arguments");
  FFTArgs A = fft_args(*g, plan);
  PrecondLayout pl = precond_layout(*g, 1);
  if (static_fft(*g)) {
    // the compile-time passes read s in its (1, 3N) node-interleaved layout
    A.s = s;
    A.soa = 0;
    A.spec = static_cast<double2*>(s_hat);
    A.pr = reinterpret_cast<double*>(static_cast<char*>(ws) + pl.pr);
    A.pg = reinterpret_cast<double*>(static_cast<char*>(ws) + pl.pg);
    hipStream_t st = as_stream(stream);
    if (int e = launch_fxy_static(A, 1, kFxyScale, st)) return e;
    return launch_fz_static(A, 1, kFzScale, st);
  }
  A.s = s;
  A.spec = static_cast<double2*>(s_hat);
  A.pr = reinterpret_cast<double*>(static_cast<char*>(ws) + pl.pr);
  A.pg = reinterpret_cast<double*>(static_cast<char*>(ws) + pl.pg);
  hipStream_t st = as_stream(stream);
  if (int e = launch_fxy<kFxyScale>(A, 1, st)) return e;
  return launch_fz<kFzScale>(A, 1, st);
}

extern "C" int pgj_precond_apply(const pgj_grid* g, int L, const void* plan, const double* r, const double* s,
                                 const void* s_hat, double* z, void* ws, void* stream) {
  if (int e = check_fft(g)) return e;
  PGJ_REQUIRE(L >= 1 && L <= 65535 && plan && r && z && ws, "invalid preconditioner arguments");
  PGJ_REQUIRE((s == nullptr) == (s_hat == nullptr), "s and s_hat must be given together");
  FFTArgs A = fft_args(*g, plan);
  PrecondLayout pl = precond_layout(*g, L);
  if (static_fft(*g)) {
    // the compile-time passes read r and s and write z in their (L, 3N) node-interleaved layout; r is read by
    // the forward pass only, so z may alias it
    char* w = static_cast<char*>(ws);
    hipStream_t st = as_stream(stream);
    A.rin = r;
    A.s = s;
    A.shat = static_cast<const double2*>(s_hat);
    A.spec = reinterpret_cast<double2*>(w + pl.spec);
    A.pr = reinterpret_cast<double*>(w + pl.pr);
    A.pg = reinterpret_cast<double*>(w + pl.pg);
    A.zout = z;
    A.soa = 0;
    if (int e = launch_fxy_static(A, L, kFxyPre, st)) return e;
    if (int e = launch_fz_static(A, L, kFzPre, st)) return e;
    return launch_fxyi_static(A, L, st, kFxyiPre);
  }
  const int kc = precond_kc(A.f, g->ny, L);
  if (kc != A.f.kc) {
    A.f.kc = kc;
    A.f.nkb = (A.f.nkx + kc - 1) / kc;
    A.f.lds_z = size_t(kc) * g->nz * 3 * 16 + 64 * 8;
  }
  char* w = static_cast<char*>(ws);
  A.rin = r;
  A.s = s;
  A.shat = static_cast<const double2*>(s_hat);
  A.spec = reinterpret_cast<double2*>(w + pl.spec);
  A.pr = reinterpret_cast<double*>(w + pl.pr);
  A.pg = reinterpret_cast<double*>(w + pl.pg);
  A.zout = z;
  hipStream_t st = as_stream(stream);
  if (int e = launch_fxy<kFxyPre>(A, L, st)) return e;
  if (int e = launch_fz<kFzPre>(A, L, st)) return e;
  return launch_fxyi<kFxyiPre>(A, L, st);
}
