// Compile-time-shaped FFT/Green kernels of the resident solver for cubic
// grids n in {16, 24, 32, 48, 64, 96, 128}. Vectors and the scale s are SoA
// (L, 3, N): one block transforms one component plane, so a plane needs only
// n (n/2+1) complex values of LDS (34 KiB at 64^3) and every index is a
// constant expression. Semantics are those of the runtime kernels in fft.hip.
#include <cmath>

#include "fft_kernels.h"
#include "stockham.h"

namespace pgj {

// Waves per SIMD that the LDS of a block allows (160 KiB per CU, 4 SIMDs),
// capped at 4: the register budget the stream batching is compiled for.
constexpr int lds_occupancy(size_t lds, int threads) {
  const int w = int(size_t(160 * 1024) / lds) * (threads / 64) / 4;
  return w < 1 ? 1 : (w > 4 ? 4 : w);
}

template <int N>
struct Cube {
  static constexpr int M = N / 2;             // complex length of the packed x transform
  static constexpr int NKX = N / 2 + 1;       // half-spectrum width
  static constexpr int RS = NKX;              // LDS row stride (complex)
  static constexpr int PLANE = N * N;         // nodes per z-plane
  static constexpr int64_t NF = int64_t(N) * N * NKX;
  static constexpr int TXY = N >= 96 ? 512 : 256;
  // LDS bytes for complex values of ES bytes: the data, the twiddle table W_N^j
  // (j < N; the same table serves x, y and z on a cube) and a reduction tail
  // of 4 values x 8 waves
  static constexpr size_t lds_xy(size_t es) { return (size_t(N) * RS + N) * es + 4 * 8 * 8; }
  template <typename R>
  static constexpr int occ_xy() { return lds_occupancy(lds_xy(2 * sizeof(R)), TXY); }
  using Rows = Lines<N, N, 0, RS, 1, false>;          // x transforms (length M) on the N rows
  using Cols = Lines<NKX, NKX, 0, 1, RS, true>;       // y transforms (length N) on the NKX columns
  // vector streams of one z-plane: pairs h = threadIdx.x + i TXY (i < ITER) of
  // consecutive x, in chunks of CH pairs per thread whose loads are issued together
  static constexpr int HALF = PLANE / 2;
  static constexpr int ITER = (HALF + TXY - 1) / TXY;
  static constexpr int CH = ITER <= 8 ? ITER : (ITER % 8 == 0 ? 8 : (ITER + 1) / 2);
  static constexpr bool EXACT = HALF % TXY == 0;
};

// z pass of a cube in precision R: KC kx columns of all three components per
// block; a row of KC complex values is the contiguous run a block reads per
// (c, kz). A last chunk narrower than KC is supported (columns past NKX are
// zero in LDS and skipped). 128^3 takes 13 columns in FP32 (104-byte runs
// instead of 40) but keeps 5 in FP64, where 13 would need 147 VGPRs at 2 blocks
// per CU; 17 columns at 96^3 measured no faster than 7. KCX > 0 overrides the
// chunk width (the wide 64^3 FP64 instance, kWideKC).
template <int N, typename R, int KCX = 0>
struct ZPass {
  static constexpr int NKX = N / 2 + 1;
  static constexpr int KC = KCX > 0    ? KCX
                            : N == 16  ? 9
                            : N == 24  ? 13
                            : N == 32  ? 17
                            : N == 48  ? 5
                            : N == 64  ? 11
                            : N == 96  ? 7
                            : sizeof(R) == 8 ? 5 : 13;
  static constexpr int TZ = KC >= 13 && N >= 96 ? 512 : 256;
  static constexpr int NKB = (NKX + KC - 1) / KC;
  static constexpr bool KFULL = NKX % KC == 0;
  // columns and the twiddle table; the reduction reuses the column storage once
  // the columns are written back (2 blocks of the 128^3 FP32 pass fit one CU)
  static constexpr size_t lds = (size_t(3) * N * KC + N) * 2 * sizeof(R);
  static constexpr int occ = lds_occupancy(lds, TZ);
  static constexpr int per_cu = occ * 4 / (TZ / 64) > 0 ? occ * 4 / (TZ / 64) : 1;  // resident blocks per CU
  using ZCols = Lines<3 * KC, KC, N * KC, 1, KC, true>;  // z transforms on (component, column)
};

// Wide z chunks: at 64^3 in FP64 the default 11 columns give 3 chunks per ky row
// in 34 KiB of LDS (4 blocks per CU), 17 columns give 2 chunks in 52 KiB (3 per
// CU), and the symbol and s_hat are read by 2 blocks per ky row instead of 3. A
// 6-load launch is 1152 blocks against 1024 slots of 256 CUs with the default, a
// partial second round, and 768 blocks in exactly one round with the wide
// chunks: the pass takes 31 us instead of 39 us. Launches of many rounds (a
// batched set of 72 loads) gain from the fewer re-reads; launches that fit one
// round with the default (up to 5 loads) keep it, since a wide block does more
// work in that single round. zpass_wide decides.
template <int N, typename R>
constexpr int kWideKC = N == 64 && sizeof(R) == 8 ? 17 : 0;

// Register-resident z pass of the 64^3 cube (fz64_reg below): blocks per load, which lay out its partials:
// 4 chunks of 8 kx columns per ky row, then 8 blocks for the Nyquist column kx = 32 (8 ky rows each).
constexpr int kZRegBlocks = 4 * 64 + 8;

// Per-load partial slots written by the z pass variant (FFTArgs::zwide: 0 default chunks, 1 wide chunks,
// kZReg the register-resident pass); fxy and fxyi reduce exactly these.
template <int N, typename R>
__host__ __device__ constexpr int z_slots(int variant) {
  if (N == 64 && variant == kZReg) return kZRegBlocks;
  return (kWideKC<N, R> > 0 && variant == 1 ? ZPass<N, R, (kWideKC<N, R> > 0 ? kWideKC<N, R> : 1)>::NKB
                                            : ZPass<N, R>::NKB) * N;
}

// Global -> LDS copy of COUNT values in two phases, so that independent work
// can run while the loads are in flight: load() issues every load of a thread
// (surplus lanes load a clamped, in-bounds index), store() writes the LDS slots.
// map(e) is the global index of LDS slot e.
template <int COUNT, int T, typename X>
struct StageIn {
  static constexpr int PER = (COUNT + T - 1) / T;
  X v[PER];
  template <typename Map>
  __device__ inline void load(const X* __restrict__ src, Map map) {
#pragma unroll
    for (int k = 0; k < PER; ++k) {
      const int e = COUNT % T == 0 ? int(threadIdx.x) + k * T : min(int(threadIdx.x) + k * T, COUNT - 1);
      v[k] = src[map(e)];
    }
  }
  __device__ inline void store(X* lds) const {
#pragma unroll
    for (int k = 0; k < PER; ++k) {
      const int e = int(threadIdx.x) + k * T;
      if (COUNT % T == 0 || e < COUNT) lds[e] = v[k];
    }
  }
};

// Pair index of chunk slot i of a vector stream, clamped in bounds for loads.
template <int N>
__device__ inline int pair_of(int i) {
  using C = Cube<N>;
  const int h = int(threadIdx.x) + i * C::TXY;
  return C::EXACT ? h : min(h, C::HALF - 1);
}

// ------------------------------------------------------------ forward xy
template <int N, int MODE, typename R, int PLAN = 0>
__global__ __launch_bounds__(Cube<N>::TXY) __attribute__((amdgpu_waves_per_eu(Cube<N>::template occ_xy<R>())))
void fxy_static(FFTArgsT<R> A) {
  using C = Cube<N>;
  using X = cplx<R>;
  constexpr int T = C::TXY;
  extern __shared__ __align__(16) unsigned char smem[];
  X* plane = reinterpret_cast<X*>(smem);
  X* tw = plane + N * C::RS;
  double* red = reinterpret_cast<double*>(tw + N);
  const int z = blockIdx.x, c = blockIdx.y, l = blockIdx.z;
  const int64_t NN = int64_t(N) * N * N;
  if (MODE == kFxyPCG && !A.st.active[l]) return;
  const int64_t off = (int64_t(l) * 3 + c) * NN + int64_t(z) * C::PLANE;
  const int64_t soff = (int64_t(design_of(l, A.group)) * 3 + c) * NN + int64_t(z) * C::PLANE;
  double sq = 0.0, sm = 0.0;
  // two nodes (x even, x+1) per pair: one packed complex of the x transform
  const X* __restrict__ sp = reinterpret_cast<const X*>(A.s + soff);
  const X* __restrict__ rinp = reinterpret_cast<const X*>(A.rin + off);
  X* __restrict__ rp = reinterpret_cast<X*>(A.r + off);
  const X* __restrict__ qp = reinterpret_cast<const X*>(A.q + off);
  const bool scaled = A.s != nullptr;
  // the scale and preconditioner passes also read the drop-in layout (A.soa == 0): xyz triples per node, so
  // component c of the pair's two nodes sits 24 bytes apart (the other components' blocks share the lines in L2)
  const bool aos = MODE != kFxyPCG && !A.soa;
  const int64_t aoff = int64_t(z) * C::PLANE * 3 + c;
  const R* __restrict__ s_aos = aos && scaled ? A.s + int64_t(design_of(l, A.group)) * 3 * NN + aoff : nullptr;
  const R* __restrict__ r_aos = aos && MODE == kFxyPre && A.rin ? A.rin + int64_t(l) * 3 * NN + aoff : nullptr;
  // FP32 apply of FP64 vectors: the residual is read in FP64 and rounded to R
  const bool wide_in = sizeof(R) == 4 && aos && MODE == kFxyPre && A.rin64 != nullptr;
  const double* __restrict__ r64 = wide_in ? A.rin64 + int64_t(l) * 3 * NN + aoff : nullptr;
  auto pair_aos = [](const R* base, int h) {
    const R* q = base + 6 * int64_t(h);
    return c2(q[0], q[3]);
  };
  double alpha = 0.0;
#pragma unroll
  for (int i0 = 0; i0 < C::ITER; i0 += C::CH) {
    // every load of the chunk is issued before the first use
    X va[C::CH], vb[C::CH], vs[C::CH];
#pragma unroll
    for (int i = 0; i < C::CH; ++i) {
      const int h = pair_of<N>(i0 + i < C::ITER ? i0 + i : C::ITER - 1);
      if (MODE == kFxyScale) {
        va[i] = aos ? pair_aos(s_aos, h) : sp[h];
      } else if (MODE == kFxyPre) {
        if (wide_in) {
          const double* q = r64 + 6 * int64_t(h);
          va[i] = c2(R(q[0]), R(q[3]));
        } else {
          va[i] = aos ? pair_aos(r_aos, h) : rinp[h];
        }
      } else {
        va[i] = rp[h];
        vb[i] = qp[h];
      }
      vs[i] = c2(R(1), R(1));
    }
    StageIn<N, T, X> twl;  // the twiddle table, staged with the first chunk
    if (i0 == 0) twl.load(A.twx, [](int e) { return e; });
    // PCG: the step length is reduced while the r and q loads are in flight
    if (MODE == kFxyPCG && i0 == 0) {
      double pq, rz;
      // per-load partials are laid out with the geometry's chunk count (slices of a set are offset by it)
      wave_reduce_partials_2(A.ppq + int64_t(l) * A.ppq_stride, A.ppq_count, 1, A.pg + int64_t(l) * A.f.nkb * N * 4,
                             z_slots<N, R>(A.zwide), 4, pq, rz);
      const bool valid = isfinite(pq) && pq > 0.0 && isfinite(rz) && rz > 0.0;
      alpha = valid ? rz / pq : 0.0;
      if (threadIdx.x == 0 && z == 0 && c == 0) {
        A.st.rz[l] = rz;
        A.st.alpha[l] = alpha;
        A.st.valid[l] = valid;
        A.st.first[l] = 0;
      }
    }
    if (i0 == 0) twl.store(tw);
    if (MODE != kFxyScale && scaled) {
#pragma unroll
      for (int i = 0; i < C::CH; ++i) {
        const int h = pair_of<N>(i0 + i < C::ITER ? i0 + i : C::ITER - 1);
        vs[i] = aos ? pair_aos(s_aos, h) : sp[h];
      }
    }
    const R ra = R(alpha);
#pragma unroll
    for (int i = 0; i < C::CH; ++i) {
      const int h = int(threadIdx.x) + (i0 + i) * T;
      if (i0 + i >= C::ITER || (!C::EXACT && h >= C::HALF)) continue;
      X v = va[i];
      if (MODE != kFxyScale) {
        X rn = va[i];
        if (MODE == kFxyPCG) {
          rn = c2(fma(-ra, vb[i].x, va[i].x), fma(-ra, vb[i].y, va[i].y));
          rp[h] = rn;
        }
        sq = fma(double(rn.x), double(rn.x), fma(double(rn.y), double(rn.y), sq));
        sm += double(rn.x) + double(rn.y);
        v = c2(vs[i].x * rn.x, vs[i].y * rn.y);
      }
      const int e = 2 * h;
      const int x = e % N, y = e / N;
      plane[y * C::RS + x / 2] = v;
    }
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
  sfft<false, C::M, typename C::Rows, T, PLAN>(plane, tw, 2);
  // split the packed spectrum of every row into its half spectrum
  constexpr int HALF = C::M / 2, PAIRS = N * (HALF + 1);
  for (int b = threadIdx.x; b < PAIRS; b += T) {
    const int row = b / (HALF + 1), k = b % (HALF + 1);
    X* L = plane + row * C::RS;
    const X zk = L[k % C::M], zm = L[(C::M - k) % C::M];
    const X E = c2(R(0.5) * (zk.x + zm.x), R(0.5) * (zk.y - zm.y));
    const X O = c2(R(0.5) * (zk.y + zm.y), R(-0.5) * (zk.x - zm.x));
    const X wo = cmul(tw[k], O);
    L[k] = cadd(E, wo);
    if (C::M - k != k) L[C::M - k] = conjd(csub(E, wo));
  }
  __syncthreads();
  sfft<false, N, typename C::Cols, T, PLAN>(plane, tw, 1);
  X* out = A.spec + ((int64_t(l) * 3 + c) * N + z) * int64_t(N) * C::NKX;
  for (int e = threadIdx.x; e < N * C::NKX; e += T) out[e] = plane[e];
}

// ---------------------------------------------------------------- z + Green
template <int N, int MODE, typename R, int KCX = 0, int PLAN = 0>
__global__ __launch_bounds__((ZPass<N, R, KCX>::TZ)) __attribute__((amdgpu_waves_per_eu((ZPass<N, R, KCX>::occ))))
void fz_static(FFTArgsT<R> A) {
  using C = Cube<N>;
  using X = cplx<R>;
  using Z = ZPass<N, R, KCX>;
  constexpr int T = Z::TZ, KC = Z::KC, NKX = C::NKX;
  extern __shared__ __align__(16) unsigned char smem[];
  X* cols = reinterpret_cast<X*>(smem);
  X* tw = cols + 3 * N * KC;
  double* red = reinterpret_cast<double*>(cols);  // used only after the columns are written back
  // ky is the fastest grid index: the kx chunks of one ky row (which share
  // cache lines) get block ids N apart, a multiple of 8 for every supported N,
  // and so run on the same XCD / L2.
  const int y = blockIdx.x, kb = blockIdx.y, l = blockIdx.z;
  const int kx0 = kb * KC;
  const int blk = y * Z::NKB + kb;
  const bool lead = threadIdx.x == 0 && kb == 0 && y == 0;
  const int64_t NN = int64_t(N) * N * N;
  // the spectrum columns are loaded first (for every load that may continue)
  // and stay in flight through the convergence test and the mean reduction
  if (MODE == kFzPCG && !A.st.active[l]) {
    // an inactive load stays inactive: the test below with active = false
    if (lead) {
      A.st.active_next[l] = 0;
      A.st.it_next[l] = A.st.it[l];
      A.st.bad_next[l] = A.st.bad[l];
      A.st.stopped[l] = 0;
    }
    return;
  }
  // the spectrum columns and the twiddle table go to LDS first (LDS slot
  // e = cz * KC + kcol with cz = c * N + zz); they are read after the test below
  X* S = A.spec + int64_t(l) * 3 * C::NF;
  // columns of this chunk (the last chunk may hold fewer than KC; its missing columns are zero in LDS)
  const int kcn = Z::KFULL ? KC : min(KC, NKX - kx0);
  {
    StageIn<3 * N * KC, T, X> in;
    in.load(S, [&](int e) { return (int64_t(e / KC) * N + y) * NKX + kx0 + min(e % KC, kcn - 1); });
    if (!Z::KFULL) {
#pragma unroll
      for (int k = 0; k < in.PER; ++k)
        if ((int(threadIdx.x) + k * T) % KC >= kcn) in.v[k] = c2(R(0), R(0));
    }
    StageIn<N, T, X> twl;
    twl.load(A.twz, [](int e) { return e; });
    in.store(cols);
    twl.store(tw);
  }
  double mean[3] = {0.0, 0.0, 0.0};
  if (MODE == kFzInit || MODE == kFzPCG) {
    const int lane = threadIdx.x & 63;
    // one pass over the per-plane partials (|r|^2 and the component sums of r)
    double v = 0.0, vm[3] = {0.0, 0.0, 0.0};
    for (int i = lane; i < N; i += 64) {
      const double* p = A.pr + (int64_t(l) * N + i) * 6;
      v += (p[0] + p[2]) + p[4];
      vm[0] += p[1];
      vm[1] += p[3];
      vm[2] += p[5];
    }
    const double res2 = wave_sum(v);
    if (A.shat)
      for (int c = 0; c < 3; ++c) mean[c] = wave_sum(vm[c]) / double(NN);
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
    for (int c = 0; c < 3; ++c)
      mean[c] = wave_reduce_partials(A.pr + (int64_t(l) * N * 3 + c) * 2 + 1, N, 6) / double(NN);
  }
  // a fixed number of frequency slots per thread, in groups of GS slots whose
  // symbol and s_hat loads are all issued before the first is used. In FP32 one
  // group holds every slot and is issued before the forward transform, so the
  // loads stay in flight while it runs; FP64, and FP32 with more than 3 slots,
  // lack the registers for that and issue their groups after it (the scaling
  // pass needs none).
  constexpr int EPT = (N * KC + T - 1) / T;
  constexpr bool EARLY = sizeof(R) == 4 && EPT <= 3;
  constexpr int GS = (EARLY || EPT <= 3) ? EPT : 2;
  constexpr bool FULL = (N * KC) % T == 0;
  const X* __restrict__ shat_d = A.shat ? A.shat + int64_t(design_of(l, A.group)) * 3 * C::NF : nullptr;
  const R* __restrict__ sym = A.sym;
  auto slot = [&](int k) { return FULL ? int(threadIdx.x) + k * T : min(int(threadIdx.x) + k * T, N * KC - 1); };
  auto freq = [&](int e) { return (int64_t(e / KC) * N + y) * NKX + kx0 + min(e % KC, kcn - 1); };
  R g[GS][6];
  X sh[GS][3];
  auto load_group = [&](int k0) {
#pragma unroll
    for (int i = 0; i < GS; ++i) {
      const int64_t fidx = freq(slot(k0 + i < EPT ? k0 + i : EPT - 1));
#pragma unroll
      for (int j = 0; j < 6; ++j) g[i][j] = sym[j * C::NF + fidx];
#pragma unroll
      for (int c = 0; c < 3; ++c) sh[i][c] = c2(R(0), R(0));
    }
    if (shat_d) {
#pragma unroll
      for (int i = 0; i < GS; ++i) {
        const int64_t fidx = freq(slot(k0 + i < EPT ? k0 + i : EPT - 1));
#pragma unroll
        for (int c = 0; c < 3; ++c) sh[i][c] = shat_d[int64_t(c) * C::NF + fidx];
      }
    }
  };
  if (MODE != kFzScale && EARLY) load_group(0);
  __syncthreads();
  sfft<false, N, typename Z::ZCols, T, PLAN>(cols, tw, 1);
  if (MODE == kFzScale) {
    for (int e = threadIdx.x; e < 3 * N * KC; e += T) {
      const int kcol = e % KC, cz = e / KC;
      if (Z::KFULL || kcol < kcn) S[(int64_t(cz) * N + y) * NKX + kx0 + kcol] = cols[e];
    }
    return;
  }
  double rz = 0.0, sw[3] = {0.0, 0.0, 0.0};
#pragma unroll
  for (int k0 = 0; k0 < EPT; k0 += GS) {
    if (!(EARLY && k0 == 0)) load_group(k0);
#pragma unroll
    for (int i = 0; i < GS; ++i) {
      const int k = k0 + i;
      const int e = threadIdx.x + k * T;
      if (k >= EPT || (!FULL && e >= N * KC)) continue;
      const int kcol = e % KC, zz = e / KC;
      if (!Z::KFULL && kcol >= kcn) continue;
      const int kx = kx0 + kcol;
      X v[3];
#pragma unroll
      for (int c = 0; c < 3; ++c) v[c] = cols[(c * N + zz) * KC + kcol];
      if (shat_d) {
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          const R mc = R(mean[c]);
          v[c] = c2(fma(-mc, sh[i][c].x, v[c].x), fma(-mc, sh[i][c].y, v[c].y));
        }
      }
      const R g0 = g[i][0], g1 = g[i][1], g2 = g[i][2], g3 = g[i][3], g4 = g[i][4], g5 = g[i][5];
      X w[3];
      w[0] = c2(g0 * v[0].x + g3 * v[1].x + g4 * v[2].x, g0 * v[0].y + g3 * v[1].y + g4 * v[2].y);
      w[1] = c2(g3 * v[0].x + g1 * v[1].x + g5 * v[2].x, g3 * v[0].y + g1 * v[1].y + g5 * v[2].y);
      w[2] = c2(g4 * v[0].x + g5 * v[1].x + g2 * v[2].x, g4 * v[0].y + g5 * v[1].y + g2 * v[2].y);
      const double wgt = (kx == 0 || kx == N / 2) ? 1.0 : 2.0;
      double t = 0.0;
#pragma unroll
      for (int c = 0; c < 3; ++c) t += double(v[c].x) * double(w[c].x) + double(v[c].y) * double(w[c].y);
      rz = fma(wgt, t, rz);
      if (shat_d) {
#pragma unroll
        for (int c = 0; c < 3; ++c)
          sw[c] = fma(wgt, double(sh[i][c].x) * double(w[c].x) + double(sh[i][c].y) * double(w[c].y), sw[c]);
      }
#pragma unroll
      for (int c = 0; c < 3; ++c) cols[(c * N + zz) * KC + kcol] = w[c];
    }
  }
  __syncthreads();
  sfft<true, N, typename Z::ZCols, T, PLAN>(cols, tw, 1);
  for (int e = threadIdx.x; e < 3 * N * KC; e += T) {
    const int kcol = e % KC, cz = e / KC;
    if (Z::KFULL || kcol < kcn) S[(int64_t(cz) * N + y) * NKX + kx0 + kcol] = cols[e];
  }
  double t[4] = {rz, sw[0], sw[1], sw[2]};
  block_sums<4>(t, red);
  if (threadIdx.x == 0) {
    // per-load partials are laid out with the geometry's chunk count (slices of a set are offset by it)
    double* pg = A.pg + (int64_t(l) * A.f.nkb * N + blk) * 4;
    pg[0] = t[0];
    pg[1] = t[1];
    pg[2] = t[2];
    pg[3] = t[3];
  }
}

// ------------------------------------------------------- z + Green, in registers (n = 64)
// The z pass of the 64^3 cube with its transforms in registers. A block is three waves, wave c holding
// component c of 8 kx columns of one ky row (the Nyquist blocks: column kx = 32 of 8 ky rows); lane
// (kcol = lane % 8, j = lane / 8) holds the points kz = j + 8 r (r < 8) of its column. The 64-point
// transform is two radix-8 Stockham passes: pass 1 is a DFT over the lane's own points, an 8 x 8 transpose
// among the 8 lanes of a column goes through LDS, and pass 2 (twiddles W^(j r)) leaves the lane with the
// frequencies kz = j + 8 r: the same points the first inverse pass reads. LDS carries the two transposes of
// each transform and the exchange of the three components for the 3x3 Green multiply; the default pass
// stages every point through LDS for each of its radix stages. Semantics are those of fz_static.
template <int MODE, typename R>
__global__ __launch_bounds__(192) __attribute__((amdgpu_waves_per_eu(3))) void fz64_reg(FFTArgsT<R> A) {
  // LDS rows of one column, padded so that the 16 lanes (8 columns x 2 j) of a transposed or exchange read
  // of 16-byte values fall on distinct banks
  constexpr int N = 64, NKX = N / 2 + 1, RS = N + 2;
  constexpr int64_t NF = int64_t(N) * N * NKX;
  using X = cplx<R>;
  extern __shared__ __align__(16) unsigned char smem[];
  X* buf = reinterpret_cast<X*>(smem);                       // [component][kcol][RS]
  X* twl = buf + 3 * 8 * RS;                                  // W_64^k, k < 64
  double* red = reinterpret_cast<double*>(twl + N);           // block sums: 4 values x 3 waves
  const int c = __builtin_amdgcn_readfirstlane(int(threadIdx.x) >> 6);
  const int lane = int(threadIdx.x) & 63, kcol = lane & 7, j = lane >> 3;
  const int b = int(blockIdx.x), l = int(blockIdx.z);
  const int ky = b < 256 ? b % 64 : 8 * (b - 256) + kcol;
  const int kx = b < 256 ? 8 * (b / 64) + kcol : N / 2;
  const bool lead = threadIdx.x == 0 && b == 0;
  const int64_t NN = int64_t(N) * N * N;
  if (MODE == kFzPCG && !A.st.active[l]) {
    if (lead) {
      A.st.active_next[l] = 0;
      A.st.it_next[l] = A.st.it[l];
      A.st.bad_next[l] = A.st.bad[l];
      A.st.stopped[l] = 0;
    }
    return;
  }
  X* S = A.spec + int64_t(l) * 3 * NF;
  auto fidx = [&](int kz) { return (int64_t(kz) * N + ky) * NKX + kx; };
  // the column's points and the twiddle table are in flight through the convergence test
  X v[8];
#pragma unroll
  for (int r = 0; r < 8; ++r) v[r] = S[int64_t(c) * NF + fidx(j + 8 * r)];
  const X twv = A.twz[lane];
  // s_hat_c at the lane's frequencies, in flight through the convergence test and the forward transform
  const X* __restrict__ shat_d = A.shat ? A.shat + int64_t(design_of(l, A.group)) * 3 * NF : nullptr;
  X sh[8];
  if (MODE != kFzScale) {
#pragma unroll
    for (int r = 0; r < 8; ++r) sh[r] = shat_d ? shat_d[int64_t(c) * NF + fidx(j + 8 * r)] : c2(R(0), R(0));
  }
  double mean_c = 0.0;
  if (MODE == kFzInit || MODE == kFzPCG) {
    double res2 = 0.0, vm = 0.0;
    {
      double t = 0.0, m = 0.0;
      for (int i = lane; i < N; i += 64) {
        const double* p = A.pr + (int64_t(l) * N + i) * 6;
        t += (p[0] + p[2]) + p[4];
        m += p[2 * c + 1];
      }
      res2 = wave_sum(t);
      vm = wave_sum(m);
    }
    if (A.shat) mean_c = vm / double(NN);
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
      const bool valid = A.st.valid[l] != 0;
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
    mean_c = wave_reduce_partials(A.pr + (int64_t(l) * N * 3 + c) * 2 + 1, N, 6) / double(NN);
  }
  if (c == 0) twl[lane] = twv;
  X* row = buf + (c * 8 + kcol) * RS;
  const X* none = nullptr;
  // forward: pass 1 over x[j + 8 r] -> y[8 j + r], transpose, pass 2 (twiddles W^(j r)) -> Y[j + 8 r]
  dft<8, false>(v, none, N, 1);
#pragma unroll
  for (int r = 0; r < 8; ++r) row[8 * j + r] = v[r];
  // this wave's row of the symbol (G_c0, G_c1, G_c2 of [G00, G11, G22, G01, G02, G12]), in flight through
  // the transpose, the second pass and the component exchange
  R g[8][3];
  if (MODE != kFzScale) {
#pragma unroll
    for (int r = 0; r < 8; ++r) {
      const int64_t f = fidx(j + 8 * r);
#pragma unroll
      for (int d = 0; d < 3; ++d) g[r][d] = A.sym[(d == c ? c : c + d + 2) * NF + f];
    }
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < 8; ++r) v[r] = r == 0 ? row[j] : cmul(row[j + 8 * r], twl[(j * r) & (N - 1)]);
  dft<8, false>(v, none, N, 1);
  if (MODE == kFzScale) {
#pragma unroll
    for (int r = 0; r < 8; ++r) S[int64_t(c) * NF + fidx(j + 8 * r)] = v[r];
    return;
  }
  // projection mean correction, then the three components meet in LDS for the Green multiply
  if (shat_d) {
    const R mc = R(mean_c);
#pragma unroll
    for (int r = 0; r < 8; ++r) v[r] = c2(fma(-mc, sh[r].x, v[r].x), fma(-mc, sh[r].y, v[r].y));
  }
  // a lane writes back exactly the slots it read after the transpose: no other lane reads them again
#pragma unroll
  for (int r = 0; r < 8; ++r) row[j + 8 * r] = v[r];
  __syncthreads();
  const double wgt = (kx == 0 || kx == N / 2) ? 1.0 : 2.0;
  double rz = 0.0, sw = 0.0;
  X w[8];
#pragma unroll
  for (int r = 0; r < 8; ++r) {
    const int kz = j + 8 * r;
    const X v0 = buf[(0 * 8 + kcol) * RS + kz], v1 = buf[(1 * 8 + kcol) * RS + kz], v2 = buf[(2 * 8 + kcol) * RS + kz];
    w[r] = c2(g[r][0] * v0.x + g[r][1] * v1.x + g[r][2] * v2.x, g[r][0] * v0.y + g[r][1] * v1.y + g[r][2] * v2.y);
    rz = fma(wgt, double(v[r].x) * double(w[r].x) + double(v[r].y) * double(w[r].y), rz);
    if (shat_d) sw = fma(wgt, double(sh[r].x) * double(w[r].x) + double(sh[r].y) * double(w[r].y), sw);
  }
  __syncthreads();  // the exchange rows are read before the inverse transposes overwrite them
  // inverse: pass 1 over W[j + 8 r] -> 8 j + r, transpose, pass 2 with conjugate twiddles -> z = j + 8 r
  dft<8, true>(w, none, N, 1);
#pragma unroll
  for (int r = 0; r < 8; ++r) row[8 * j + r] = w[r];
  __syncthreads();
#pragma unroll
  for (int r = 0; r < 8; ++r) w[r] = r == 0 ? row[j] : cmul(row[j + 8 * r], conjd(twl[(j * r) & (N - 1)]));
  dft<8, true>(w, none, N, 1);
#pragma unroll
  for (int r = 0; r < 8; ++r) S[int64_t(c) * NF + fidx(j + 8 * r)] = w[r];
  double t[4] = {rz, c == 0 ? sw : 0.0, c == 1 ? sw : 0.0, c == 2 ? sw : 0.0};
  block_sums<4>(t, red);
  if (threadIdx.x == 0) {
    double* pg = A.pg + (int64_t(l) * A.f.nkb * N + b) * 4;
    pg[0] = t[0];
    pg[1] = t[1];
    pg[2] = t[2];
    pg[3] = t[3];
  }
}

// --------------------------------------------------------------- inverse xy
// MODE kFxyiPre: the preconditioner apply alone (z = P(s w), no PCG state).
template <int N, int MODE, typename R, int PLAN = 0>
__global__ __launch_bounds__(Cube<N>::TXY) __attribute__((amdgpu_waves_per_eu(Cube<N>::template occ_xy<R>())))
void fxyi_static(FFTArgsT<R> A) {
  using C = Cube<N>;
  using X = cplx<R>;
  constexpr int T = C::TXY;
  constexpr bool PCG = MODE == kFxyiPCG;
  extern __shared__ __align__(16) unsigned char smem[];
  X* plane = reinterpret_cast<X*>(smem);
  X* tw = plane + N * C::RS;
  const int z = blockIdx.x, c = blockIdx.y, l = blockIdx.z;
  const int64_t NN = int64_t(N) * N * N;
  if (PCG && threadIdx.x == 0 && z == 0 && c == 0) {
    A.st.active[l] = A.st.active_next[l];
    A.st.it[l] = A.st.it_next[l];
    A.st.bad[l] = A.st.bad_next[l];
  }
  const int64_t off = (int64_t(l) * 3 + c) * NN + int64_t(z) * C::PLANE;
  // PCG: the displacement update u += alpha p of this iteration (alpha from the forward pass)
  const bool move = PCG && A.p != nullptr && A.u != nullptr;
  const R ra = move ? R(A.st.alpha[l]) : R(0);
  if (PCG && !A.st.active_next[l]) {
    // a load stopped by this iteration's test still takes its last displacement update
    if (move && A.st.stopped[l]) {
      const X* __restrict__ pp = reinterpret_cast<const X*>(A.p + off);
      X* __restrict__ up = reinterpret_cast<X*>(A.u + off);
      for (int h = threadIdx.x; h < C::PLANE / 2; h += T) {
        const X po = pp[h], uo = up[h];
        up[h] = c2(fma(ra, po.x, uo.x), fma(ra, po.y, uo.y));
      }
    }
    return;
  }
  // the half spectrum of the plane stays in flight through the mean and beta reductions
  StageIn<N * C::NKX, T, X> in;
  StageIn<N, T, X> twl;
  in.load(A.spec + ((int64_t(l) * 3 + c) * N + z) * int64_t(N) * C::NKX, [](int e) { return e; });
  twl.load(A.twy, [](int e) { return e; });
  const int NSLOT = z_slots<N, R>(A.zwide);
  // per-load partials are laid out with the geometry's chunk count (slices of a set are offset by it)
  const double* pgl = A.pg + int64_t(l) * A.f.nkb * N * 4;
  // the output is the next direction p = z + beta p (z alone right after init or without a direction input)
  const bool update = PCG && A.p != nullptr && !A.st.first[l];
  double mean = 0.0, beta = 0.0;
  if (A.shat || update) {
    // one pass: r.z (offset 0) and this component's s.w (offset 1 + c) partials
    const int off[2] = {0, 1 + c};
    double sums[2];
    wave_reduce_partials_n<2>(pgl, off, NSLOT, 4, sums);
    if (A.shat) mean = sums[1] / double(NN);
    if (update) {
      const double rz = A.st.rz[l];
      beta = sums[0] / (rz != 0.0 ? rz : 1.0);
    }
  }
  in.store(plane);
  twl.store(tw);
  // output streams: every load of a chunk is issued before the first use; in
  // FP32 the first chunk is issued before the inverse transforms and stays in
  // flight while they run (FP64, and planes of more than one chunk, lack the
  // registers for that)
  constexpr bool EARLY = sizeof(R) == 4 && C::ITER == C::CH;
  const int64_t soff = (int64_t(design_of(l, A.group)) * 3 + c) * NN + int64_t(z) * C::PLANE;
  const X* __restrict__ sp = reinterpret_cast<const X*>(A.s + soff);
  const X* __restrict__ pp = PCG && A.p != nullptr ? reinterpret_cast<const X*>(A.p + off) : nullptr;
  X* __restrict__ up = move ? reinterpret_cast<X*>(A.u + off) : nullptr;
  X* __restrict__ zp = reinterpret_cast<X*>(A.zout + off);
  const bool scaled = A.s != nullptr;
  // the preconditioner pass also reads s and writes z in the drop-in layout (A.soa == 0): xyz triples per node
  const bool aos = MODE == kFxyiPre && !A.soa;
  const int64_t aoff = int64_t(z) * C::PLANE * 3 + c;
  const R* __restrict__ s_aos = aos && scaled ? A.s + int64_t(design_of(l, A.group)) * 3 * NN + aoff : nullptr;
  R* __restrict__ z_aos = aos && A.zout ? A.zout + int64_t(l) * 3 * NN + aoff : nullptr;
  // FP32 apply of FP64 vectors: z is written in FP64
  double* __restrict__ z64 = sizeof(R) == 4 && aos && A.zout64 ? A.zout64 + int64_t(l) * 3 * NN + aoff : nullptr;
  const R m = R(mean), tb = R(beta);
  X vs[C::CH], vp[C::CH], vu[C::CH];
  auto load_chunk = [&](int i0) {
#pragma unroll
    for (int i = 0; i < C::CH; ++i) {
      vs[i] = c2(R(1), R(1));
      vp[i] = c2(R(0), R(0));
      vu[i] = c2(R(0), R(0));
    }
    if (scaled) {
#pragma unroll
      for (int i = 0; i < C::CH; ++i) {
        const int h = pair_of<N>(i0 + i < C::ITER ? i0 + i : C::ITER - 1);
        if (aos) {
          const R* q = s_aos + 6 * int64_t(h);
          vs[i] = c2(q[0], q[3]);
        } else {
          vs[i] = sp[h];
        }
      }
    }
    if (update || move) {
#pragma unroll
      for (int i = 0; i < C::CH; ++i) vp[i] = pp[pair_of<N>(i0 + i < C::ITER ? i0 + i : C::ITER - 1)];
    }
    if (move) {
#pragma unroll
      for (int i = 0; i < C::CH; ++i) vu[i] = up[pair_of<N>(i0 + i < C::ITER ? i0 + i : C::ITER - 1)];
    }
  };
  if (EARLY) load_chunk(0);
  __syncthreads();
  sfft<true, N, typename C::Cols, T, PLAN>(plane, tw, 1);
  constexpr int HALF = C::M / 2, PAIRS = N * (HALF + 1);
  for (int b = threadIdx.x; b < PAIRS; b += T) {
    const int row = b / (HALF + 1), k = b % (HALF + 1);
    X* L = plane + row * C::RS;
    X xk = L[k], xm = L[C::M - k];
    if (k == 0) {
      xk.y = R(0);
      xm.y = R(0);
    }
    const X wk = conjd(tw[k]);
    const X e1 = c2(xk.x + xm.x, xk.y - xm.y);
    const X o1 = cmul(c2(xk.x - xm.x, xk.y + xm.y), wk);
    const X zk = c2(e1.x - o1.y, e1.y + o1.x);
    if (k != 0 && C::M - k != k) {
      const X wm = conjd(tw[C::M - k]);
      const X e2 = c2(xm.x + xk.x, xm.y - xk.y);
      const X o2 = cmul(c2(xm.x - xk.x, xm.y + xk.y), wm);
      L[C::M - k] = c2(e2.x - o2.y, e2.y + o2.x);
    }
    L[k] = zk;
  }
  __syncthreads();
  sfft<true, C::M, typename C::Rows, T, PLAN>(plane, tw, 2);
#pragma unroll
  for (int i0 = 0; i0 < C::ITER; i0 += C::CH) {
    if (!(EARLY && i0 == 0)) load_chunk(i0);
#pragma unroll
    for (int i = 0; i < C::CH; ++i) {
      const int h = int(threadIdx.x) + (i0 + i) * T;
      if (i0 + i >= C::ITER || (!C::EXACT && h >= C::HALF)) continue;
      const int e = 2 * h;
      const int x = e % N, y = e / N;
      const X w = plane[y * C::RS + x / 2];
      X v = scaled ? c2(vs[i].x * w.x - m, vs[i].y * w.y - m) : w;
      if (update) v = c2(fma(tb, vp[i].x, v.x), fma(tb, vp[i].y, v.y));
      if (z64) {
        double* q = z64 + 6 * int64_t(h);
        q[0] = double(v.x);
        q[3] = double(v.y);
      } else if (aos) {
        R* q = z_aos + 6 * int64_t(h);
        q[0] = v.x;
        q[3] = v.y;
      } else {
        zp[h] = v;
      }
      if (move) up[h] = c2(fma(ra, vp[i].x, vu[i].x), fma(ra, vp[i].y, vu[i].y));
    }
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

template <int N, typename R>
int fxy_static_n(FFTArgsT<R>& A, int L, int mode, hipStream_t s) {
  using C = Cube<N>;
  const size_t lds = C::lds_xy(sizeof(cplx<R>));
  dim3 grid(N, 3, L), block(C::TXY);
  switch (mode) {
    case kFxyScale:
      if (int e = allow_lds_static(fxy_static<N, kFxyScale, R>, lds)) return e;
      hipLaunchKernelGGL((fxy_static<N, kFxyScale, R>), grid, block, lds, s, A);
      break;
    case kFxyPre:
      if (int e = allow_lds_static(fxy_static<N, kFxyPre, R>, lds)) return e;
      hipLaunchKernelGGL((fxy_static<N, kFxyPre, R>), grid, block, lds, s, A);
      break;
    default:
      if (fft_plan(N, int(sizeof(R)), 0) == 1) {
        if (int e = allow_lds_static(fxy_static<N, kFxyPCG, R, 1>, lds)) return e;
        hipLaunchKernelGGL((fxy_static<N, kFxyPCG, R, 1>), grid, block, lds, s, A);
      } else {
        if (int e = allow_lds_static(fxy_static<N, kFxyPCG, R>, lds)) return e;
        hipLaunchKernelGGL((fxy_static<N, kFxyPCG, R>), grid, block, lds, s, A);
      }
  }
  PGJ_LAUNCHED("fxy_static");
  return PGJ_OK;
}

template <int N, typename R, int KCX>
int fz_static_kc(FFTArgsT<R>& A, int L, int mode, hipStream_t s) {
  using Z = ZPass<N, R, KCX>;
  const size_t lds = Z::lds;
  // the per-block partials are laid out for the geometry's chunk count
  PGJ_REQUIRE(Z::NKB <= A.f.nkb, "z pass chunks (%d) exceed the partial layout (%d)", Z::NKB, A.f.nkb);
  dim3 grid(N, Z::NKB, L), block(Z::TZ);
  switch (mode) {
    case kFzScale:
      if (int e = allow_lds_static(fz_static<N, kFzScale, R, KCX>, lds)) return e;
      hipLaunchKernelGGL((fz_static<N, kFzScale, R, KCX>), grid, block, lds, s, A);
      break;
    case kFzPre:
      if (int e = allow_lds_static(fz_static<N, kFzPre, R, KCX>, lds)) return e;
      hipLaunchKernelGGL((fz_static<N, kFzPre, R, KCX>), grid, block, lds, s, A);
      break;
    case kFzInit:
      if (int e = allow_lds_static(fz_static<N, kFzInit, R, KCX>, lds)) return e;
      hipLaunchKernelGGL((fz_static<N, kFzInit, R, KCX>), grid, block, lds, s, A);
      break;
    default:
      if (fft_plan(N, int(sizeof(R)), 1) == 1) {
        if (int e = allow_lds_static(fz_static<N, kFzPCG, R, KCX, 1>, lds)) return e;
        hipLaunchKernelGGL((fz_static<N, kFzPCG, R, KCX, 1>), grid, block, lds, s, A);
      } else {
        if (int e = allow_lds_static(fz_static<N, kFzPCG, R, KCX>, lds)) return e;
        hipLaunchKernelGGL((fz_static<N, kFzPCG, R, KCX>), grid, block, lds, s, A);
      }
  }
  PGJ_LAUNCHED("fz_static");
  return PGJ_OK;
}

template <typename R>
int fz64_reg_launch(FFTArgsT<R>& A, int L, int mode, hipStream_t s) {
  constexpr size_t lds = (size_t(3) * 8 * 66 + 64) * sizeof(cplx<R>) + 4 * 8 * sizeof(double);
  PGJ_REQUIRE(kZRegBlocks <= A.f.nkb * 64, "z pass blocks (%d) exceed the partial layout (%d)", kZRegBlocks,
              A.f.nkb * 64);
  const dim3 grid(kZRegBlocks, 1, L), block(192);
  switch (mode) {
    case kFzScale: hipLaunchKernelGGL((fz64_reg<kFzScale, R>), grid, block, lds, s, A); break;
    case kFzPre: hipLaunchKernelGGL((fz64_reg<kFzPre, R>), grid, block, lds, s, A); break;
    case kFzInit: hipLaunchKernelGGL((fz64_reg<kFzInit, R>), grid, block, lds, s, A); break;
    default: hipLaunchKernelGGL((fz64_reg<kFzPCG, R>), grid, block, lds, s, A);
  }
  PGJ_LAUNCHED("fz64_reg");
  return PGJ_OK;
}

template <int N, typename R>
int fz_static_n(FFTArgsT<R>& A, int L, int mode, hipStream_t s) {
  if constexpr (N == 64) {
    if (A.zwide == kZReg) return fz64_reg_launch<R>(A, L, mode, s);
  }
  if constexpr (kWideKC<N, R> > 0) {
    if (A.zwide == 1) return fz_static_kc<N, R, kWideKC<N, R>>(A, L, mode, s);
  }
  return fz_static_kc<N, R, 0>(A, L, mode, s);
}

// CUs of the current device, cached per device; 0 if the query fails.
static int device_cus() {
  static int cache[64] = {};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 0;
  if (cache[dev] == 0) {
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
    cache[dev] = cus;
  }
  return cache[dev];
}

bool zpass_wide(const pgj_grid& g, int L, int es) {
  if (es != 8 || !static_cube(g) || g.nx != 64 || L < 1) return false;
  const int64_t cus = device_cus();
  using Zn = ZPass<64, double>;
  const int64_t narrow = int64_t(Zn::NKB) * 64 * L;  // blocks of the default pass
  return cus > 0 && narrow > cus * Zn::per_cu;
}

int zpass_variant(const pgj_grid& g, int L, int es) {
  if (static_cube(g) && g.nx == 64 && fft_plan(64, es, 1) == 2) return kZReg;
  return zpass_wide(g, L, es) ? 1 : 0;
}

}  // namespace pgj

extern "C" int pgj_zpass_wide(const pgj_grid* g, int L, int es) {
  return g && L >= 1 && pgj::zpass_wide(*g, L, es) ? 1 : 0;
}

namespace pgj {

template <int N, typename R>
int fxyi_static_n(FFTArgsT<R>& A, int L, int mode, hipStream_t s) {
  using C = Cube<N>;
  const size_t lds = C::lds_xy(sizeof(cplx<R>));
  if (mode == kFxyiPre) {
    if (int e = allow_lds_static(fxyi_static<N, kFxyiPre, R>, lds)) return e;
    hipLaunchKernelGGL((fxyi_static<N, kFxyiPre, R>), dim3(N, 3, L), dim3(C::TXY), lds, s, A);
  } else if (fft_plan(N, int(sizeof(R)), 2) == 1) {
    if (int e = allow_lds_static(fxyi_static<N, kFxyiPCG, R, 1>, lds)) return e;
    hipLaunchKernelGGL((fxyi_static<N, kFxyiPCG, R, 1>), dim3(N, 3, L), dim3(C::TXY), lds, s, A);
  } else {
    if (int e = allow_lds_static(fxyi_static<N, kFxyiPCG, R>, lds)) return e;
    hipLaunchKernelGGL((fxyi_static<N, kFxyiPCG, R>), dim3(N, 3, L), dim3(C::TXY), lds, s, A);
  }
  PGJ_LAUNCHED("fxyi_static");
  return PGJ_OK;
}

bool static_fft(const pgj_grid& g) { return static_cube(g); }

#define PGJ_STATIC_SWITCH(CALL)                                       \
  switch (A.nx) {                                                     \
    case 16: return CALL(16);                                         \
    case 24: return CALL(24);                                         \
    case 32: return CALL(32);                                         \
    case 48: return CALL(48);                                         \
    case 64: return CALL(64);                                         \
    case 96: return CALL(96);                                         \
    case 128: return CALL(128);                                       \
    default: set_error("no static FFT for n=%d", A.nx); return PGJ_ERR_UNSUPPORTED; \
  }

template <typename R>
int launch_fxy_static(FFTArgsT<R>& A, int L, int mode, hipStream_t s) {
#define CALL(n) fxy_static_n<n, R>(A, L, mode, s)
  PGJ_STATIC_SWITCH(CALL)
#undef CALL
}

template <typename R>
int launch_fz_static(FFTArgsT<R>& A, int L, int mode, hipStream_t s) {
#define CALL(n) fz_static_n<n, R>(A, L, mode, s)
  PGJ_STATIC_SWITCH(CALL)
#undef CALL
}

template <typename R>
int launch_fxyi_static(FFTArgsT<R>& A, int L, hipStream_t s, int mode) {
#define CALL(n) fxyi_static_n<n, R>(A, L, mode, s)
  PGJ_STATIC_SWITCH(CALL)
#undef CALL
}

template int launch_fxy_static<double>(FFTArgsT<double>&, int, int, hipStream_t);
template int launch_fz_static<double>(FFTArgsT<double>&, int, int, hipStream_t);
template int launch_fxyi_static<double>(FFTArgsT<double>&, int, hipStream_t, int);
template int launch_fxy_static<float>(FFTArgsT<float>&, int, int, hipStream_t);
template int launch_fz_static<float>(FFTArgsT<float>&, int, int, hipStream_t);
template int launch_fxyi_static<float>(FFTArgsT<float>&, int, hipStream_t, int);

}  // namespace pgj
