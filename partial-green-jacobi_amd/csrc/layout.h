// Launch configurations and the byte layouts of the caller-owned plan,
// workspace and resident-state buffers. Host and device agree on these
// through the structs below (they are passed to kernels by value).
#pragma once

#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <cstdlib>

#include <algorithm>

#include "../../include/pgj.h"

namespace pgj {

constexpr int kCap = 4000;   // default PCG iteration cap (exp/asg_gpu_resident_20261103/resident.py:50)

inline size_t align256(size_t v) { return (v + 255) & ~size_t(255); }
inline size_t align8(size_t v) { return (v + 7) & ~size_t(7); }

// ---------------------------------------------------------------- row-march
// The HEX8 row-march kernel maps x to lanes (XPL consecutive x per lane), a
// wavefront marches R element rows over the z-layers of a block, a block has
// W row-waves plus one halo wave and covers W*R node rows and ZC planes.
struct RowMarch {
  int valid;
  int xpl, nl;      // x per lane, active lanes = nx / xpl
  int R, W, ZC;
  int plane;        // 1: plane-carry march (shared in-plane butterflies), 0: node-carry march
  int by, bz;       // blocks along y and z per load
};

// Tuned row-march shapes, keyed by (nz, ny, nx, L, element bytes); set by
// pgj_hex8_set_shape (the autotuner) and consulted by every launch.
// Returns false if unset.
bool tuned_shape(const pgj_grid& g, int L, int es, int* R, int* W, int* ZC, int* plane);

// Radix plan of the compile-time FFT pass `pass` (0 forward xy, 1 z with the Green multiply, 2 inverse xy) of
// the PCG iteration on an n^3 cube in precision es: 0 (radix 4 first, the default) or 1 (radix 8 first); set by
// pgj_fft_set_plan (the autotuner) and read at every launch of that pass.
int fft_plan(int n, int es, int pass);

// Shape preferences (R, W, ZC) for L loads per launch. ZC (planes marched by
// one block) is the largest of 16, 8, 4, 2, 1 dividing nz that still gives
// at least 192 blocks: longer marches re-read fewer halo planes, but the
// launch must keep most of the 256 CUs busy. A tuned shape replaces the
// preferences.
inline RowMarch rowmarch_config(const pgj_grid& g, int L, int es = 8) {
  RowMarch c{};
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  c.xpl = (nx + 63) / 64;
  if (c.xpl > 2 || nx % c.xpl != 0 || ny < 2 || nz < 2 || L < 1) return c;
  c.nl = nx / c.xpl;
  // Default shape (measured on the PCG pass over 32^3..128^3, 1 and 6 loads):
  // the plane-carry march with two row waves (four for FP32 with one x per
  // lane), except FP64 with two x per lane, which runs the node march.
  int want_r = 1, want_w = (es == 4 && c.xpl == 1) ? 4 : 2, want_z = 0;
  int want_plane = (es == 4 || c.xpl == 1) ? 1 : 0;
  {
    int r = 0, w = 0, z = 0, p = 0;
    if (tuned_shape(g, L, es, &r, &w, &z, &p)) {
      want_r = r; want_w = w; want_z = z; want_plane = p;
    }
  }
  c.plane = want_plane ? 1 : 0;
  // two rows of two x per lane, and two rows of the plane-carry march, exceed 256 VGPRs
  c.R = (want_r == 2 && ny % 2 == 0 && c.xpl == 1 && !c.plane) ? 2 : 1;
  // block = W row-waves + 1 halo wave: at most 5 waves (320 threads, the launch
  // bound of the R = 2, two-x-per-lane and FP64 plane-carry kernels) or 9 waves (576)
  const int wcap = (c.R == 2 || c.xpl == 2 || (c.plane && es == 8)) ? 4 : 8;
  const int wmax = want_w < wcap ? want_w : wcap;
  const int rows = ny / c.R;
  c.W = 1;
  for (int w = wmax; w >= 1; --w)
    if (rows % w == 0) { c.W = w; break; }
  c.by = ny / (c.W * c.R);
  c.ZC = 1;
  if (want_z > 0) {
    for (int z = want_z; z >= 1; --z)
      if (nz % z == 0) { c.ZC = z; break; }
  } else {
    // the longest march (<= 16 planes) that still launches >= 768 blocks (3 per CU), at least 2 planes
    for (int z = 16; z >= 1; z /= 2) {
      if (nz % z != 0) continue;
      c.ZC = z;
      if (int64_t(c.by) * (nz / z) * L >= 768 || z <= 2) break;
    }
  }
  c.bz = nz / c.ZC;
  c.valid = 1;
  return c;
}

inline int hex8_blocks_per_load(const pgj_grid& g, int L, int es = 8) {
  RowMarch c = rowmarch_config(g, L, es);
  if (c.valid) return c.by * c.bz;
  const int64_t n = int64_t(g.nx) * g.ny * g.nz;
  return int((n + 255) / 256);
}

// Upper bound of hex8_blocks_per_load over every shape (W = ZC = R = 1), so
// workspaces stay valid whatever shape is tuned later.
inline int hex8_blocks_bound(const pgj_grid& g) {
  const int64_t n = int64_t(g.nx) * g.ny * g.nz;
  const int64_t rm = int64_t(g.ny) * g.nz;
  const int64_t ref = (n + 255) / 256;
  return int(rm > ref ? rm : ref);
}

// --------------------------------------------------------------------- FFT
// One-dimensional transforms of length n are factored into radices 4, 2, 3, 5, 7.
constexpr int kMaxStages = 12;
struct Factors {
  int n, count;
  int r[kMaxStages];
};

inline Factors factorize(int n) {
  Factors f{};
  f.n = n;
  int m = n;
  while (m % 4 == 0 && f.count < kMaxStages) { f.r[f.count++] = 4; m /= 4; }
  while (m % 2 == 0 && f.count < kMaxStages) { f.r[f.count++] = 2; m /= 2; }
  for (int p : {3, 5, 7})
    while (m % p == 0 && f.count < kMaxStages) { f.r[f.count++] = p; m /= p; }
  if (m != 1) f.count = -1;
  return f;
}

struct FFTGeom {
  int nz, ny, nx;
  int nkx;          // nx/2 + 1
  int M;            // complex length of the x transform: nx/2 (even nx) or nx (odd)
  int rs;           // row stride of a plane row in LDS (complex): max(nkx, M)
  int64_t nf;       // nz*ny*nkx
  int comps;        // components per plane block in the xy kernels (3 or 1)
  int kc;           // kx columns per block in the z kernel
  int nkb;          // ceil(nkx/kc)
  int threads_xy, threads_z;
  size_t lds_xy, lds_z;
  Factors fx, fy, fz;
  int valid;        // some FFT path supports the grid
  int runtime_ok;   // the runtime-shaped kernels (fft.hip) support it
};

// Cubic grids with compile-time-shaped resident FFT kernels (fft_static.hip).
inline bool static_cube(const pgj_grid& g) {
  if (g.nx != g.ny || g.ny != g.nz) return false;
  switch (g.nx) {
    case 16: case 24: case 32: case 48: case 64: case 96: case 128: return true;
    default: return false;
  }
}

constexpr size_t kLdsLimit = 160 * 1024 - 2048;
constexpr int kStageValues = 16;  // complex values a thread holds in one in-place LDS stage

// True when every stage of a batched transform fits the register budget of
// one in-place LDS stage (each thread holds kStageValues/R butterflies).
inline bool stages_fit(const Factors& f, int lines, int threads) {
  for (int s = 0; s < f.count; ++s) {
    const int R = f.r[s];
    if (int64_t(lines) * (f.n / R) > int64_t(threads) * (kStageValues / R)) return false;
  }
  return true;
}

inline FFTGeom fft_geom(const pgj_grid& g) {
  FFTGeom f{};
  f.nz = g.nz; f.ny = g.ny; f.nx = g.nx;
  f.nkx = g.nx / 2 + 1;
  f.M = (g.nx % 2 == 0) ? g.nx / 2 : g.nx;
  f.rs = f.nkx > f.M ? f.nkx : f.M;
  f.nf = int64_t(g.nz) * g.ny * f.nkx;
  f.fx = factorize(f.M);
  f.fy = factorize(g.ny);
  f.fz = factorize(g.nz);
  f.valid = f.fx.count >= 0 && f.fy.count >= 0 && f.fz.count >= 0 && g.nx >= 2 && g.ny >= 2 && g.nz >= 2 &&
            g.nx <= 256 && g.ny <= 256 && g.nz <= 256;
  const size_t plane = size_t(g.ny) * f.rs * 16;
  f.comps = (3 * plane + 64 <= kLdsLimit) ? 3 : 1;
  if (plane + 64 > kLdsLimit) f.valid = 0;
  f.lds_xy = f.comps * plane + 64 * 8;
  f.threads_xy = 512;
  // z kernel: kc columns x 3 comps x nz complex
  const size_t col = size_t(g.nz) * 3 * 16;
  int kc = 1;
  while (kc < f.nkx && size_t(kc + 1) * col + 64 * 8 <= 56 * 1024) ++kc;
  if (g.nx == g.ny && g.ny == g.nz) {
    // cubes: at least as many chunks as the compile-time kernels use (fft_static.hip, ZPass<N, R>::NKB; at 64^3
    // 5 x 64 rows for the 264 blocks of the register-resident pass), so the per-block partials laid out with
    // nkb below hold every kind of z pass
    switch (g.nx) {
      case 16: kc = 9; break;
      case 24: kc = 13; break;
      case 32: kc = 17; break;
      case 48: kc = 5; break;
      case 64: kc = 7; break;
      case 96: kc = 7; break;
      case 128: kc = 5; break;
      default: break;
    }
  }
  f.kc = kc;
  f.nkb = (f.nkx + kc - 1) / kc;
  f.lds_z = size_t(kc) * col + 64 * 8;
  f.threads_z = 256;
  if (f.valid) {
    const int rows = f.comps * g.ny;
    f.valid = stages_fit(f.fx, rows, f.threads_xy) && stages_fit(f.fy, f.comps * f.nkx, f.threads_xy) &&
              stages_fit(f.fz, 3 * f.kc, f.threads_z);
  }
  f.runtime_ok = f.valid;
  if (static_cube(g)) f.valid = 1;
  return f;
}

// Plan buffer: twiddles for x (length M), y, z; then the 6-coefficient symbol.
// rs is the real size: 8 for the FP64 plan, 4 for its FP32 copy.
struct PlanLayout {
  size_t tw_x, tw_y, tw_z, sym, bytes;
};

inline PlanLayout plan_layout(const pgj_grid& g, size_t rs = 8) {
  FFTGeom f = fft_geom(g);
  PlanLayout p{};
  size_t off = 0;
  p.tw_x = off; off = align256(off + size_t(g.nx) * 2 * rs);
  p.tw_y = off; off = align256(off + size_t(g.ny) * 2 * rs);
  p.tw_z = off; off = align256(off + size_t(g.nz) * 2 * rs);
  p.sym = off; off = align256(off + size_t(6) * f.nf * rs);
  p.bytes = off;
  return p;
}

// Preconditioner workspace (also the Green part of the resident workspace).
struct PrecondLayout {
  size_t spec;      // L x 3 x nf complex
  size_t pr;        // L x nz x 3 x 2 doubles  (per plane/component: sum r^2, sum r)
  size_t pg;        // L x nkb*ny x 4 doubles  (rz part, sum s_c w_c)
  size_t bytes;
};

// kx columns per block of the preconditioner apply's z pass: the geometry's
// chunk, or narrower when the (load, ky, chunk) blocks would not cover 512
// (few loads), which raises the block count of that pass.
inline int precond_kc(const FFTGeom& f, int ny, int L) {
  const int64_t rows = int64_t(L) * ny;
  if (rows * f.nkb >= 512) return f.kc;
  const int64_t want = std::min<int64_t>(f.nkx, (512 + rows - 1) / rows);
  return std::max(1, int((f.nkx + want - 1) / want));
}

inline PrecondLayout precond_layout(const pgj_grid& g, int L, size_t cs = 16) {
  FFTGeom f = fft_geom(g);
  PrecondLayout p{};
  size_t off = 0;
  // partial rows for the geometry's chunking and for the preconditioner apply's
  const int kc = precond_kc(f, g.ny, L);
  const int nkb = std::max(f.nkb, (f.nkx + kc - 1) / kc);
  p.spec = off; off = align256(off + size_t(L) * 3 * f.nf * cs);
  p.pr = off; off = align256(off + size_t(L) * g.nz * 6 * 8);
  p.pg = off; off = align256(off + size_t(L) * nkb * g.ny * 4 * 8);
  p.bytes = off;
  return p;
}

// Resident workspace: preconditioner part + operator partials + certify
// partials + per-load refinement scales (cs: complex size of the spectrum).
struct ResidentLayout {
  PrecondLayout pre;
  size_t ppq;       // L x hex8_blocks doubles
  size_t pcert;     // L x nblk x 4 doubles (sum u_c, |r|^2)
  size_t rscale;    // L doubles (FP32 inner solver: |r| of each load's outer residual)
  int nblk_cert;
  size_t bytes;
};

inline ResidentLayout resident_layout(const pgj_grid& g, int L, size_t cs = 16) {
  ResidentLayout r{};
  r.pre = precond_layout(g, L, cs);
  size_t off = r.pre.bytes;
  r.ppq = off; off = align256(off + size_t(L) * hex8_blocks_bound(g) * 8);
  const int64_t n = int64_t(g.nx) * g.ny * g.nz;
  r.nblk_cert = int((n + 255) / 256);
  r.pcert = off; off = align256(off + size_t(L) * r.nblk_cert * 4 * 8);
  r.rscale = off; off = align256(off + size_t(L) * 8);
  r.bytes = off;
  return r;
}

// Loads one resident launch may carry (designs x loads per design).
constexpr int kMaxResidentLoads = 4096;

// Resident state: public blocks first (documented in pgj.h), then private.
struct StateLayout {
  size_t rz, target2, norm, absolute, it, active, bad;
  size_t it_next, active_next, bad_next, first, alpha, stopped, valid, cap;
  size_t bytes;
};

inline StateLayout state_layout(int L) {
  StateLayout s{};
  size_t off = 0;
  s.rz = off; off += align8(size_t(L) * 8);
  s.target2 = off; off += align8(size_t(L) * 8);
  s.norm = off; off += align8(size_t(L) * 8);
  s.absolute = off; off += align8(size_t(L) * 8);
  s.it = off; off += align8(size_t(L) * 8);
  s.active = off; off += align8(size_t(L));
  s.bad = off; off += align8(size_t(L));
  s.it_next = off; off += align8(size_t(L) * 8);
  s.active_next = off; off += align8(size_t(L));
  s.bad_next = off; off += align8(size_t(L));
  s.first = off; off += align8(size_t(L));
  s.alpha = off; off += align8(size_t(L) * 8);
  s.stopped = off; off += align8(size_t(L));
  s.valid = off; off += align8(size_t(L));
  s.cap = off; off += align8(size_t(L) * 8);
  s.bytes = align256(off);
  return s;
}

// Device view of the resident state (pointers into one buffer).
struct StateView {
  double* rz;
  double* target2;
  double* norm;
  double* absolute;
  int64_t* it;
  uint8_t* active;
  uint8_t* bad;
  int64_t* it_next;
  uint8_t* active_next;
  uint8_t* bad_next;
  uint8_t* first;
  double* alpha;
  uint8_t* stopped;   // active at the start of this iteration and stopped by its test
  uint8_t* valid;     // this iteration's p.q and r.z were finite and positive (forward pass)
  int64_t* cap;       // iteration cap of the load (set by the set-up)
};

// View of loads [l0, ...) of a state block laid out for L loads.
inline StateView state_view(void* base, int L, int l0 = 0) {
  StateLayout s = state_layout(L);
  char* b = static_cast<char*>(base);
  StateView v;
  v.rz = reinterpret_cast<double*>(b + s.rz) + l0;
  v.target2 = reinterpret_cast<double*>(b + s.target2) + l0;
  v.norm = reinterpret_cast<double*>(b + s.norm) + l0;
  v.absolute = reinterpret_cast<double*>(b + s.absolute) + l0;
  v.it = reinterpret_cast<int64_t*>(b + s.it) + l0;
  v.active = reinterpret_cast<uint8_t*>(b + s.active) + l0;
  v.bad = reinterpret_cast<uint8_t*>(b + s.bad) + l0;
  v.it_next = reinterpret_cast<int64_t*>(b + s.it_next) + l0;
  v.active_next = reinterpret_cast<uint8_t*>(b + s.active_next) + l0;
  v.bad_next = reinterpret_cast<uint8_t*>(b + s.bad_next) + l0;
  v.first = reinterpret_cast<uint8_t*>(b + s.first) + l0;
  v.alpha = reinterpret_cast<double*>(b + s.alpha) + l0;
  v.stopped = reinterpret_cast<uint8_t*>(b + s.stopped) + l0;
  v.valid = reinterpret_cast<uint8_t*>(b + s.valid) + l0;
  v.cap = reinterpret_cast<int64_t*>(b + s.cap) + l0;
  return v;
}

}  // namespace pgj
