// Error reporting and version for libpgj.
#include <cstdarg>
#include <cstdio>
#include <map>
#include <mutex>
#include <string>
#include <tuple>

#include "layout.h"
#include "pgj_internal.h"

namespace pgj {
namespace {
thread_local std::string g_error;
}

void set_error(const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_error = buf;
}

void clear_error() { g_error.clear(); }

namespace {
using ShapeKey = std::tuple<int, int, int, int, int>;  // nz, ny, nx, L, element bytes
std::mutex g_shape_mutex;
std::map<ShapeKey, std::tuple<int, int, int, int>> g_shapes;
}  // namespace

namespace {
using PlanKey = std::tuple<int, int, int>;  // cube n, element bytes, pass (0 forward xy, 1 z, 2 inverse xy)
std::mutex g_plan_mutex;
std::map<PlanKey, int> g_plans;
}  // namespace

int fft_plan(int n, int es, int pass) {
  std::lock_guard<std::mutex> lock(g_plan_mutex);
  auto it = g_plans.find(PlanKey(n, es, pass));
  return it == g_plans.end() ? 0 : it->second;
}

bool tuned_shape(const pgj_grid& g, int L, int es, int* R, int* W, int* ZC, int* plane) {
  std::lock_guard<std::mutex> lock(g_shape_mutex);
  auto it = g_shapes.find(ShapeKey(g.nz, g.ny, g.nx, L, es));
  if (it == g_shapes.end()) return false;
  std::tie(*R, *W, *ZC, *plane) = it->second;
  return true;
}

}  // namespace pgj

extern "C" int pgj_hex8_set_shape(const pgj_grid* g, int L, int es, int R, int W, int ZC, int plane) {
  PGJ_REQUIRE(g && L >= 1 && (es == 4 || es == 8), "invalid shape key");
  std::lock_guard<std::mutex> lock(pgj::g_shape_mutex);
  const pgj::ShapeKey key(g->nz, g->ny, g->nx, L, es);
  if (R == 0 && W == 0 && ZC == 0) {
    pgj::g_shapes.erase(key);
    return PGJ_OK;
  }
  PGJ_REQUIRE((R == 1 || R == 2) && W >= 1 && W <= 8 && ZC >= 1 && ZC <= g->nz && (plane == 0 || plane == 1),
              "row-march shape needs R in {1,2}, W in 1..8, ZC in 1..nz, plane in {0,1}, got (%d,%d,%d,%d)", R, W,
              ZC, plane);
  pgj::g_shapes[key] = std::make_tuple(R, W, ZC, plane);
  return PGJ_OK;
}

extern "C" int pgj_hex8_get_shape(const pgj_grid* g, int L, int es, int* out5) {
  PGJ_REQUIRE(g && L >= 1 && (es == 4 || es == 8) && out5, "invalid shape query");
  const pgj::RowMarch c = pgj::rowmarch_config(*g, L, es);
  PGJ_REQUIRE(c.valid, "grid has no row-march shape");
  out5[0] = c.R; out5[1] = c.W; out5[2] = c.ZC; out5[3] = c.plane; out5[4] = c.by * c.bz;
  return PGJ_OK;
}

extern "C" int pgj_fft_set_plan(const pgj_grid* g, int es, int pass, int plan) {
  PGJ_REQUIRE(g && g->nx == g->ny && g->ny == g->nz && (es == 4 || es == 8) && pass >= 0 && pass <= 2 &&
                  (plan == -1 || plan == 0 || plan == 1 || (plan == 2 && pass == 1 && g->nx == 64)),
              "FFT plan needs a cube, es in {4,8}, pass in 0..2 and plan in {-1,0,1} (plan 2: the z pass at 64^3)");
  std::lock_guard<std::mutex> lock(pgj::g_plan_mutex);
  if (plan < 0)
    pgj::g_plans.erase(pgj::PlanKey(g->nx, es, pass));
  else
    pgj::g_plans[pgj::PlanKey(g->nx, es, pass)] = plan;
  return PGJ_OK;
}

extern "C" int pgj_fft_get_plan(const pgj_grid* g, int es, int* out3) {
  PGJ_REQUIRE(g && (es == 4 || es == 8) && out3, "invalid FFT plan query");
  for (int pass = 0; pass < 3; ++pass) out3[pass] = pgj::fft_plan(g->nx, es, pass);
  return PGJ_OK;
}

extern "C" int pgj_version(void) { return 1; }

extern "C" const char* pgj_last_error(void) { return pgj::g_error.c_str(); }
