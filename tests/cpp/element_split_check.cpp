// The plane-split element (plane_forward, hex8_modal_core, plane_transposed) used by the
// plane-carry row march equals the fused hex8_element bitwise, on random elements.
#include <cstdio>
#include <cmath>
#include <random>
#define PGJ_HD
#include "hex8_element.h"
using namespace pgj;
int main() {
  std::mt19937 g(1); std::uniform_real_distribution<double> U(-1, 1);
  ElementCoef c = make_coef(1/64., 1/64., 1/64., 15/26., 5/13.);
  double maxd = 0;
  for (int trial = 0; trial < 1000; ++trial) {
    double u[24], f[24]; for (auto& v : u) v = U(g);
    double a = 0.5 + 0.5 * U(g);
    hex8_element<true>(u, a, c, f);
    // plane path
    double PL[3][4], PU[3][4];
    for (int d = 0; d < 3; ++d) {
      plane_forward(u[0*3+d], u[1*3+d], u[2*3+d], u[3*3+d], PL[d]);
      plane_forward(u[4*3+d], u[5*3+d], u[6*3+d], u[7*3+d], PU[d]);
    }
    double Um[3][8], F[3][8];
    for (int d = 0; d < 3; ++d) for (int m = 0; m < 4; ++m) { Um[d][m] = PL[d][m] + PU[d][m]; Um[d][m | 4] = PU[d][m] - PL[d][m]; }
    hex8_modal_core<true>(Um, a, c, F);
    double f2[24];
    for (int d = 0; d < 3; ++d) {
      double ML[4], MU[4], fl[4], fu[4];
      for (int m = 0; m < 4; ++m) { ML[m] = F[d][m] - F[d][m | 4]; MU[m] = F[d][m] + F[d][m | 4]; }
      plane_transposed(ML, fl); plane_transposed(MU, fu);
      for (int n = 0; n < 4; ++n) { f2[3 * n + d] = fl[n]; f2[3 * (n + 4) + d] = fu[n]; }
    }
    for (int k = 0; k < 24; ++k) maxd = std::fmax(maxd, std::fabs(f[k] - f2[k]));
  }
  printf("max |diff| = %g\n", maxd);
  return maxd == 0 ? 0 : 1;
}
