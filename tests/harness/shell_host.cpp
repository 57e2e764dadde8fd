// Host run of the device ANDES4 element code (csrc/shell_andes4.hpp): the
// kernel's phases executed lane by lane, in phase order.  Reads elements
// from stdin, one per line: E nu rho t1 t2 t3 t4 x1..x4 y1..y4 z1..z4;
// prints per element: bad, then the unprojected global stiffness (576,
// row-major), the projector P (576, row-major), the packed projected
// blocks (360) and the lumped-mass diagonal (24), all as %.17e.
#define SHELL_HD
#include "../../fedem-solvers_amd/csrc/shell_andes4.hpp"
#include <cstdio>
#include <cstring>

int main() {
  using namespace fgpu::shell;
  double v[19];
  while (true) {
    for (int k = 0; k < 19; k++)
      if (std::scanf("%lf", &v[k]) != 1) return 0;
    static Work w;
    std::memset(&w, 0, sizeof w);
    w.E = v[0]; w.nu = v[1]; w.rho = v[2];
    w.thk = (v[3] + v[4] + v[5] + v[6]) * 0.25;
    for (int n = 0; n < 4; n++)
      for (int c = 0; c < 3; c++) w.G[n][c] = v[7 + 4 * c + n];
    w.bad = 0;
    for (int k = 0; k < NPHASE; k++)
      for (int lane = 0; lane < 64; lane++) phase(k, lane, w);
    std::printf("%d\n", w.bad);
    for (int i = 0; i < 24; i++)
      for (int j = 0; j < 24; j++) std::printf("%.17e\n", w.K[i][j]);
    for (int i = 0; i < 24; i++)
      for (int j = 0; j < 24; j++) std::printf("%.17e\n", w.P[i][j]);
    for (int idx = 0; idx < 360; idx++) {
      int r, c;
      std::printf("%.17e\n", projected_entry(idx, w, r, c));
    }
    for (int k = 0; k < 24; k++) std::printf("%.17e\n", w.em[k]);
  }
}
