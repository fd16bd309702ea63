// Reads cases from stdin: "O x0 y0 z0 ... (4 points)", "S ... (5 points)" or "F ... (4 points,
// the first three float-valued)"; prints the sign of orient / insphere5 (unperturbed) /
// orient_exact_fp (the tetra walk's exact test) per line, for comparison with exact rational
// arithmetic.
#include <cstdio>

#include "../../fomo_amd/csrc/orient_exact.h"
#include "../../fomo_amd/csrc/predicates.h"

int main() {
  char kind;
  using fomo::pred::P3;
  while (std::scanf(" %c", &kind) == 1) {
    const int np = kind == 'S' ? 5 : 4;
    P3 p[5];
    for (int i = 0; i < np; ++i)
      if (std::scanf("%lf %lf %lf", &p[i].x, &p[i].y, &p[i].z) != 3) return 2;
    int s;
    if (kind == 'F') {
      double q[4][3];
      for (int i = 0; i < 4; ++i) q[i][0] = p[i].x, q[i][1] = p[i].y, q[i][2] = p[i].z;
      s = fomo::orient_exact_fp(q[0], q[1], q[2], q[3]);
    } else {
      s = kind == 'O' ? fomo::pred::orient(p[0], p[1], p[2], p[3])
                      : fomo::pred::insphere5(p[0], p[1], p[2], p[3], p[4]);
    }
    std::printf("%d\n", s);
  }
  return 0;
}
