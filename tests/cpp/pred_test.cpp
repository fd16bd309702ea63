// Reads cases from stdin: "O x0 y0 z0 ... (4 points)" or "S ... (5 points)"; prints the sign of
// orient / insphere5 (unperturbed) per line, for comparison with exact rational arithmetic.
#include <cstdio>

#include "../../fomo_amd/csrc/predicates.h"

int main() {
  char kind;
  using fomo::pred::P3;
  while (std::scanf(" %c", &kind) == 1) {
    const int np = kind == 'O' ? 4 : 5;
    P3 p[5];
    for (int i = 0; i < np; ++i)
      if (std::scanf("%lf %lf %lf", &p[i].x, &p[i].y, &p[i].z) != 3) return 2;
    const int s = kind == 'O' ? fomo::pred::orient(p[0], p[1], p[2], p[3])
                              : fomo::pred::insphere5(p[0], p[1], p[2], p[3], p[4]);
    std::printf("%d\n", s);
  }
  return 0;
}
