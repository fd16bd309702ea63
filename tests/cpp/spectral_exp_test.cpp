// Accuracy of fomo::exp3 and fomo::spectral_kappa (fomo_amd/csrc/spectral_exp.h, the code the
// integration kernels run) against long double expl and division, over the argument ranges the
// kernels use: (-600, 0] for the first anchor, (-100, 100) for the second, (-30, 0] for the third,
// and the whole (-709, 709) for good measure.
#include <cmath>
#include <cstdio>
#include <random>

#include "../../fomo_amd/csrc/spectral_exp.h"

int main() {
  std::mt19937_64 rng(7);
  const double ranges[4][2] = {{-600.0, 0.0}, {-100.0, 100.0}, {-30.0, 0.0}, {-708.0, 708.0}};
  long double worst = 0;
  for (const auto& rg : ranges) {
    std::uniform_real_distribution<double> U(rg[0], rg[1]);
    for (int n = 0; n < 2000000; ++n) {
      const double x[3] = {U(rng), U(rng), n % 97 == 0 ? 0.0 : U(rng)};
      double e[3];
      fomo::exp3(x, e);
      for (int i = 0; i < 3; ++i) {
        const long double ref = expl((long double)x[i]);
        const long double rel = fabsl((long double)e[i] - ref) / ref;
        if (rel > worst) worst = rel;
      }
    }
  }
  // exact multiples of ln2 / 2 (the reduction's ties) and small arguments
  for (int k = -2000; k <= 2000; ++k) {
    const double x[3] = {k * 0.34657359027997264, k * 1e-9, k * 0.01};
    double e[3];
    fomo::exp3(x, e);
    for (int i = 0; i < 3; ++i) {
      const long double ref = expl((long double)x[i]);
      const long double rel = fabsl((long double)e[i] - ref) / ref;
      if (rel > worst) worst = rel;
    }
  }
  std::printf("exp3 max relative error %.3Le\n", worst);
  if (!(worst < 2e-14L)) return 1;
  long double kworst = 0;
  std::mt19937_64 rng2(11);
  std::uniform_real_distribution<double> L(-6.0, 0.0);
  for (int n = 0; n < 1000000; ++n) {
    const double w = (double)(float)std::pow(10.0, L(rng2));  // fwhm values are floats
    const long double ref = (4.0L * 0.6931471805599453L) / ((long double)w * w);
    const long double rel = fabsl((long double)fomo::spectral_kappa(w) - ref) / ref;
    if (rel > kworst) kworst = rel;
  }
  std::printf("spectral_kappa max relative error %.3Le\n", kworst);
  if (!(kworst < 1e-15L)) return 1;
  std::printf("spectral exp ok\n");
  return 0;
}
