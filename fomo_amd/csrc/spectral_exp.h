// Exponentials for the recurrence anchors of the Gaussian line profile (host and device).
//
// The LOS integration evaluates, per sample, exp of three arguments that are finite and lie in
// (-709, 709). The library exp evaluates its polynomial by Horner's rule, a chain of a dozen
// dependent FP64 operations, and three such chains in a row dominate the latency of the anchor
// step. Here: x = k ln2 + r with k = rint(x / ln2), |r| <= ln2/2 (two-term ln2, FMA reduction),
// exp(r) by its Taylor polynomial of degree 11 (truncation error below 7e-15 relative at
// |r| = ln2/2), split into even and odd parts (dependency depth 7 instead of 11), and the three
// evaluations side by side so their chains overlap. Measured error: tests/cpp/spectral_exp_test.cpp.
#pragma once

#include <cmath>

#ifndef FOMO_HD
#if defined(__HIPCC__)
#define FOMO_HD __host__ __device__
#else
#define FOMO_HD
#endif
#endif

namespace fomo {

// ln 2 = kLn2Hi + kLn2Lo (kLn2Hi = ln 2 rounded to double, kLn2Lo = the rest rounded to double),
// 1 / ln 2, and 1/i! for i = 2..11, each rounded to double.
constexpr double kLog2E = 1.4426950408889634;
constexpr double kLn2Hi = 0.6931471805599453;
constexpr double kLn2Lo = 2.3190468138462996e-17;
constexpr double kInvFact[12] = {1.0,
                                 1.0,
                                 0.5,
                                 0.16666666666666666,
                                 0.041666666666666664,
                                 0.008333333333333333,
                                 0.001388888888888889,
                                 0.0001984126984126984,
                                 2.48015873015873e-05,
                                 2.7557319223985893e-06,
                                 2.755731922398589e-07,
                                 2.505210838544172e-08};

// e[i] = exp(x[i]) for three arguments in (-709, 709). The polynomial is split into its even
// and odd parts in s = r^2, p(r) = E(s) + r O(s), each by Horner's rule: six short chains across
// the three arguments, with few values live at once.
FOMO_HD inline void exp3(const double (&x)[3], double (&e)[3]) {
  double k[3], r[3], s[3], ev[3], od[3];
  for (int i = 0; i < 3; ++i) k[i] = rint(x[i] * kLog2E);
  for (int i = 0; i < 3; ++i) r[i] = fma(-k[i], kLn2Hi, x[i]);
  for (int i = 0; i < 3; ++i) r[i] = fma(-k[i], kLn2Lo, r[i]);
  for (int i = 0; i < 3; ++i) s[i] = r[i] * r[i];
  for (int i = 0; i < 3; ++i) {
    ev[i] = fma(kInvFact[10], s[i], kInvFact[8]);
    od[i] = fma(kInvFact[11], s[i], kInvFact[9]);
  }
  for (int j = 6; j >= 0; j -= 2)
    for (int i = 0; i < 3; ++i) {
      ev[i] = fma(ev[i], s[i], kInvFact[j]);
      od[i] = fma(od[i], s[i], kInvFact[j + 1]);
    }
  for (int i = 0; i < 3; ++i) e[i] = ldexp(fma(r[i], od[i], ev[i]), (int)k[i]);
}

// 4 ln2 / w^2 from a reciprocal and two Newton steps (within 2 ulp of the division).
FOMO_HD inline double spectral_kappa(double w) {
  const double w2 = w * w;
#if defined(__HIP_DEVICE_COMPILE__)
  double y = __builtin_amdgcn_rcp(w2);
#else
  double y = 1.0 / w2;
#endif
  y = fma(fma(-w2, y, 1.0), y, y);
  y = fma(fma(-w2, y, 1.0), y, y);
  return (4.0 * 0.6931471805599453) * y;
}

}  // namespace fomo
