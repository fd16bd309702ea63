// Exact orientation test for the tetra walk, host and device: the sign of det[b - a; c - a; p - a]
// for float vertices a, b, c and a double point p, with no rounding anywhere.
//
// Multilinearity in the rows gives det(b - a, c - a, p - a) = det(b, c, p) - det(a, c, p)
// - det(b, a, p) - det(b, c, a) (every term with a repeated row vanishes): 24 triple products
// u_i v_j w_k. In each, u and v are float coordinates, so u_i v_j is exact in double (48 bits),
// and (u_i v_j) w_k is exactly x + y with x = fl(u_i v_j w_k), y = fma(u_i v_j, w_k, -x) (one
// correctly rounded fused multiply-add). The 48 components are summed into a floating-point
// expansion (non-overlapping components in increasing magnitude) with Shewchuk's Grow-Expansion
// (two-sum of each new value through the expansion, zero components dropped; J. R. Shewchuk,
// "Adaptive precision floating-point arithmetic and fast robust geometric predicates", 1997); the
// sign of the sum is that of the largest component.
#pragma once

#include <cmath>

#if defined(__HIPCC__)
#define FOMO_XD __host__ __device__
#else
#define FOMO_XD
#endif

namespace fomo {

FOMO_XD inline void ex_two_sum(double a, double b, double& x, double& y) {
  x = a + b;
  const double bv = x - a;
  const double av = x - bv;
  y = (a - av) + (b - bv);
}

// h[0..m) += b; returns the new length (at most m + 1)
FOMO_XD inline int ex_grow(double* h, int m, double b) {
  double q = b;
  int k = 0;
  for (int i = 0; i < m; ++i) {
    double s, t;
    ex_two_sum(q, h[i], s, t);
    if (t != 0.0) h[k++] = t;
    q = s;
  }
  if (q != 0.0 || k == 0) h[k++] = q;
  return k;
}

// h += sgn * det(u, v, w) for rows u, v with float coordinates and any row w
FOMO_XD inline int ex_det3_add(double* h, int m, const double* u, const double* v, const double* w, double sgn) {
  // the six terms of det(u, v, w): (i, j, k, sign)
  const int I[6] = {0, 0, 1, 1, 2, 2}, J[6] = {1, 2, 0, 2, 0, 1}, K[6] = {2, 1, 2, 0, 1, 0};
  const double S[6] = {1.0, -1.0, -1.0, 1.0, 1.0, -1.0};
  for (int t = 0; t < 6; ++t) {
    const double uv = u[I[t]] * v[J[t]];  // exact: two floats
    const double x = uv * w[K[t]];
    const double y = std::fma(uv, w[K[t]], -x);
    m = ex_grow(h, m, sgn * S[t] * y);
    m = ex_grow(h, m, sgn * S[t] * x);
  }
  return m;
}

// sign det[b - a; c - a; p - a] (a, b, c: float coordinates held in doubles; p: any doubles)
FOMO_XD inline int orient_exact_fp(const double* a, const double* b, const double* c, const double* p) {
  double h[49];
  int m = 0;
  m = ex_det3_add(h, m, b, c, p, 1.0);
  m = ex_det3_add(h, m, a, c, p, -1.0);
  m = ex_det3_add(h, m, b, a, p, -1.0);
  m = ex_det3_add(h, m, b, c, a, -1.0);
  const double top = h[m - 1];
  return top > 0.0 ? 1 : (top < 0.0 ? -1 : 0);
}

}  // namespace fomo
