// Nearest point of a rectilinear point set: every point lies on a product grid xs × ys × zs of
// distinct axis values, and every combination occurs exactly once (structured simulation output).
// Such a set needs no candidate index. The squared distance the engine compares,
// (d0^2 + d1^2) + d2^2 in double from float coordinates and a float target, is non-decreasing in
// each axis term, because rounded addition is monotone. So the nearest point combines, along
// each axis, values whose term is the smallest or within rounding of it, and an axis's terms
// only grow away from the target. Shared by the device kernels (fomo_common.h) and a host unit
// test, which checks it against brute force.
//
// Per axis the device keeps the values and a bin table: bins of equal width over
// [v[0], v[n-1]], each holding {i, v[i], v[i+1], v[i+2]}. Here i is the largest index whose value
// is <= the bin's start (0 when none), and +inf stands for values past the end. With twice as
// many bins as values, the table usually holds the two values bracketing a target, so a query
// reads one 16-byte entry per axis. The winner's record is the only read that leaves the cache.
#pragma once

#include <cmath>
#include <cstdint>

#include "octree.h"

namespace fomo {

// Bins per axis value.
constexpr int kRectBinsPerValue = 2;
// An axis value two steps from the target can tie the nearest one only when the axis's smallest
// gap g satisfies g^2 <= ~1.1e-15 D (D: the smallest squared distance; the bound allows for the
// rounding of the differences, squares and sums). Below that tolerance the query gives the
// target to the exact search.
constexpr double kRectGapTol = 1e-14;

// Host: the bin table of an axis of n >= 1 distinct ascending values, kRectBinsPerValue * n
// entries of type F4 (fields x, y, z, w), with the bins' origin o, inverse width inv and the
// square of the smallest gap between consecutive values, rounded down (+inf for n = 1).
template <class F4>
inline void rect_axis_table(const float* v, int n, F4* bins, double* o, double* inv, double* gap2) {
  const int nb = kRectBinsPerValue * n;
  const double lo = v[0], w = ((double)v[n - 1] - lo) / nb;
  *o = lo;
  *inv = w > 0 ? 1.0 / w : 0.0;
  int i = 0;
  for (int b = 0; b < nb; ++b) {
    const double start = lo + b * w;
    while (i + 1 < n && (double)v[i + 1] <= start) ++i;
    bins[b].x = i2f(i);
    bins[b].y = v[i];
    bins[b].z = i + 1 < n ? v[i + 1] : INFINITY;
    bins[b].w = i + 2 < n ? v[i + 2] : INFINITY;
  }
  double g2 = INFINITY;
  for (int k = 0; k + 1 < n; ++k) {
    const double g = (double)v[k + 1] - (double)v[k];
    g2 = std::fmin(g2, g * g);
  }
  *gap2 = g2 * (1.0 - 1e-9);
}

// Bin of target t on an axis whose bins start at o with inverse width inv.
FOMO_HD inline int rect_bin(double o, double inv, int nb, float t) {
  const double b = std::floor(((double)t - o) * inv);
  if (!(b >= 0.0)) return 0;
  if (b >= (double)(nb - 1)) return nb - 1;
  return (int)b;
}

// Largest index lo with v[lo] <= t (-1 when none) from the entry e = {i, v[i], v[i+1], v[i+2]}
// of t's bin; ld(k) reads v[k]. t is finite.
template <class F4, class Ld>
FOMO_HD inline int rect_floor(const F4& e, int n, float t, const Ld& ld) {
  const int i = f2i(e.x);
  if (t < e.y) {
    // below v[i]: the largest lo in [-1, i) (bins are conservative, so this is rare)
    int a = -1, b = i;
    while (b - a > 1) {
      const int m = (a + b) >> 1;
      if (ld(m) <= t) a = m;
      else b = m;
    }
    return a;
  }
  if (t < e.z) return i;
  if (t < e.w) return i + 1;
  // v[i+2] <= t (i + 2 < n, since past the end the entry holds +inf)
  int a = i + 2, b = n;
  while (b - a > 1) {
    const int m = (a + b) >> 1;
    if (ld(m) <= t) a = m;
    else b = m;
  }
  return a;
}

// The nearest point to the search box's target: returns its grid position gid = i + n0 (j + n1 k)
// when it lies strictly inside the open search box, else -2; also -2 (the caller then runs the
// exact query) for targets that are not finite and when a value two steps from the target could
// tie (an axis's smallest gap tiny against the distance). e: the three bin entries of the target;
// n: axis lengths; gap2: per axis the square of the smallest gap between consecutive values,
// rounded down (+inf for one value); ld(a, k): value k of axis a; orig(gid): original index of a
// point (read only on an exact distance tie, lowest wins).
//
// Along each axis the two values bracketing the target are the only candidates: the nearer one
// (term eb) and, when its term eo still gives the total D = (eb0 + eb1) + eb2, the other one. D is
// the smallest distance (every combination's total is >= D, rounded sums being monotone), so the
// answer is the combination with total D and the lowest original index.
// One axis of rect_pick: the value nearer to t among the two bracketing it (index kb, value vb,
// squared difference eb) and the other one (ko, vo, eo; eo = +inf when t lies beyond the last or
// before the first value). ld(k) reads value k.
template <class F4, class Ld>
FOMO_HD inline void rect_axis(const F4& e, int n, float t, const Ld& ld, int& kb, int& ko, float& vb, float& vo,
                              double& eb, double& eo) {
  const int i = f2i(e.x);
  const int lo = rect_floor(e, n, t, ld);
  float vl = 0.f, vr = 0.f;
  double el = INFINITY, er = INFINITY;
  if (lo >= 0) {
    vl = lo == i ? e.y : (lo == i + 1 ? e.z : ld(lo));
    const double d = (double)vl - (double)t;
    el = d * d;
  }
  if (lo + 1 < n) {
    vr = lo == i ? e.z : (lo == i + 1 ? e.w : (lo + 1 == i ? e.y : ld(lo + 1)));
    const double d = (double)vr - (double)t;
    er = d * d;
  }
  const bool left = el <= er;
  kb = left ? lo : lo + 1;
  ko = left ? lo + 1 : lo;
  vb = left ? vl : vr;
  vo = left ? vr : vl;
  eb = left ? el : er;
  eo = left ? er : el;
}

template <class F4, class Ld, class Orig>
FOMO_HD inline int rect_pick(const F4& e0, const F4& e1, const F4& e2, int n0, int n1, int n2, double g20, double g21,
                             double g22, const SearchBox& box, const Ld& ld, const Orig& orig) {
  if (!(std::isfinite(box.t0) && std::isfinite(box.t1) && std::isfinite(box.t2))) return -2;
  int kb0, kb1, kb2, ko0, ko1, ko2;
  float vb0, vb1, vb2, vo0, vo1, vo2;
  double eb0, eb1, eb2, eo0, eo1, eo2;
  rect_axis(e0, n0, box.t0, [&](int k) { return ld(0, k); }, kb0, ko0, vb0, vo0, eb0, eo0);
  rect_axis(e1, n1, box.t1, [&](int k) { return ld(1, k); }, kb1, ko1, vb1, vo1, eb1, eo1);
  rect_axis(e2, n2, box.t2, [&](int k) { return ld(2, k); }, kb2, ko2, vb2, vo2, eb2, eo2);
  const double D = (eb0 + eb1) + eb2;
  if (!(g20 > kRectGapTol * D && g21 > kRectGapTol * D && g22 > kRectGapTol * D)) return -2;
  const bool two0 = (eo0 + eb1) + eb2 <= D;
  const bool two1 = (eb0 + eo1) + eb2 <= D;
  const bool two2 = (eb0 + eb1) + eo2 <= D;
  int64_t g = kb0 + (int64_t)n0 * (kb1 + (int64_t)n1 * kb2);
  float q0 = vb0, q1 = vb1, q2 = vb2;
  if (two0 || two1 || two2) {
    int bo = orig(g);
    for (int m = 1; m < 8; ++m) {
      const bool u = (m & 1) != 0, v = (m & 2) != 0, w = (m & 4) != 0;
      if ((u && !two0) || (v && !two1) || (w && !two2)) continue;
      const double d = ((u ? eo0 : eb0) + (v ? eo1 : eb1)) + (w ? eo2 : eb2);
      if (!(d <= D)) continue;
      const int64_t gm = (u ? ko0 : kb0) + (int64_t)n0 * ((v ? ko1 : kb1) + (int64_t)n1 * (w ? ko2 : kb2));
      const int om = orig(gm);
      if (om < bo) {
        bo = om;
        g = gm;
        q0 = u ? vo0 : vb0;
        q1 = v ? vo1 : vb1;
        q2 = w ? vo2 : vb2;
      }
    }
  }
  // strictly inside the search box (the lo < q < hi test of every query)
  if (box.lo0 < q0 && q0 < box.hi0 && box.lo1 < q1 && q1 < box.hi1 && box.lo2 < q2 && q2 < box.hi2) return (int)g;
  return -2;
}

}  // namespace fomo
