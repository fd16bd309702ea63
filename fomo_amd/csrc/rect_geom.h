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
// [v[0], v[n-1]], each holding {i, v[i], v[i+1], v[i+2]}. Here i is the last value in an earlier
// bin (0 when none), and +inf stands for values past the end. With twice as many bins as values,
// the entry usually holds the two values bracketing a target, so a query reads one 16-byte entry
// per axis and no value. The winner's record is the only read that leaves the cache.
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

// Bin of target t on an axis whose bins start at o with inverse width inv, in float (the same
// arithmetic on the host, which builds the tables, and on the device): floor((t - o) inv),
// clamped to [0, nb). Monotone in t.
FOMO_HD inline int rect_bin(float o, float inv, int nb, float t) {
  const float b = std::floor((t - o) * inv);
  return (int)std::fmin(std::fmax(b, 0.f), (float)(nb - 1));
}

// Host: the bin table of an axis of n >= 1 distinct ascending values, kRectBinsPerValue * n
// entries of type F4 (fields x, y, z, w), with the bins' origin o and inverse width inv, and the
// square of the smallest gap between consecutive values, rounded down (+inf for n = 1). Entry b
// holds {i, v[i], v[i+1], v[i+2]}: i is the largest index whose value lies in a bin before b
// (0 when none), so v[i] < t for every t of bin b unless i is 0 and t < v[0].
template <class F4>
inline void rect_axis_table(const float* v, int n, F4* bins, float* o, float* inv, double* gap2) {
  const int nb = kRectBinsPerValue * n;
  const float range = v[n - 1] - v[0];
  *o = v[0];
  *inv = range > 0.f ? (float)nb / range : 0.f;
  int i = 0, k = 0;
  for (int b = 0; b < nb; ++b) {
    while (k < n && rect_bin(*o, *inv, nb, v[k]) < b) i = k++;
    bins[b].x = i2f(i);
    bins[b].y = v[i];
    bins[b].z = i + 1 < n ? v[i + 1] : INFINITY;
    bins[b].w = i + 2 < n ? v[i + 2] : INFINITY;
  }
  double g2 = INFINITY;
  for (int k2 = 0; k2 + 1 < n; ++k2) {
    const double g = (double)v[k2 + 1] - (double)v[k2];
    g2 = std::fmin(g2, g * g);
  }
  *gap2 = g2 * (1.0 - 1e-9);
}

// The largest smallest distance D for which no axis value two steps from a target can tie
// (kRectGapTol), from the three axes' squared smallest gaps.
inline double rect_dmax(const double (&gap2)[3]) {
  return std::fmin(std::fmin(gap2[0], gap2[1]), gap2[2]) / kRectGapTol;
}

// Largest index lo with v[lo] <= t (-1 when none) by bisection of v[a, b): v[a] <= t (a = -1:
// none) and v[b] > t (b = n: none); ld(k) reads v[k].
template <class Ld>
FOMO_HD inline int rect_bisect(int a, int b, float t, const Ld& ld) {
  while (b - a > 1) {
    const int m = (a + b) >> 1;
    if (ld(m) <= t) a = m;
    else b = m;
  }
  return a;
}

// One axis of rect_pick: the value nearer to t among the two bracketing it (index kb, value vb,
// squared difference eb) and the other one (ko, vo, eo; eo = +inf when t lies beyond the last or
// before the first value). e = {i, v[i], v[i+1], v[i+2]} is the entry of t's bin; ld(k) reads
// value k, only when t lies outside [v[i], v[i+2]) (a bin holding several values, or t beyond
// the axis's ends).
template <class F4, class Ld>
FOMO_HD inline void rect_axis(const F4& e, int n, float t, const Ld& ld, int& kb, int& ko, float& vb, float& vo,
                              double& eb, double& eo) {
  const int i = f2i(e.x);
  int lo;
  float vl, vr;
  if (t >= e.y && t < e.w) {
    // lo is i or i + 1: both bracketing values are in the entry
    const bool up = t >= e.z;
    lo = up ? i + 1 : i;
    vl = up ? e.z : e.y;
    vr = up ? e.w : e.z;
  } else {
    lo = t < e.y ? (i == 0 ? -1 : rect_bisect(-1, i, t, ld)) : rect_bisect(i + 2, n, t, ld);
    vl = lo >= 0 ? ld(lo) : 0.f;
    vr = lo + 1 < n ? ld(lo + 1) : 0.f;
  }
  double el = INFINITY, er = INFINITY;
  if (lo >= 0) {
    const double d = (double)vl - (double)t;
    el = d * d;
  }
  if (lo + 1 < n) {
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

// The nearest point to the search box's target: returns its grid position gid = i + n0 (j + n1 k)
// when it lies strictly inside the open search box, else -2; also -2 (the caller then runs the
// exact query) for targets that are not finite and when a value two steps from the target could
// tie (D > dmax: an axis's smallest gap tiny against the distance). e0, e1, e2: the bin entries of
// the target; n0, n1, n2: axis lengths; dmax: rect_dmax of the axes; ld(a, k): value k of axis a;
// orig(gid): original index of a point (read only on an exact distance tie, lowest wins).
//
// Along each axis the two values bracketing the target are the only candidates: the nearer one
// (term eb) and, when its term eo still gives the total D = (eb0 + eb1) + eb2, the other one. D is
// the smallest distance (every combination's total is >= D, rounded sums being monotone), so the
// answer is the combination with total D and the lowest original index.
template <class F4, class Ld, class Orig>
FOMO_HD inline int rect_pick(const F4& e0, const F4& e1, const F4& e2, int n0, int n1, int n2, double dmax,
                             const SearchBox& box, const Ld& ld, const Orig& orig) {
  if (!(std::isfinite(box.t0) && std::isfinite(box.t1) && std::isfinite(box.t2))) return -2;
  int kb0, kb1, kb2, ko0, ko1, ko2;
  float vb0, vb1, vb2, vo0, vo1, vo2;
  double eb0, eb1, eb2, eo0, eo1, eo2;
  rect_axis(e0, n0, box.t0, [&](int k) { return ld(0, k); }, kb0, ko0, vb0, vo0, eb0, eo0);
  rect_axis(e1, n1, box.t1, [&](int k) { return ld(1, k); }, kb1, ko1, vb1, vo1, eb1, eo1);
  rect_axis(e2, n2, box.t2, [&](int k) { return ld(2, k); }, kb2, ko2, vb2, vo2, eb2, eo2);
  const double D = (eb0 + eb1) + eb2;
  if (!(D <= dmax)) return -2;
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

// Masked grids: the points are a subset of the product grid, each on its own node; gp(g) gives the
// point on grid node g (its sorted position) or -1, orig(sp) the original index of a point. The
// nearest present point is the nearest node when that node holds a point: every node, present or
// not, is at least D away, so a present node at distance D is nearest, ties going to the lowest
// original index among the present nodes at D. Returns its sorted position when it lies strictly
// inside the search box, -2 when it lies outside (the caller runs the exact query), and -3 when no
// node at distance D holds a point (or D > dmax, or the target is not finite): then the answer is
// farther away and the caller takes another path.
template <class F4, class Ld, class Gp, class Orig>
FOMO_HD inline int mrect_pick(const F4& e0, const F4& e1, const F4& e2, int n0, int n1, int n2, double dmax,
                              const SearchBox& box, const Ld& ld, const Gp& gp, const Orig& orig) {
  if (!(std::isfinite(box.t0) && std::isfinite(box.t1) && std::isfinite(box.t2))) return -3;
  int kb0, kb1, kb2, ko0, ko1, ko2;
  float vb0, vb1, vb2, vo0, vo1, vo2;
  double eb0, eb1, eb2, eo0, eo1, eo2;
  rect_axis(e0, n0, box.t0, [&](int k) { return ld(0, k); }, kb0, ko0, vb0, vo0, eb0, eo0);
  rect_axis(e1, n1, box.t1, [&](int k) { return ld(1, k); }, kb1, ko1, vb1, vo1, eb1, eo1);
  rect_axis(e2, n2, box.t2, [&](int k) { return ld(2, k); }, kb2, ko2, vb2, vo2, eb2, eo2);
  const double D = (eb0 + eb1) + eb2;
  if (!(D <= dmax)) return -3;
  const bool two0 = (eo0 + eb1) + eb2 <= D;
  const bool two1 = (eb0 + eo1) + eb2 <= D;
  const bool two2 = (eb0 + eb1) + eo2 <= D;
  int best = gp(kb0 + (int64_t)n0 * (kb1 + (int64_t)n1 * kb2));
  float q0 = vb0, q1 = vb1, q2 = vb2;
  if (two0 || two1 || two2) {
    int bo = -1;  // original index of best (-1: not read yet)
    for (int m = 1; m < 8; ++m) {
      const bool u = (m & 1) != 0, v = (m & 2) != 0, w = (m & 4) != 0;
      if ((u && !two0) || (v && !two1) || (w && !two2)) continue;
      const double d = ((u ? eo0 : eb0) + (v ? eo1 : eb1)) + (w ? eo2 : eb2);
      if (!(d <= D)) continue;
      const int sp = gp((u ? ko0 : kb0) + (int64_t)n0 * ((v ? ko1 : kb1) + (int64_t)n1 * (w ? ko2 : kb2)));
      if (sp < 0) continue;
      bool take = best < 0;
      if (!take) {
        if (bo < 0) bo = orig(best);
        const int o = orig(sp);
        take = o < bo;
        if (take) bo = o;
      }
      if (take) {
        best = sp;
        q0 = u ? vo0 : vb0;
        q1 = v ? vo1 : vb1;
        q2 = w ? vo2 : vb2;
      }
    }
  }
  if (best < 0) return -3;
  if (box.lo0 < q0 && q0 < box.hi0 && box.lo1 < q1 && q1 < box.hi1 && box.lo2 < q2 && q2 < box.hi2) return best;
  return -2;
}

}  // namespace fomo
