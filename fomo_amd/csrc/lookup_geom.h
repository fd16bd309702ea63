// Box geometry of the candidate lookup grid: distances between an axis-aligned box (a lookup cell
// or a group of cells) and a point, and the two domination tests that prune candidate lists.
// Shared by the device build kernels (fomo_device.hip) and a host unit test; V is any type with
// float members x, y, z.
//
// A candidate list for a box B must hold every point that is the nearest (or tied nearest) to
// some target in B. A point q may be left out when another point p dominates it over B: p is
// strictly closer than q to every target in B, so q never wins. Both tests are conservative: they
// claim domination only when it holds with a relative margin of 1e-10, far above the rounding of
// the double arithmetic on float coordinates.
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

// Largest squared distance from the box [lo, hi] to q.
template <class V>
FOMO_HD inline double box_maxdist2(const double lo[3], const double hi[3], const V& q) {
  const double a0 = std::fmax(std::fabs((double)q.x - lo[0]), std::fabs((double)q.x - hi[0]));
  const double a1 = std::fmax(std::fabs((double)q.y - lo[1]), std::fabs((double)q.y - hi[1]));
  const double a2 = std::fmax(std::fabs((double)q.z - lo[2]), std::fabs((double)q.z - hi[2]));
  return a0 * a0 + a1 * a1 + a2 * a2;
}

// Smallest squared distance from the box [lo, hi] to q (0 inside).
template <class V>
FOMO_HD inline double box_mindist2(const double lo[3], const double hi[3], const V& q) {
  const double a0 = std::fmax(0.0, std::fmax(lo[0] - (double)q.x, (double)q.x - hi[0]));
  const double a1 = std::fmax(0.0, std::fmax(lo[1] - (double)q.y, (double)q.y - hi[1]));
  const double a2 = std::fmax(0.0, std::fmax(lo[2] - (double)q.z, (double)q.z - hi[2]));
  return a0 * a0 + a1 * a1 + a2 * a2;
}

// q dominated by p over the box (center c, half widths hw). With u = q - c and w = p - c,
// |t - q|^2 - |t - p|^2 = |u|^2 - |w|^2 - 2 s.(u - w) for s = t - c, smallest at the corner
// s = hw * sign(u - w): |u|^2 - |w|^2 - 2 sum hw |u - w|.
template <class V>
FOMO_HD inline bool dominated(const double c[3], const double hw[3], const V& q, const V& p) {
  const double u0 = (double)q.x - c[0], u1 = (double)q.y - c[1], u2 = (double)q.z - c[2];
  const double w0 = (double)p.x - c[0], w1 = (double)p.y - c[1], w2 = (double)p.z - c[2];
  const double uu = u0 * u0 + u1 * u1 + u2 * u2, ww = w0 * w0 + w1 * w1 + w2 * w2;
  const double gmin =
      (uu - ww) - 2.0 * (hw[0] * std::fabs(u0 - w0) + hw[1] * std::fabs(u1 - w1) + hw[2] * std::fabs(u2 - w2));
  const double margin = 1e-10 * (uu + ww + 4.0 * (hw[0] * hw[0] + hw[1] * hw[1] + hw[2] * hw[2]));
  return gmin > margin;
}

// Every point of the node box [l, h] dominated by p over the box (center c, half widths hw):
// min over t in the box and q in the node of |t - q|^2 - |t - p|^2 > 0. The nearest q to t gives
// |t - q|^2 = sum over axes of d(t)^2, d the distance from t to the node's interval, so the
// expression separates by axis into f(t) = d(t)^2 - (t - p)^2: linear outside the node's interval,
// -(t - p)^2 inside it, and with the same slope 2 (p - end) on both sides of each end. f is
// therefore concave, and its minimum over the box's interval lies at one of that interval's ends.
template <class V>
FOMO_HD inline bool node_dominated(const double c[3], const double hw[3], const V& l, const V& h, const V& p) {
  const double ln[3] = {(double)l.x, (double)l.y, (double)l.z}, hn[3] = {(double)h.x, (double)h.y, (double)h.z};
  const double pp[3] = {(double)p.x, (double)p.y, (double)p.z};
  double s = 0.0, scale = 0.0;
  for (int a = 0; a < 3; ++a) {
    const double nl = ln[a] - c[a], nh = hn[a] - c[a], pa = pp[a] - c[a];
    const auto f = [&](double t) {
      const double d = std::fmax(std::fmax(nl - t, t - nh), 0.0);
      return d * d - (t - pa) * (t - pa);
    };
    s += std::fmin(f(-hw[a]), f(hw[a]));
    scale += 4.0 * (hw[a] * hw[a] + nl * nl + nh * nh + pa * pa);
  }
  return s > 1e-10 * scale;
}

}  // namespace fomo
