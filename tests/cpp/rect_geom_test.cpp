// Host test of the rectilinear nearest-point query (fomo_amd/csrc/rect_geom.h) against brute
// force. Product grids with uniform, stretched, random, clustered (values a few ulps apart) and
// offset axes (large magnitude, small spacing), and axes of one value. Targets are random, on
// axis values, on exact midpoints (distance ties between grid points, decided by the original
// index), and far outside the grid. Search boxes are random.
//
// Masked grids (random subsets of a product grid's nodes, and spherical shells) check mrect_pick the
// same way.
//
// The query must return the brute-force answer: the nearest point by (d0^2 + d1^2) + d2^2 in double
// from float coordinates, then the lowest original index, when it lies strictly inside the box,
// and -2 otherwise. It may also return -2 (the caller then runs the exact query) when an axis's
// smallest gap is tiny against the nearest distance (g^2 <= 1e-14 D: clustered axes, very far
// targets), and only then.
#include "rect_geom.h"

#include <algorithm>
#include <cstdio>
#include <numeric>
#include <random>
#include <vector>

struct F4 {
  float x, y, z, w;
};

static std::vector<float> make_axis(std::mt19937& rng, int n, int kind) {
  std::uniform_real_distribution<double> u(-5.0, 5.0), s(0.01, 2.0);
  std::vector<float> v;
  const double x0 = u(rng), h = s(rng);
  switch (kind) {
    case 0:  // uniform
      for (int i = 0; i < n; ++i) v.push_back((float)(x0 + i * h));
      break;
    case 1:  // stretched (geometric)
      for (int i = 0; i < n; ++i) v.push_back((float)(x0 + h * (std::pow(1.3, i) - 1.0)));
      break;
    case 2:  // random
      while ((int)v.size() < n) {
        v.push_back((float)u(rng));
        std::sort(v.begin(), v.end());
        v.erase(std::unique(v.begin(), v.end()), v.end());
      }
      break;
    case 3: {  // clustered: runs of values 1-3 ulps apart
      float f = (float)x0;
      for (int i = 0; i < n; ++i) {
        v.push_back(f);
        const int steps = (i % 3 == 2) ? 1 + (int)(rng() % 3) : 0;
        if (steps) {
          for (int k = 0; k < steps; ++k) f = std::nextafter(f, INFINITY);
        } else {
          f = (float)(f + h);
        }
      }
      break;
    }
    default:  // offset: large magnitude, small spacing
      for (int i = 0; i < n; ++i) v.push_back((float)(1.0e6 + x0 + i * 0.0625));
      break;
  }
  std::sort(v.begin(), v.end());
  v.erase(std::unique(v.begin(), v.end()), v.end());
  return v;
}

int main() {
  std::mt19937 rng(11);
  long queries = 0, inside = 0, ties = 0, fallbacks = 0, fails = 0, far_q = 0;
  for (int set = 0; set < 4000; ++set) {
    std::vector<float> ax[3];
    int n[3];
    for (int a = 0; a < 3; ++a) {
      const int len = (rng() % 7 == 0) ? 1 : 2 + (int)(rng() % 11);
      ax[a] = make_axis(rng, len, (int)(rng() % 5));
      n[a] = (int)ax[a].size();
    }
    const int64_t np = (int64_t)n[0] * n[1] * n[2];
    std::vector<int> orig(np);
    std::iota(orig.begin(), orig.end(), 0);
    std::shuffle(orig.begin(), orig.end(), rng);
    std::vector<F4> bins[3];
    float o[3], inv[3];
    double gap2[3];
    int nb[3];
    for (int a = 0; a < 3; ++a) {
      nb[a] = fomo::kRectBinsPerValue * n[a];
      bins[a].resize(nb[a]);
      fomo::rect_axis_table(ax[a].data(), n[a], bins[a].data(), &o[a], &inv[a], &gap2[a]);
    }
    const auto ld = [&](int a, int k) {
      if (k < 0 || k >= n[a]) {
        std::printf("FAIL: axis read out of range (%d of %d)\n", k, n[a]);
        std::exit(1);
      }
      return ax[a][k];
    };
    const auto og = [&](int64_t g) {
      if (g < 0 || g >= np) {
        std::printf("FAIL: grid position out of range\n");
        std::exit(1);
      }
      return orig[g];
    };
    double span = 0, spacing = 1e300;
    for (int a = 0; a < 3; ++a) {
      span = std::max(span, (double)ax[a].back() - ax[a].front());
      for (int k = 0; k + 1 < n[a]; ++k) spacing = std::min(spacing, (double)ax[a][k + 1] - ax[a][k]);
    }
    if (span == 0) span = 1;
    if (spacing > 1e299) spacing = 1;
    for (int qn = 0; qn < 150; ++qn) {
      float t[3];
      const int kind = (int)(rng() % 6);
      for (int a = 0; a < 3; ++a) {
        const int k = (int)(rng() % n[a]);
        const double lo = ax[a].front(), hi = ax[a].back();
        std::uniform_real_distribution<double> u(lo - 0.3 * span - 1, hi + 0.3 * span + 1);
        switch (kind) {
          case 0: t[a] = (float)u(rng); break;
          case 1: t[a] = ax[a][k]; break;  // on an axis value
          case 2:                          // on a midpoint (exact when representable)
            t[a] = k + 1 < n[a] ? (float)(0.5 * ((double)ax[a][k] + ax[a][k + 1])) : ax[a][k];
            break;
          case 3: t[a] = (rng() & 1) ? ax[a][k] : (float)u(rng); break;
          case 4: t[a] = (float)((rng() & 1 ? 1 : -1) * (1e3 + u(rng)) * (span + 1)); break;  // far
          default: t[a] = (float)((rng() & 1 ? 1 : -1) * 1e7 * (span + 1) + u(rng)); break;   // very far
        }
      }
      if (kind >= 4) ++far_q;
      const double md = (rng() % 5 == 0) ? 1e30 : spacing * std::uniform_real_distribution<double>(0.05, 3.0)(rng);
      fomo::SearchBox box;
      box.t0 = t[0], box.t1 = t[1], box.t2 = t[2];
      box.lo0 = (float)(t[0] - md), box.lo1 = (float)(t[1] - md), box.lo2 = (float)(t[2] - md);
      box.hi0 = (float)(t[0] + md), box.hi1 = (float)(t[1] + md), box.hi2 = (float)(t[2] + md);
      // brute force over the product grid
      double best = INFINITY;
      int64_t bg = -1;
      int nbest = 0;
      for (int w = 0; w < n[2]; ++w)
        for (int v = 0; v < n[1]; ++v)
          for (int u = 0; u < n[0]; ++u) {
            const double d0 = (double)ax[0][u] - (double)t[0];
            const double d1 = (double)ax[1][v] - (double)t[1];
            const double d2 = (double)ax[2][w] - (double)t[2];
            const double d = (d0 * d0 + d1 * d1) + d2 * d2;
            const int64_t g = u + (int64_t)n[0] * (v + (int64_t)n[1] * w);
            if (d < best || (d == best && orig[g] < orig[bg])) {
              nbest = d < best ? 1 : nbest + 1;
              best = d;
              bg = g;
            } else if (d == best) {
              ++nbest;
            }
          }
      const int gu = (int)(bg % n[0]), gv = (int)((bg / n[0]) % n[1]), gw = (int)(bg / ((int64_t)n[0] * n[1]));
      const float q[3] = {ax[0][gu], ax[1][gv], ax[2][gw]};
      const bool in = box.lo0 < q[0] && q[0] < box.hi0 && box.lo1 < q[1] && q[1] < box.hi1 && box.lo2 < q[2] &&
                      q[2] < box.hi2;
      const int64_t expect = in ? bg : -2;
      F4 e[3];
      for (int a = 0; a < 3; ++a) e[a] = bins[a][fomo::rect_bin(o[a], inv[a], nb[a], t[a])];
      const int r = fomo::rect_pick(e[0], e[1], e[2], n[0], n[1], n[2], fomo::rect_dmax(gap2), box, ld, og);
      ++queries;
      inside += in;
      ties += nbest > 1;
      if (r == expect) continue;
      if (r == -2) {
        ++fallbacks;
        const bool allowed = !(best <= fomo::rect_dmax(gap2));
        if (!allowed && ++fails <= 10)
          std::printf("FAIL set %d query %d: fallback without a tiny gap (expected %lld)\n", set, qn,
                      (long long)expect);
        continue;
      }
      if (++fails <= 10)
        std::printf("FAIL set %d query %d: got %d expected %lld (ties %d)\n", set, qn, r, (long long)expect, nbest);
    }
  }
  std::printf("queries %ld, nearest inside the box %ld, distance ties %ld, far targets %ld, fallbacks %ld, "
              "failures %ld\n", queries, inside, ties, far_q, fallbacks, fails);
  // masked grids: a random subset of the nodes (a quarter to all of them, or a spherical shell)
  // holds points; mrect_pick must give the nearest present point's sorted position (inside the
  // box), -2 (outside it), or -3 only when no node at the grid's nearest distance holds a point
  long mq = 0, mhit = 0, mpass = 0, mfails = 0;
  for (int set = 0; set < 2000; ++set) {
    std::vector<float> ax[3];
    int n[3];
    for (int a = 0; a < 3; ++a) {
      ax[a] = make_axis(rng, 2 + (int)(rng() % 11), (int)(rng() % 5));
      n[a] = (int)ax[a].size();
    }
    const int64_t np = (int64_t)n[0] * n[1] * n[2];
    std::vector<int> gp(np, -1), orig_of_sp;
    const int shell = set % 4 == 0;
    const double fill = 0.25 + 0.75 * std::uniform_real_distribution<double>(0, 1)(rng);
    double c[3];
    for (int a = 0; a < 3; ++a) c[a] = 0.5 * ((double)ax[a].front() + ax[a].back());
    for (int64_t g = 0; g < np; ++g) {
      const int u = (int)(g % n[0]), v = (int)((g / n[0]) % n[1]), w = (int)(g / ((int64_t)n[0] * n[1]));
      bool keep;
      if (shell) {
        const double r2 = std::pow(ax[0][u] - c[0], 2) + std::pow(ax[1][v] - c[1], 2) + std::pow(ax[2][w] - c[2], 2);
        keep = r2 > 0.5 && r2 < 9.0;
      } else {
        keep = std::uniform_real_distribution<double>(0, 1)(rng) < fill;
      }
      if (keep) {
        gp[g] = (int)orig_of_sp.size();
        orig_of_sp.push_back(0);
      }
    }
    if (orig_of_sp.empty()) continue;
    std::vector<int> oidx(orig_of_sp.size());
    std::iota(oidx.begin(), oidx.end(), 0);
    std::shuffle(oidx.begin(), oidx.end(), rng);
    orig_of_sp = oidx;
    std::vector<F4> bins[3];
    float o[3], inv[3];
    double gap2[3];
    int nb[3];
    for (int a = 0; a < 3; ++a) {
      nb[a] = fomo::kRectBinsPerValue * n[a];
      bins[a].resize(nb[a]);
      fomo::rect_axis_table(ax[a].data(), n[a], bins[a].data(), &o[a], &inv[a], &gap2[a]);
    }
    const auto ld = [&](int a, int k) { return ax[a][k]; };
    const auto gpf = [&](int64_t g) {
      if (g < 0 || g >= np) {
        std::printf("FAIL: grid node out of range\n");
        std::exit(1);
      }
      return gp[g];
    };
    const auto og = [&](int sp) { return orig_of_sp[sp]; };
    for (int qn = 0; qn < 100; ++qn) {
      float t[3];
      for (int a = 0; a < 3; ++a) {
        const int k = (int)(rng() % n[a]);
        const double span = (double)ax[a].back() - ax[a].front() + 1;
        std::uniform_real_distribution<double> u(ax[a].front() - 0.3 * span, ax[a].back() + 0.3 * span);
        switch (qn % 3) {
          case 0: t[a] = (float)u(rng); break;
          case 1: t[a] = k + 1 < n[a] ? (float)(0.5 * ((double)ax[a][k] + ax[a][k + 1])) : ax[a][k]; break;
          default: t[a] = ax[a][k]; break;
        }
      }
      const double md = std::uniform_real_distribution<double>(0.05, 4.0)(rng);
      fomo::SearchBox box;
      box.t0 = t[0], box.t1 = t[1], box.t2 = t[2];
      box.lo0 = (float)(t[0] - md), box.lo1 = (float)(t[1] - md), box.lo2 = (float)(t[2] - md);
      box.hi0 = (float)(t[0] + md), box.hi1 = (float)(t[1] + md), box.hi2 = (float)(t[2] + md);
      double dall = INFINITY, best = INFINITY;
      int bsp = -1;
      int64_t bg = -1;
      for (int64_t g = 0; g < np; ++g) {
        const int u = (int)(g % n[0]), v = (int)((g / n[0]) % n[1]), w = (int)(g / ((int64_t)n[0] * n[1]));
        const double d0 = (double)ax[0][u] - (double)t[0], d1 = (double)ax[1][v] - (double)t[1],
                     d2 = (double)ax[2][w] - (double)t[2];
        const double d = (d0 * d0 + d1 * d1) + d2 * d2;
        dall = std::fmin(dall, d);
        if (gp[g] < 0) continue;
        if (d < best || (d == best && orig_of_sp[gp[g]] < orig_of_sp[bsp])) best = d, bsp = gp[g], bg = g;
      }
      const int u = (int)(bg % n[0]), v = (int)((bg / n[0]) % n[1]), w = (int)(bg / ((int64_t)n[0] * n[1]));
      const bool in = box.lo0 < ax[0][u] && ax[0][u] < box.hi0 && box.lo1 < ax[1][v] && ax[1][v] < box.hi1 &&
                      box.lo2 < ax[2][w] && ax[2][w] < box.hi2;
      F4 e[3];
      for (int a = 0; a < 3; ++a) e[a] = bins[a][fomo::rect_bin(o[a], inv[a], nb[a], t[a])];
      const int r = fomo::mrect_pick(e[0], e[1], e[2], n[0], n[1], n[2], fomo::rect_dmax(gap2), box, ld, gpf, og);
      ++mq;
      if (r == -3) {
        ++mpass;
        if (best == dall && dall <= fomo::rect_dmax(gap2) && ++mfails <= 10)
          std::printf("FAIL masked set %d query %d: -3 with a point on a nearest node\n", set, qn);
        continue;
      }
      ++mhit;
      const int expect = in ? bsp : -2;
      if (r != expect && ++mfails <= 10)
        std::printf("FAIL masked set %d query %d: got %d expected %d\n", set, qn, r, expect);
    }
  }
  std::printf("masked: queries %ld, answered %ld, passed on %ld, failures %ld\n", mq, mhit, mpass, mfails);
  fails += mfails;
  if (mhit * 3 < mq) {
    std::printf("FAIL: masked grids answered too few queries\n");
    return 1;
  }
  if (fails) return 1;
  if (ties < 1000) {
    std::printf("FAIL: too few tie cases exercised\n");
    return 1;
  }
  if (fallbacks * 3 > far_q) {
    std::printf("FAIL: more fallbacks than a third of the far targets\n");
    return 1;
  }
  std::printf("rect geometry ok\n");
  return 0;
}
