// Host test of the lookup grid's box geometry (fomo_amd/csrc/lookup_geom.h) against brute force:
// box distances against sampled points of the box, the point domination test against sampled
// targets, and the node domination test against sampled targets with the analytically nearest
// point of the node. A domination claim that a sample contradicts would drop a true candidate
// (an NN choice error), so soundness is checked on every case; completeness (no clear case
// missed) only where the sampled minimum is well above zero.
#include "lookup_geom.h"

#include <algorithm>
#include <cstdio>
#include <random>

struct P {
  float x, y, z;
};

int main() {
  std::mt19937 rng(7);
  std::uniform_real_distribution<double> u(-2.0, 2.0), w(0.005, 1.0);
  const int K = 6;  // targets per box edge: (K+1)^3 samples, corners included
  long fails = 0, claims_node = 0, claims_point = 0, missed = 0;
  for (int it = 0; it < 60000; ++it) {
    double c[3], hw[3], lo[3], hi[3];
    for (int a = 0; a < 3; ++a) {
      c[a] = u(rng);
      hw[a] = 0.5 * w(rng);
      lo[a] = c[a] - hw[a];
      hi[a] = c[a] + hw[a];
    }
    P nl, nh, p, q;
    float a0[3], a1[3];
    for (int a = 0; a < 3; ++a) {
      const double m = u(rng), e = it % 4 == 0 ? 0.0 : w(rng);  // every 4th node is a single point
      a0[a] = (float)(m - e);
      a1[a] = (float)(m + e);
    }
    nl = {a0[0], a0[1], a0[2]};
    nh = {a1[0], a1[1], a1[2]};
    p = {(float)u(rng), (float)u(rng), (float)u(rng)};
    q = {(float)u(rng), (float)u(rng), (float)u(rng)};
    double mn_node = 1e300, mn_point = 1e300, bmin = 1e300, bmax = 0.0;
    for (int i = 0; i <= K; ++i)
      for (int j = 0; j <= K; ++j)
        for (int k = 0; k <= K; ++k) {
          const double t[3] = {lo[0] + (hi[0] - lo[0]) * i / K, lo[1] + (hi[1] - lo[1]) * j / K,
                               lo[2] + (hi[2] - lo[2]) * k / K};
          const double pc[3] = {p.x, p.y, p.z}, qc[3] = {q.x, q.y, q.z}, l3[3] = {nl.x, nl.y, nl.z},
                       h3[3] = {nh.x, nh.y, nh.z};
          double dn = 0, dp = 0, dq = 0;
          for (int a = 0; a < 3; ++a) {
            const double nearest = std::min(std::max(t[a], l3[a]), h3[a]);
            dn += (t[a] - nearest) * (t[a] - nearest);
            dp += (t[a] - pc[a]) * (t[a] - pc[a]);
            dq += (t[a] - qc[a]) * (t[a] - qc[a]);
          }
          mn_node = std::min(mn_node, dn - dp);
          mn_point = std::min(mn_point, dq - dp);
          bmin = std::min(bmin, dq);
          bmax = std::max(bmax, dq);
        }
    // box distances: the sampled range lies inside [mindist, maxdist], and the corners reach maxdist
    const double md = fomo::box_mindist2(lo, hi, q), Md = fomo::box_maxdist2(lo, hi, q);
    if (bmin < md * (1 - 1e-12) - 1e-15 || bmax > Md * (1 + 1e-12) + 1e-15 || bmax < Md * (1 - 1e-9)) {
      std::printf("box distance case %d: sampled [%g, %g], computed [%g, %g]\n", it, bmin, bmax, md, Md);
      ++fails;
    }
    if (fomo::node_dominated(c, hw, nl, nh, p)) {
      ++claims_node;
      if (mn_node <= 0) {
        std::printf("node domination unsound, case %d: sampled min %g\n", it, mn_node);
        ++fails;
      }
    } else if (mn_node > 1e-3) {
      ++missed;
    }
    if (fomo::dominated(c, hw, q, p)) {
      ++claims_point;
      if (mn_point <= 0) {
        std::printf("point domination unsound, case %d: sampled min %g\n", it, mn_point);
        ++fails;
      }
    } else if (mn_point > 1e-3) {
      ++missed;
    }
    // a node that is a single point tests like that point
    if (it % 4 == 0 && fomo::node_dominated(c, hw, nl, nh, p) != fomo::dominated(c, hw, nl, p) &&
        std::fabs(mn_node) > 1e-6) {
      std::printf("point node case %d differs from the point test\n", it);
      ++fails;
    }
  }
  // a point never dominates itself or a node holding it
  const double c[3] = {0, 0, 0}, hw[3] = {0.5, 0.5, 0.5};
  const P s = {1.5f, -0.25f, 2.0f}, sl = {1.0f, -1.0f, 1.5f}, sh = {2.0f, 0.5f, 2.5f};
  if (fomo::dominated(c, hw, s, s) || fomo::node_dominated(c, hw, sl, sh, s)) ++fails;
  std::printf("node claims %ld, point claims %ld, clear cases missed %ld, failures %ld\n", claims_node,
              claims_point, missed, fails);
  if (fails || missed || claims_node < 5000 || claims_point < 5000) return 1;
  std::printf("lookup geometry ok\n");
  return 0;
}
