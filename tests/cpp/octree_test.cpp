// Host test of the octree used for the exact NN fallback (fomo_amd/csrc/octree.h): the very query
// code the device runs, here over bounds-checked arrays. Checks the tree invariants (preorder skip
// pointers, children partition their parent's range, leaves of 1..8 points, boxes enclose their
// points), the wide records made from it (each node's children in order: boxes, leaf ranges,
// ranks), and compares cold and warm-started queries, over the preorder nodes and over the wide
// records, with brute force (nearest point strictly inside the open box, squared double distance
// of float coordinates, ties to the lowest original index) on uniform, lattice (many exact ties),
// duplicate and shell point sets.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <numeric>
#include <random>
#include <vector>

#include "../../fomo_amd/csrc/octree.h"

struct F4 {
  float x, y, z, w;
};

struct Arr {
  const F4* p;
  int64_t n;
  F4 operator[](int64_t i) const {
    if (i < 0 || i >= n) {
      std::printf("out of bounds access %lld of %lld\n", (long long)i, (long long)n);
      std::exit(3);
    }
    return p[i];
  }
};

static int fails = 0;

static void check(bool c, const char* what) {
  if (!c && fails++ < 20) std::printf("FAIL: %s\n", what);
}

static int brute(const std::vector<F4>& pts, const fomo::SearchBox& sb) {
  double best = INFINITY;
  int bp = -1, bo = 0x7fffffff;
  for (size_t p = 0; p < pts.size(); ++p) {
    const F4 q = pts[p];
    if (!sb.holds(q)) continue;
    const double d0 = (double)q.x - sb.t0, d1 = (double)q.y - sb.t1, d2 = (double)q.z - sb.t2;
    const double d = (d0 * d0 + d1 * d1) + d2 * d2;
    const int oi = fomo::f2i(q.w);
    if (d < best || (d == best && oi < bo)) {
      best = d;
      bp = (int)p;
      bo = oi;
    }
  }
  return bp;
}

static void run_set(const char* name, std::vector<F4> pts, std::mt19937_64& rng, int nq) {
  const int64_t n = (int64_t)pts.size();
  // original indices: a random permutation (pts is in "cell order" in the engine)
  std::vector<int> orig(n);
  std::iota(orig.begin(), orig.end(), 0);
  std::shuffle(orig.begin(), orig.end(), rng);
  for (int64_t p = 0; p < n; ++p) pts[p].w = fomo::i2f(orig[p]);
  float bmin[3] = {INFINITY, INFINITY, INFINITY}, bmax[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (auto& q : pts) {
    const float c[3] = {q.x, q.y, q.z};
    for (int a = 0; a < 3; ++a) bmin[a] = std::min(bmin[a], c[a]), bmax[a] = std::max(bmax[a], c[a]);
  }
  std::vector<std::pair<uint32_t, int>> kv(n);
  for (int64_t p = 0; p < n; ++p) kv[p] = {fomo::morton30(pts[p].x, pts[p].y, pts[p].z, bmin, bmax), (int)p};
  std::sort(kv.begin(), kv.end());
  std::vector<uint32_t> keys(n);
  std::vector<F4> mpts(n);
  for (int64_t m = 0; m < n; ++m) {
    keys[m] = kv[m].first;
    const F4 q = pts[kv[m].second];
    mpts[m] = {q.x, q.y, q.z, fomo::i2f(kv[m].second)};
  }
  fomo::OctBuilder ob;
  ob.build(keys.data(), n);
  const int64_t nodes = (int64_t)ob.w.size();
  std::vector<F4> lo(nodes), hi(nodes);
  for (int64_t i = 0; i < nodes; ++i) {
    lo[i] = {0, 0, 0, fomo::i2f(ob.w[i])};
    hi[i] = {0, 0, 0, fomo::i2f(ob.first[i])};
  }
  // boxes, deepest first (as the device does)
  for (int d = (int)ob.bydepth.size() - 1; d >= 0; --d)
    for (int32_t i : ob.bydepth[d]) {
      float l[3] = {INFINITY, INFINITY, INFINITY}, h[3] = {-INFINITY, -INFINITY, -INFINITY};
      const int w = ob.w[i];
      auto grow = [&](float x0, float y0, float z0, float x1, float y1, float z1) {
        l[0] = std::min(l[0], x0), l[1] = std::min(l[1], y0), l[2] = std::min(l[2], z0);
        h[0] = std::max(h[0], x1), h[1] = std::max(h[1], y1), h[2] = std::max(h[2], z1);
      };
      if (w < 0) {
        for (int q = ob.first[i]; q < ob.first[i] - w; ++q) grow(mpts[q].x, mpts[q].y, mpts[q].z, mpts[q].x, mpts[q].y, mpts[q].z);
      } else {
        for (int c = i + 1; c < w;) {
          grow(lo[c].x, lo[c].y, lo[c].z, hi[c].x, hi[c].y, hi[c].z);
          const int cw = ob.w[c];
          c = cw < 0 ? c + 1 : cw;
        }
      }
      lo[i] = {l[0], l[1], l[2], lo[i].w};
      hi[i] = {h[0], h[1], h[2], hi[i].w};
    }
  // invariants
  int64_t leafpts = 0;
  for (int64_t i = 0; i < nodes; ++i) {
    const int w = ob.w[i];
    if (w < 0) {
      check(-w >= 1 && -w <= 8, "leaf size in 1..8");
      leafpts += -w;
      for (int q = ob.first[i]; q < ob.first[i] - w; ++q)
        check(lo[i].x <= mpts[q].x && mpts[q].x <= hi[i].x && lo[i].y <= mpts[q].y && mpts[q].y <= hi[i].y &&
                  lo[i].z <= mpts[q].z && mpts[q].z <= hi[i].z, "leaf box encloses its points");
    } else {
      check(w > i + 1 && w <= nodes, "skip pointer past the node");
      int64_t expect = ob.first[i], nch = 0;
      for (int c = i + 1; c < w;) {
        check(ob.first[c] == expect, "children partition the range in order");
        const int cw = ob.w[c];
        // range end of child c: the next child's first, or for the last child the parent's end
        c = cw < 0 ? c + 1 : cw;
        expect = c < w ? ob.first[c] : expect;
        ++nch;
      }
      check(nch >= 2 && nch <= 8, "2..8 children");
    }
  }
  check(leafpts == n, "leaves cover every point once");
  // wide records (the device's k_nn_far walk): made by the same functions the device build uses
  std::vector<int32_t> kid, kref;
  const int64_t nint = fomo::wide_children(ob, kid, kref);
  check(nodes == 1 ? nint == 0 : nint > 0, "wide records for every tree with an internal root");
  std::vector<F4> wide((size_t)nint * 16);
  for (int64_t t = 0; t < nint * 8; ++t) {
    const int c = kid[t];
    const F4 z{0, 0, 0, 0};
    fomo::wide_slot(c, kref[t], c >= 0 ? lo[c] : z, c >= 0 ? hi[c] : z, wide[2 * t], wide[2 * t + 1]);
  }
  {
    // every record lists its node's children in order; leaf slots hold their point ranges,
    // internal slots the child's rank, whose own record is that child's
    int64_t r = 0, wleaf = 0;
    for (int64_t i = 0; i < nodes; ++i) {
      if (ob.w[i] < 0) continue;
      int s = 0;
      for (int c = (int)i + 1; c < ob.w[i]; c = ob.w[c] < 0 ? c + 1 : ob.w[c], ++s) {
        const F4 a = wide[r * 16 + 2 * s], b = wide[r * 16 + 2 * s + 1];
        check(a.x == lo[c].x && a.y == lo[c].y && a.z == lo[c].z && b.x == hi[c].x && b.y == hi[c].y &&
                  b.z == hi[c].z, "wide slot box is the child's");
        if (ob.w[c] < 0) {
          check(fomo::f2i(a.w) == ob.first[c] && fomo::f2i(b.w) == -ob.w[c], "wide leaf slot: first, count");
          wleaf += -ob.w[c];
        } else {
          check(fomo::f2i(b.w) == 0 && fomo::f2i(a.w) == kref[r * 8 + s] && kid[r * 8 + s] == c,
                "wide internal slot: the child's rank");
        }
      }
      for (; s < 8; ++s) {
        const F4 a = wide[r * 16 + 2 * s], b = wide[r * 16 + 2 * s + 1];
        check(kid[r * 8 + s] == -1 && a.x == INFINITY && b.x == -INFINITY && fomo::f2i(b.w) == 0,
              "wide empty slot");
      }
      ++r;
    }
    check(r == nint && (nint == 0 || wleaf == n), "wide leaf slots cover every point once");
  }
  const Arr A{pts.data(), n}, M{mpts.data(), n}, L{lo.data(), nodes}, H{hi.data(), nodes};
  const Arr W{wide.data(), (int64_t)wide.size()};
  const float ext = std::max({bmax[0] - bmin[0], bmax[1] - bmin[1], bmax[2] - bmin[2], 1e-3f});
  std::uniform_real_distribution<float> U(0.f, 1.f);
  int bad = 0;
  for (int t = 0; t < nq; ++t) {
    fomo::SearchBox sb;
    // targets: inside, just outside and far outside the data; some exactly on data points
    float c[3];
    if (t % 7 == 0) {
      const F4 q = pts[rng() % n];
      c[0] = q.x, c[1] = q.y, c[2] = q.z;
    } else {
      const float s = (t % 3 == 0) ? 3.f : 1.4f;
      for (int a = 0; a < 3; ++a) c[a] = bmin[a] + (bmax[a] - bmin[a]) * 0.5f + ext * s * (U(rng) - 0.5f);
    }
    const double md = ext * (t % 5 == 0 ? 5.0 : (t % 5 == 1 ? 0.02 : 0.3 * U(rng)));
    sb.t0 = c[0], sb.t1 = c[1], sb.t2 = c[2];
    sb.lo0 = (float)(c[0] - md), sb.lo1 = (float)(c[1] - md), sb.lo2 = (float)(c[2] - md);
    sb.hi0 = (float)(c[0] + md), sb.hi1 = (float)(c[1] + md), sb.hi2 = (float)(c[2] + md);
    const int want = brute(pts, sb);
    const int cold = fomo::oct_query(A, M, L, H, nodes, sb, INFINITY, -1, 0x7fffffff);
    // warm: any point strictly inside the box as the starting bound
    int seed = -1;
    for (int tries = 0; tries < 50 && seed < 0; ++tries) {
      const int p = (int)(rng() % n);
      if (sb.holds(pts[p])) seed = p;
    }
    int warm = cold;
    if (seed >= 0) {
      const F4 q = pts[seed];
      const double d0 = (double)q.x - sb.t0, d1 = (double)q.y - sb.t1, d2 = (double)q.z - sb.t2;
      warm = fomo::oct_query(A, M, L, H, nodes, sb, (d0 * d0 + d1 * d1) + d2 * d2, seed, fomo::f2i(q.w));
    }
    // the wide records, cold and from the same warm start
    int wcold = cold, wwarm = warm;
    if (nint > 0) {
      wcold = fomo::wide_query(A, M, W, sb, INFINITY, -1, 0x7fffffff);
      if (seed >= 0) {
        const F4 q = pts[seed];
        const double d0 = (double)q.x - sb.t0, d1 = (double)q.y - sb.t1, d2 = (double)q.z - sb.t2;
        wwarm = fomo::wide_query(A, M, W, sb, (d0 * d0 + d1 * d1) + d2 * d2, seed, fomo::f2i(q.w));
      } else {
        wwarm = wcold;
      }
    }
    if (cold != want || warm != want || wcold != want || wwarm != want) {
      if (bad++ < 5)
        std::printf("%s: query %d cold %d warm %d wide %d %d brute %d\n", name, t, cold, warm, wcold, wwarm, want);
    }
  }
  check(bad == 0, "queries equal brute force");
  std::printf("%s: n %lld nodes %lld depth %zu queries %d mismatches %d\n", name, (long long)n, (long long)nodes,
              ob.bydepth.size(), nq, bad);
}

int main() {
  std::mt19937_64 rng(2024);
  std::uniform_real_distribution<float> U(-1.f, 1.f);
  {
    std::vector<F4> p(20000);
    for (auto& q : p) q = {U(rng) * 5.f, U(rng) * 2.f, U(rng) * 0.5f + 3.f, 0};
    run_set("uniform", p, rng, 3000);
  }
  {
    std::vector<F4> p;  // lattice: exact distance ties everywhere
    for (int i = 0; i < 24; ++i)
      for (int j = 0; j < 20; ++j)
        for (int k = 0; k < 18; ++k) p.push_back({i * 0.5f, j * 0.5f, k * 0.5f, 0});
    run_set("lattice", p, rng, 3000);
  }
  {
    std::vector<F4> p;  // duplicates: 300 copies of a few points plus noise
    for (int i = 0; i < 300; ++i)
      for (int c = 0; c < 4; ++c) p.push_back({(float)c, 1.f, 2.f, 0});
    for (int i = 0; i < 500; ++i) p.push_back({U(rng) * 3.f, U(rng) * 3.f, U(rng) * 3.f, 0});
    run_set("duplicates", p, rng, 2000);
  }
  {
    std::vector<F4> p;  // spherical shell on a grid (hole and outside queries)
    const int m = 40;
    for (int i = 0; i < m; ++i)
      for (int j = 0; j < m; ++j)
        for (int k = 0; k < m; ++k) {
          const float x = -3.f + 6.f * i / (m - 1), y = -3.f + 6.f * j / (m - 1), z = -3.f + 6.f * k / (m - 1);
          const float r = std::sqrt(x * x + y * y + z * z);
          if (r > 1.f && r <= 3.f) p.push_back({x * 695.7f, y * 695.7f, z * 695.7f, 0});
        }
    run_set("shell", p, rng, 3000);
  }
  {
    std::vector<F4> p(1, F4{1.f, 2.f, 3.f, 0});
    run_set("single", p, rng, 200);
    std::vector<F4> q(9, F4{0.f, 0.f, 0.f, 0});
    run_set("nine identical", q, rng, 200);
  }
  if (fails) {
    std::printf("%d failures\n", fails);
    return 1;
  }
  std::printf("octree ok\n");
  return 0;
}
