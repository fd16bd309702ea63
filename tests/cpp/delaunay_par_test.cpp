// Host test of the parallel Delaunay build (fomo_amd/csrc/delaunay.h): on jittered, uniform and
// exact-grid (cospherical) inputs, the build with several worker threads must give exactly the
// serial build's tets (vertices, order and neighbours: the canonical output), and most points
// must go in through the parallel passes. Built with ThreadSanitizer by the Python test, so a data
// race between the workers (the build shares the tet array without locks) fails it too.
#include <cstdio>
#include <cstring>
#include <random>
#include <vector>

#include "delaunay.h"

static bool check(const char* name, const std::vector<float>& x, const std::vector<float>& y,
                  const std::vector<float>& z, int threads) {
  fomo::Delaunay a, b;
  a.threads = 1;
  b.threads = threads;
  const int64_t n = (int64_t)x.size();
  if (!a.build(x.data(), y.data(), z.data(), n) || !b.build(x.data(), y.data(), z.data(), n)) {
    std::printf("%s: build failed: %s %s\n", name, a.error.c_str(), b.error.c_str());
    return false;
  }
  const bool same = a.T.size() == b.T.size() && a.nfinite == b.nfinite &&
                    std::memcmp(a.T.data(), b.T.data(), sizeof(fomo::Tet) * a.T.size()) == 0;
  std::printf("%s: %zu tets, %lld of %lld points in parallel passes, %s\n", name, b.T.size(),
              (long long)b.nparallel, (long long)b.ninserted, same ? "identical" : "DIFFERENT");
  return same && b.nparallel > b.ninserted / 2;
}

int main() {
  std::mt19937 rng(11);
  std::uniform_real_distribution<double> j(-0.3, 0.3), u(0.0, 1.0);
  const int m = 24;
  std::vector<float> x, y, z, ux, uy, uz, gx, gy, gz;
  for (int k = 0; k < m; ++k)
    for (int jj = 0; jj < m; ++jj)
      for (int i = 0; i < m; ++i) {
        x.push_back((float)(i + j(rng))), y.push_back((float)(jj + j(rng))), z.push_back((float)(k + j(rng)));
        ux.push_back((float)u(rng)), uy.push_back((float)u(rng)), uz.push_back((float)u(rng));
        gx.push_back((float)i), gy.push_back((float)jj), gz.push_back((float)k);
      }
  bool ok = check("jittered 24^3", x, y, z, 4);
  ok = check("uniform 13824", ux, uy, uz, 3) && ok;
  ok = check("exact grid 24^3", gx, gy, gz, 4) && ok;
  if (!ok) return 1;
  std::printf("delaunay parallel ok\n");
  return 0;
}
