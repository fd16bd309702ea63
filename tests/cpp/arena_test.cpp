// Randomised check of the server's arena allocator against a byte-map model: blocks never
// overlap, stay aligned and inside the capacity, frees coalesce (after freeing everything one
// block of the full capacity fits again) and the bookkeeping stays consistent.
#include <cstdio>
#include <cstdlib>
#include <map>
#include <random>
#include <vector>

#include "../../fomo_amd/csrc/arena.h"

int main() {
  std::mt19937_64 rng(12345);
  for (int trial = 0; trial < 40; ++trial) {
    const int64_t cap = (int64_t)(1 + rng() % 64) * 4096 + (int64_t)(rng() % 300);
    fomo::Arena a(cap);
    const int64_t c = a.capacity();
    if (c % fomo::Arena::kAlign || c > cap) return std::printf("bad capacity\n"), 1;
    std::vector<char> used((size_t)c, 0);
    std::map<int64_t, int64_t> live;  // offset -> requested bytes
    for (int step = 0; step < 4000; ++step) {
      if (live.empty() || rng() % 3) {
        const int64_t want = 1 + (int64_t)(rng() % (uint64_t)(c / 6 + 1));
        const int64_t off = a.alloc(want);
        if (off < 0) {
          // first fit must not miss a free aligned run that is large enough
          const int64_t need = fomo::Arena::round_up(want);
          for (int64_t s = 0; s + need <= c; s += fomo::Arena::kAlign) {
            bool free_run = true;
            for (int64_t k = s; k < s + need && free_run; ++k) free_run = !used[(size_t)k];
            if (free_run) return std::printf("alloc(%lld) failed with a free run at %lld\n", (long long)want, (long long)s), 1;
          }
          continue;
        }
        if (off % fomo::Arena::kAlign || off + want > c) return std::printf("bad block\n"), 1;
        for (int64_t k = off; k < off + fomo::Arena::round_up(want) && k < c; ++k) {
          if (used[(size_t)k]) return std::printf("overlap at %lld\n", (long long)k), 1;
          used[(size_t)k] = 1;
        }
        live[off] = want;
      } else {
        auto it = live.begin();
        std::advance(it, (long)(rng() % live.size()));
        if (!a.release(it->first)) return std::printf("release failed\n"), 1;
        for (int64_t k = it->first; k < it->first + fomo::Arena::round_up(it->second) && k < c; ++k) used[(size_t)k] = 0;
        live.erase(it);
      }
      if (!a.consistent()) return std::printf("inconsistent at step %d\n", step), 1;
    }
    if (a.release(1)) return std::printf("released a non-block\n"), 1;
    for (auto& kv : live) a.release(kv.first);
    if (a.in_use() != 0 || a.largest_free() != c || a.alloc(c) != 0) return std::printf("no coalescing\n"), 1;
  }
  std::printf("arena ok\n");
  return 0;
}
