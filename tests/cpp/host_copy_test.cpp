// CPU test of fomo::HostCopyPool (fomo_amd/csrc/host_copy.h): copies of many sizes and offsets,
// with 0..7 helper threads, reused pools, and a byte pattern checked around each destination.
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <random>
#include <vector>

#include "host_copy.h"

int main() {
  std::mt19937_64 rng(7);
  std::vector<unsigned char> src(1 << 22), dst(src.size() + 256);
  for (auto& c : src) c = (unsigned char)rng();
  const size_t sizes[] = {0, 1, 7, 63, 64, 65, 127, 1000, 4095, 65536, 1000003, (size_t)1 << 22};
  int bad = 0, cases = 0;
  for (int helpers = 0; helpers <= 7; ++helpers) {
    fomo::HostCopyPool pool(helpers);
    for (size_t len : sizes) {
      for (size_t off : {(size_t)0, (size_t)3, (size_t)64}) {
        if (off + len > src.size()) continue;
        std::memset(dst.data(), 0xA5, dst.size());
        pool.copy(dst.data() + 128, src.data() + off, len);
        ++cases;
        if (std::memcmp(dst.data() + 128, src.data() + off, len) != 0) ++bad;
        for (size_t i = 0; i < 128; ++i)
          if (dst[i] != 0xA5) ++bad;
        for (size_t i = 128 + len; i < dst.size(); ++i)
          if (dst[i] != 0xA5) {
            ++bad;
            break;
          }
      }
    }
  }
  std::printf("host copy: %d cases, %d bad\n", cases, bad);
  if (bad) return 1;
  std::printf("host copy ok\n");
  return 0;
}
