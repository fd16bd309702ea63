// First-fit offset allocator over one preallocated region (device or pinned host memory).
// Used by the render server: each admitted request takes one block sized by its footprint, and
// admission fails (the request waits) when no free block is large enough. Free blocks are kept
// sorted by offset and coalesced on release. Host-only, not thread-safe (the server locks).
#pragma once

#include <cstdint>
#include <iterator>
#include <map>

namespace fomo {

class Arena {
 public:
  static constexpr int64_t kAlign = 256;

  Arena() = default;
  explicit Arena(int64_t capacity) { reset(capacity); }

  void reset(int64_t capacity) {
    cap_ = capacity < 0 ? 0 : capacity / kAlign * kAlign;
    free_.clear();
    used_.clear();
    if (cap_ > 0) free_[0] = cap_;
    in_use_ = 0;
    peak_ = 0;
  }

  static int64_t round_up(int64_t b) { return (b + kAlign - 1) / kAlign * kAlign; }

  // Offset of a block of at least `bytes` (aligned to kAlign), or -1 when nothing fits.
  int64_t alloc(int64_t bytes) {
    if (bytes <= 0) bytes = 1;
    const int64_t need = round_up(bytes);
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second < need) continue;
      const int64_t off = it->first, len = it->second;
      free_.erase(it);
      if (len > need) free_[off + need] = len - need;
      used_[off] = need;
      in_use_ += need;
      if (in_use_ > peak_) peak_ = in_use_;
      return off;
    }
    return -1;
  }

  // Returns false for an offset that is not an allocated block.
  bool release(int64_t off) {
    auto u = used_.find(off);
    if (u == used_.end()) return false;
    int64_t start = off, len = u->second;
    used_.erase(u);
    in_use_ -= len;
    auto next = free_.lower_bound(start);
    if (next != free_.end() && next->first == start + len) {
      len += next->second;
      next = free_.erase(next);
    }
    if (next != free_.begin()) {
      auto prev = std::prev(next);
      if (prev->first + prev->second == start) {
        start = prev->first;
        len += prev->second;
        free_.erase(prev);
      }
    }
    free_[start] = len;
    return true;
  }

  int64_t capacity() const { return cap_; }
  int64_t in_use() const { return in_use_; }
  int64_t peak() const { return peak_; }
  int64_t blocks() const { return (int64_t)used_.size(); }
  int64_t largest_free() const {
    int64_t m = 0;
    for (auto& f : free_) m = f.second > m ? f.second : m;
    return m;
  }
  // Free blocks never overlap, never touch (they are coalesced) and sum to capacity - in_use.
  bool consistent() const {
    int64_t sum = 0, end = -1;
    for (auto& f : free_) {
      if (f.second <= 0 || f.first <= end || f.first % kAlign) return false;
      end = f.first + f.second;
      if (end > cap_) return false;
      sum += f.second;
    }
    for (auto& u : used_) {
      auto nx = free_.upper_bound(u.first);
      if (nx != free_.end() && nx->first < u.first + u.second) return false;
      if (nx != free_.begin()) {
        auto pv = std::prev(nx);
        if (pv->first + pv->second > u.first) return false;
      }
    }
    return sum == cap_ - in_use_;
  }

 private:
  int64_t cap_ = 0, in_use_ = 0, peak_ = 0;
  std::map<int64_t, int64_t> free_;  // offset -> length
  std::map<int64_t, int64_t> used_;  // offset -> length
};

}  // namespace fomo
