// Parallel host memcpy for large device-to-host results.
//
// A device-to-host copy into the caller's pageable memory runs at the speed of one host thread
// copying out of the runtime's pinned bounce buffers. Large images instead go through pinned
// staging of our own, and each staged chunk is copied to its destination by a few threads at
// once (the calling thread takes one slice).
#pragma once

#include <algorithm>
#include <condition_variable>
#include <cstddef>
#include <cstring>
#include <mutex>
#include <thread>
#include <vector>

namespace fomo {

class HostCopyPool {
 public:
  explicit HostCopyPool(int helpers) {
    for (int i = 0; i < helpers; ++i) th_.emplace_back([this] { run(); });
  }
  ~HostCopyPool() {
    {
      std::lock_guard<std::mutex> lk(mu_);
      stop_ = true;
    }
    cv_.notify_all();
    for (auto& t : th_) t.join();
  }
  HostCopyPool(const HostCopyPool&) = delete;
  HostCopyPool& operator=(const HostCopyPool&) = delete;

  // dst[0, len) = src[0, len); returns when every slice is done. One caller at a time.
  void copy(void* dst, const void* src, size_t len) {
    const size_t parts = th_.size() + 1;
    size_t slice = (len + parts - 1) / parts;
    slice = (slice + 63) & ~static_cast<size_t>(63);
    char* d = static_cast<char*>(dst);
    const char* s = static_cast<const char*>(src);
    size_t off = std::min(slice, len);  // the first slice is the caller's
    {
      std::lock_guard<std::mutex> lk(mu_);
      for (; off < len; off += slice) {
        jobs_.push_back({d + off, s + off, std::min(slice, len - off)});
        ++pending_;
      }
    }
    cv_.notify_all();
    std::memcpy(d, s, std::min(slice, len));
    std::unique_lock<std::mutex> lk(mu_);
    done_.wait(lk, [&] { return pending_ == 0; });
  }

 private:
  struct Job {
    char* dst;
    const char* src;
    size_t len;
  };
  void run() {
    for (;;) {
      Job j;
      {
        std::unique_lock<std::mutex> lk(mu_);
        cv_.wait(lk, [&] { return stop_ || !jobs_.empty(); });
        if (jobs_.empty()) return;
        j = jobs_.back();
        jobs_.pop_back();
      }
      std::memcpy(j.dst, j.src, j.len);
      std::lock_guard<std::mutex> lk(mu_);
      if (--pending_ == 0) done_.notify_all();
    }
  }
  std::mutex mu_;
  std::condition_variable cv_, done_;
  std::vector<Job> jobs_;
  size_t pending_ = 0;
  bool stop_ = false;
  std::vector<std::thread> th_;
};

}  // namespace fomo
