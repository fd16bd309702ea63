// fomo_writer.h — background writer of RenderCubes (the asynchronous per-view output of
// FoMoObject::render, fomo-nearestneighbour.cpp:294-302 -> fomo-io.cpp:248-287).
#pragma once

#include <condition_variable>
#include <cstddef>
#include <deque>
#include <mutex>
#include <string>
#include <thread>
#include <utility>

#include "../../include/FoMo.h"

namespace fomo {

// Writes each submitted cube with writegoftcube(filename) on one background thread, in submission
// order, so the files are those the synchronous calls would write. At most `depth` cubes wait;
// submit() blocks while the queue is full. drain() returns once every submitted cube is written;
// the destructor drains too.
class CubeWriter {
 public:
  explicit CubeWriter(size_t depth = 2);
  ~CubeWriter();
  CubeWriter(const CubeWriter&) = delete;
  CubeWriter& operator=(const CubeWriter&) = delete;
  void submit(FoMo::RenderCube cube, std::string filename);
  void drain();

 private:
  void run();
  size_t depth_;
  std::mutex mu_;
  std::condition_variable cv_, room_, idle_;
  std::deque<std::pair<FoMo::RenderCube, std::string>> q_;
  bool busy_ = false, stop_ = false;
  std::thread th_;
};

}  // namespace fomo
