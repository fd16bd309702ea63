// fomo_io.cpp — GoftCube/RenderCube file output and input.
//
// Formats follow fomo-io.cpp so files stay readable by the reference's readers
// (python/readfomo.py:9-119, idl/readgoftcube*.pro):
//  * text (.txt): "FoMo-<version>", dim, ng, nvars, the dim+nvars units each followed by a tab,
//    chiantifile, abundfile, then one row per point: dim grid values and nvars values, each
//    followed by a space, printed with 8 significant digits (fomo-io.cpp:26-79);
//  * binary (.dat): "FoMo-<version>#", int32 dim, ng, nvars, for each unit a size_t length and the
//    bytes, the same for chiantifile and abundfile, then dim float32[ng] grid columns and nvars
//    float32[ng] variable columns (fomo-io.cpp:89-150);
//  * gzip (.gz) of each written file on request, optionally deleting the original (:154-170,248-287).
//
// The text rows are formatted by several threads at once, each value with std::to_chars in the
// general format at precision 8, which gives the bytes of the reference's `out << setprecision(8)
// << value` (printf's %.8g of the float widened to double); rows end in '\n' as std::endl writes
// it. The gzip file is a sequence of gzip members compressed in parallel (zlib), whose
// decompressed bytes are the uncompressed file. fomo::CubeWriter writes cubes in a background
// thread (the async writer of FoMoObject::render), in submission order, with the same bytes.

#include <zlib.h>

#include <algorithm>
#include <charconv>
#include <condition_variable>
#include <cstdio>
#include <cstring>
#include <deque>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <mutex>
#include <sstream>
#include <string>
#include <thread>
#include <vector>

#include "../../include/FoMo.h"
#include "fomo_writer.h"

namespace {

const char* kVersion = "v3.3-mi355x";
const std::vector<std::string> kExt = {".dat", ".txt"};

unsigned io_threads() {
  const unsigned hc = std::thread::hardware_concurrency();
  return std::max(1u, std::min(16u, hc ? hc : 4u));
}

// Runs fn(t) for t in [0, n) on up to io_threads() threads.
template <class Fn>
void parallel_for(size_t n, Fn fn) {
  const size_t nt = std::min<size_t>(io_threads(), n);
  if (nt <= 1) {
    for (size_t t = 0; t < n; ++t) fn(t);
    return;
  }
  std::vector<std::thread> th;
  for (size_t w = 0; w < nt; ++w)
    th.emplace_back([&, w] {
      for (size_t t = w; t < n; t += nt) fn(t);
    });
  for (auto& x : th) x.join();
}

// One value as `ostream << setprecision(8) << value` writes it, then a space.
inline char* put_value(char* p, float v) {
  p = std::to_chars(p, p + 48, (double)v, std::chars_format::general, 8).ptr;
  *p++ = ' ';
  return p;
}

void write_txt(FoMo::GoftCube& c, const std::string& filename) {
  std::ofstream out(filename, std::ios::ate);
  if (!out.is_open()) {
    std::cerr << "Unable to write to " << filename << std::endl;
    return;
  }
  const int nvars = c.readnvars(), dim = c.readdim(), ng = c.readngrid();
  out << "FoMo-" << kVersion << std::endl;
  out << dim << std::endl << ng << std::endl << nvars << std::endl;
  std::vector<std::string> units = c.readunit();
  units.resize(dim + nvars);
  for (int u = 0; u < dim + nvars; ++u) out << units[u] << "\t";
  out << std::endl;
  out << c.readchiantifile() << std::endl;
  out << c.readabundfile() << std::endl;
  FoMo::tgrid grid = c.readgrid();
  FoMo::tvars vars;
  for (int i = 0; i < nvars; ++i) vars.push_back(c.readvar(i));
  std::vector<const float*> cols;
  for (int i = 0; i < dim; ++i) cols.push_back(grid[i].data());
  for (int i = 0; i < nvars; ++i) cols.push_back(vars[i].data());
  // rows in blocks, each block formatted by one thread; a wave of blocks is written in order
  const int64_t block = 1 << 15;
  const int64_t nblocks = (ng + block - 1) / block;
  const int64_t wave = 4 * (int64_t)io_threads();
  const size_t row_max = cols.size() * 48 + 1;
  std::vector<std::string> buf((size_t)std::min<int64_t>(wave, std::max<int64_t>(nblocks, 1)));
  for (int64_t w0 = 0; w0 < nblocks; w0 += wave) {
    const int64_t nw = std::min(wave, nblocks - w0);
    parallel_for((size_t)nw, [&](size_t t) {
      const int64_t r0 = (w0 + (int64_t)t) * block, r1 = std::min<int64_t>(ng, r0 + block);
      std::string& s = buf[t];
      s.resize((size_t)(r1 - r0) * row_max);
      char* p = &s[0];
      for (int64_t j = r0; j < r1; ++j) {
        for (const float* col : cols) p = put_value(p, col[j]);
        *p++ = '\n';
      }
      s.resize((size_t)(p - s.data()));
    });
    for (int64_t t = 0; t < nw; ++t) out.write(buf[t].data(), (std::streamsize)buf[t].size());
  }
}

void write_bin(FoMo::GoftCube& c, const std::string& filename) {
  std::ofstream out(filename, std::ios::binary | std::ios::ate);
  if (!out.is_open()) {
    std::cerr << "Unable to write to " << filename << std::endl;
    return;
  }
  std::string ver = std::string("FoMo-") + kVersion + "#";
  out.write(ver.data(), ver.size());
  const int nvars = c.readnvars(), dim = c.readdim(), ng = c.readngrid();
  out.write(reinterpret_cast<const char*>(&dim), sizeof dim);
  out.write(reinterpret_cast<const char*>(&ng), sizeof ng);
  out.write(reinterpret_cast<const char*>(&nvars), sizeof nvars);
  std::vector<std::string> units = c.readunit();
  units.resize(dim + nvars);
  auto wstr = [&](const std::string& s) {
    size_t n = s.size();
    out.write(reinterpret_cast<const char*>(&n), sizeof n);
    out.write(s.data(), n);
  };
  for (int i = 0; i < dim + nvars; ++i) wstr(units[i]);
  wstr(c.readchiantifile());
  wstr(c.readabundfile());
  FoMo::tgrid grid = c.readgrid();
  for (int i = 0; i < dim; ++i) out.write(reinterpret_cast<const char*>(grid[i].data()), sizeof(float) * ng);
  for (int i = 0; i < nvars; ++i) {
    FoMo::tphysvar v = c.readvar(i);
    out.write(reinterpret_cast<const char*>(v.data()), sizeof(float) * ng);
  }
}

// One gzip member of src[0, n) into out.
bool gzip_member(const char* src, size_t n, std::string& out) {
  z_stream z{};
  if (deflateInit2(&z, Z_DEFAULT_COMPRESSION, Z_DEFLATED, 15 + 16, 8, Z_DEFAULT_STRATEGY) != Z_OK) return false;
  out.resize(deflateBound(&z, (uLong)n) + 64);
  z.next_in = reinterpret_cast<Bytef*>(const_cast<char*>(src));
  z.avail_in = (uInt)n;
  z.next_out = reinterpret_cast<Bytef*>(&out[0]);
  z.avail_out = (uInt)out.size();
  const int rc = deflate(&z, Z_FINISH);
  out.resize(out.size() - z.avail_out);
  deflateEnd(&z);
  return rc == Z_STREAM_END;
}

// src + ".gz": gzip members of 16 MiB pieces, compressed in parallel, written in order.
bool gzip_file(const std::string& src) {
  std::ifstream in(src, std::ios::binary);
  if (!in) return false;
  std::ofstream out(src + ".gz", std::ios::binary | std::ios::trunc);
  if (!out) return false;
  const size_t piece = 16u << 20;
  const size_t wave = io_threads();
  std::vector<std::string> raw(wave), packed(wave);
  bool ok = true;
  for (;;) {
    size_t nw = 0;
    for (; nw < wave; ++nw) {
      raw[nw].resize(piece);
      in.read(&raw[nw][0], (std::streamsize)piece);
      raw[nw].resize((size_t)in.gcount());
      if (raw[nw].empty()) break;
    }
    if (nw == 0) break;
    std::vector<char> good(nw, 1);
    parallel_for(nw, [&](size_t t) { good[t] = gzip_member(raw[t].data(), raw[t].size(), packed[t]) ? 1 : 0; });
    for (size_t t = 0; t < nw; ++t) {
      ok = ok && good[t];
      out.write(packed[t].data(), (std::streamsize)packed[t].size());
    }
    if (nw < wave || !in) break;
  }
  return ok && (bool)out;
}

}  // namespace

// fomo-io.cpp:248-287
void FoMo::GoftCube::writegoftcube(const std::string filename) {
  size_t pos = filename.rfind('.');
  std::string root = (pos == std::string::npos || pos == 0) ? filename : filename.substr(0, pos);
  if (writeoptions[0]) write_bin(*this, root + kExt[0]);
  if (writeoptions[1]) write_txt(*this, root + kExt[1]);
  if (writeoptions[2] && (writeoptions[0] || writeoptions[1])) {
    for (int i = 0; i < 2; ++i)
      if (writeoptions[i]) gzip_file(root + kExt[i]);
    if (writeoptions[3])
      for (int i = 0; i < 2; ++i)
        if (writeoptions[i]) std::remove((root + kExt[i]).c_str());
  } else if (writeoptions[2]) {
    std::cerr << "Zipping of output only possible if output set with goftcube.setwriteouttext or "
                 "goftcube.setwriteoutbinary."
              << std::endl
              << std::flush;
  }
}

namespace fomo {

CubeWriter::CubeWriter(size_t depth) : depth_(std::max<size_t>(1, depth)), th_([this] { run(); }) {}

CubeWriter::~CubeWriter() {
  {
    std::lock_guard<std::mutex> lk(mu_);
    stop_ = true;
  }
  cv_.notify_all();
  th_.join();
}

void CubeWriter::submit(FoMo::RenderCube cube, std::string filename) {
  std::unique_lock<std::mutex> lk(mu_);
  room_.wait(lk, [&] { return q_.size() < depth_; });
  q_.emplace_back(std::move(cube), std::move(filename));
  cv_.notify_all();
}

void CubeWriter::drain() {
  std::unique_lock<std::mutex> lk(mu_);
  idle_.wait(lk, [&] { return q_.empty() && !busy_; });
}

void CubeWriter::run() {
  for (;;) {
    std::pair<FoMo::RenderCube, std::string> job(FoMo::RenderCube(FoMo::GoftCube(3)), std::string());
    {
      std::unique_lock<std::mutex> lk(mu_);
      cv_.wait(lk, [&] { return stop_ || !q_.empty(); });
      if (q_.empty()) return;
      job = std::move(q_.front());
      q_.pop_front();
      busy_ = true;
    }
    room_.notify_all();
    job.first.writegoftcube(job.second);
    {
      std::lock_guard<std::mutex> lk(mu_);
      busy_ = false;
    }
    idle_.notify_all();
  }
}

}  // namespace fomo

// Reads the text format written above (and the older header without version and units line,
// which is what fomo-io.cpp:295-337 parses). lambda0 is re-read from the chiantifile when it exists.
void FoMo::GoftCube::readgoftcube(const std::string filename) {
  std::ifstream in(filename);
  if (!in.is_open()) {
    std::cerr << "Unable to read " << filename << std::endl;
    return;
  }
  std::string first;
  std::getline(in, first);
  std::vector<std::string> units;
  int idim = 0, ing = 0, invars = 0;
  if (first.rfind("FoMo-", 0) == 0) {
    in >> idim >> ing >> invars;
    in.ignore(1 << 20, '\n');
    std::string ul;
    std::getline(in, ul);
    std::stringstream us(ul);
    std::string u;
    while (std::getline(us, u, '\t')) units.push_back(u);
    std::getline(in, chiantifile);
    std::getline(in, abundfile);
  } else {
    std::stringstream fs(first);
    fs >> idim;
    in >> ing >> invars >> chiantifile >> abundfile;
  }
  dim = idim;
  ng = ing;
  nvars = invars;
  grid.assign(dim, tcoord(ng));
  vars.assign(nvars, tphysvar(ng));
  units.resize(dim + nvars);
  unit = units;
  {
    std::ifstream cf(chiantifile);
    std::string ion;
    double l0;
    if (cf >> ion >> l0) lambda0 = l0;
  }
  for (unsigned int j = 0; j < ng; ++j) {
    for (unsigned int i = 0; i < dim; ++i) in >> grid[i][j];
    for (unsigned int i = 0; i < nvars; ++i) in >> vars[i][j];
  }
}
