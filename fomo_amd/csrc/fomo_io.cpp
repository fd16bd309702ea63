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
// The gzip stream comes from zlib; its decompressed bytes equal the uncompressed file.

#include <zlib.h>

#include <cstdio>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/FoMo.h"

namespace {

const char* kVersion = "v3.3-mi355x";
const std::vector<std::string> kExt = {".dat", ".txt"};

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
  out << std::setprecision(8);
  for (int j = 0; j < ng; ++j) {
    for (int i = 0; i < dim; ++i) out << grid[i][j] << " ";
    for (int i = 0; i < nvars; ++i) out << vars[i][j] << " ";
    out << std::endl;
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

bool gzip_file(const std::string& src) {
  std::ifstream in(src, std::ios::binary);
  if (!in) return false;
  gzFile gz = gzopen((src + ".gz").c_str(), "wb");
  if (!gz) return false;
  std::vector<char> buf(1 << 20);
  while (in) {
    in.read(buf.data(), buf.size());
    std::streamsize got = in.gcount();
    if (got > 0 && gzwrite(gz, buf.data(), (unsigned)got) != got) {
      gzclose(gz);
      return false;
    }
  }
  gzclose(gz);
  return true;
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
