// Writer checks on the CPU (no GPU needed).
// usage: writer_test <dir>
//  1. A GoftCube with awkward values (NaN, infinities, signed zero, subnormals, 8- and 9-digit
//     values) written as text, binary and gzip by writegoftcube into <dir>/sync.*; its .txt must
//     equal, byte for byte, the text the reference's writer produces (fomo-io.cpp:26-79: an
//     ostream, setprecision(8), std::endl after every row), restated here with iostreams.
//  2. Six RenderCubes written by fomo::CubeWriter (the background writer of FoMoObject::render)
//     into <dir>/async_*.txt must equal the same cubes written synchronously into <dir>/seq_*.txt.
// Exit status 0 when both hold.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <limits>
#include <sstream>
#include <string>
#include <vector>

#include "FoMo.h"
#include "../../fomo_amd/csrc/fomo_writer.h"

static std::string slurp(const std::string& p) {
  std::ifstream in(p, std::ios::binary);
  std::stringstream ss;
  ss << in.rdbuf();
  return ss.str();
}

// The reference's text writer, restated with the same stream operations.
static std::string reference_text(FoMo::GoftCube& c, const std::string& version) {
  std::ostringstream out;
  const int nvars = c.readnvars(), dim = c.readdim(), ng = c.readngrid();
  out << "FoMo-" << version << std::endl;
  out << dim << std::endl << ng << std::endl << nvars << std::endl;
  std::vector<std::string> u = c.readunit();
  for (int k = 0; k < dim + nvars; ++k) out << u.at(k) << "\t";
  out << std::endl << c.readchiantifile() << std::endl << c.readabundfile() << std::endl;
  FoMo::tgrid g = c.readgrid();
  out << std::setprecision(8);
  for (int j = 0; j < ng; ++j) {
    for (int i = 0; i < dim; ++i) out << g[i][j] << " ";
    for (int i = 0; i < nvars; ++i) out << c.readvar(i)[j] << " ";
    out << std::endl;
  }
  return out.str();
}

static FoMo::RenderCube make_cube(int seed, int ng) {
  FoMo::GoftCube base(3);
  FoMo::tgrid g(3, FoMo::tcoord(ng));
  FoMo::tvars v(2, FoMo::tphysvar(ng));
  uint32_t s = 2463534242u + seed;
  const float special[] = {0.f,
                           -0.f,
                           std::numeric_limits<float>::quiet_NaN(),
                           std::numeric_limits<float>::infinity(),
                           -std::numeric_limits<float>::infinity(),
                           std::numeric_limits<float>::denorm_min(),
                           -1.17549435e-38f,
                           3.40282347e38f,
                           171.0f,
                           99999999.f,
                           100000000.f,
                           123456789.f,
                           0.0001f,
                           1e-5f,
                           0.1f};
  for (int j = 0; j < ng; ++j)
    for (int a = 0; a < 5; ++a) {
      s ^= s << 13;
      s ^= s >> 17;
      s ^= s << 5;
      float f;
      if (j < 15) f = special[(j + a) % 15];
      else {
        uint32_t bits = s;
        std::memcpy(&f, &bits, 4);
      }
      if (a < 3) g[a][j] = f;
      else v[a - 3][j] = f;
    }
  std::vector<std::string> units{"Mm", "Mm", "\\AA{}", "erg cm^{-2} s^{-1} \\AA{}^{-1}", "x"};
  base.setdata(g, v, &units);
  base.setchiantifile("chianti_table.dat");
  base.setabundfile("/empty");
  FoMo::RenderCube rc(base);
  rc.setdata(g, v, &units);
  return rc;
}

int main(int argc, char** argv) {
  if (argc < 2) return 2;
  const std::string dir = argv[1];
  int bad = 0;
  {
    FoMo::RenderCube c = make_cube(0, 200003);
    c.setwriteoptions(std::bitset<FoMo::noptions>(std::string("0111")));  // binary, text, gzip
    c.writegoftcube(dir + "/sync.txt");
    const std::string got = slurp(dir + "/sync.txt");
    const std::string first = got.substr(0, got.find('\n'));
    const std::string want = reference_text(c, first.substr(5));
    if (got != want) {
      size_t k = 0;
      while (k < got.size() && k < want.size() && got[k] == want[k]) ++k;
      std::cerr << "text differs from the reference writer's at byte " << k << " of " << want.size() << std::endl;
      ++bad;
    }
  }
  {
    std::vector<FoMo::RenderCube> cubes;
    for (int k = 0; k < 6; ++k) {
      cubes.push_back(make_cube(k + 1, 5000 + 977 * k));
      cubes.back().setwriteoptions(std::bitset<FoMo::noptions>(std::string("0011")));  // binary + text
    }
    for (int k = 0; k < 6; ++k) cubes[k].writegoftcube(dir + "/seq_" + std::to_string(k) + ".txt");
    {
      fomo::CubeWriter w(2);
      for (int k = 0; k < 6; ++k) w.submit(cubes[k], dir + "/async_" + std::to_string(k) + ".txt");
      w.drain();
    }
    for (int k = 0; k < 6; ++k)
      for (const char* ext : {".txt", ".dat"}) {
        const std::string a = slurp(dir + "/seq_" + std::to_string(k) + ext);
        const std::string b = slurp(dir + "/async_" + std::to_string(k) + ext);
        if (a.empty() || a != b) {
          std::cerr << "async cube " << k << ext << " differs from the synchronous write" << std::endl;
          ++bad;
        }
      }
  }
  std::cout << (bad ? "FAIL" : "OK") << std::endl;
  return bad ? 1 : 0;
}
