// Drives the FoMo C++ API the way fomo-c/example/example.cpp does: push_back_datapoint,
// setchiantifile, setrendermethod, setresolution, render, readrendering, writegoftcube.
// usage: api_render <table> <outprefix> <method> <m> <nx> <ny> <nz> <nl> <l,...> <b,...> [options [setdata]]
// l and b are comma-separated lists (render(lvec, bvec) renders every pair); options is the
// writeoptions bitset string (default "0001": binary only); "setdata" after it fills the object with
// setdata (with units) instead of push_back_datapoint. After rendering it prints the
// rendering's shape and the GoftCube of the render as "goft <ng> <nvars> <sum var0> <sum var1>"
// and "units|<unit>|<unit>...", once from a copy of the object made before the GoftCube was read,
// once from the object.
#include <cstdlib>
#include <iomanip>
#include <iostream>
#include <sstream>
#include <vector>

#include "FoMo.h"
#include "fomo_hip.h"

static std::vector<double> parse_list(const char* s) {
  std::vector<double> v;
  std::stringstream ss(s);
  std::string tok;
  while (std::getline(ss, tok, ',')) v.push_back(std::atof(tok.c_str()));
  return v;
}

static void print_goft(FoMo::GoftCube g) {
  double s0 = 0, s1 = 0;
  if (g.readnvars() >= 2) {
    for (float v : g.readvar(0)) s0 += v;
    for (float v : g.readvar(1)) s1 += v;
  }
  std::cout << std::setprecision(17) << "goft " << g.readngrid() << " " << g.readnvars() << " " << s0 << " " << s1
            << std::endl;
  std::cout << "units";
  for (const std::string& u : g.readunit()) std::cout << "|" << u;
  std::cout << std::endl;
}

int main(int argc, char** argv) {
  if (argc < 11) {
    std::cerr << "usage: api_render table outprefix method m nx ny nz nl l,... b,... [options]" << std::endl;
    return 2;
  }
  const std::string table = argv[1], out = argv[2], method = argv[3];
  const int m = std::atoi(argv[4]), nx = std::atoi(argv[5]), ny = std::atoi(argv[6]);
  const int nz = std::atoi(argv[7]), nl = std::atoi(argv[8]);
  const std::vector<double> lvec = parse_list(argv[9]), bvec = parse_list(argv[10]);
  const std::string options = argc > 11 ? argv[11] : "0001";
  const size_t n = (size_t)m * m * m;
  std::vector<float> x(n), y(n), z(n), rho(n), T(n), vx(n), vy(n), vz(n);
  if (fomo_gen_loop_cube(m, 1, 1e6, 2e6, 0.05, 2e4, 0.0, 0.0, x.data(), y.data(), z.data(), rho.data(), T.data(),
                         vx.data(), vy.data(), vz.data()) != 0)
    return 3;
  FoMo::FoMoObject Object;
  std::vector<std::string> unitvec{"Mm", "Mm", "Mm", "cm^{-3}", "K", "m s^{-1}", "m s^{-1}", "m s^{-1}"};
  if (argc > 12 && std::string(argv[12]) == "setdata") {
    // the whole cube at once, with units (push_back_datapoint ignores them, fomo-object.cpp:139-141)
    FoMo::tgrid g{FoMo::tcoord(x.begin(), x.end()), FoMo::tcoord(y.begin(), y.end()), FoMo::tcoord(z.begin(), z.end())};
    FoMo::tvars v{FoMo::tphysvar(rho.begin(), rho.end()), FoMo::tphysvar(T.begin(), T.end()),
                  FoMo::tphysvar(vx.begin(), vx.end()), FoMo::tphysvar(vy.begin(), vy.end()),
                  FoMo::tphysvar(vz.begin(), vz.end())};
    Object.setdata(g, v, &unitvec);
  } else {
    for (size_t i = 0; i < n; ++i) {
      std::vector<double> c{x[i], y[i], z[i]};
      std::vector<double> v{rho[i], T[i], vx[i], vy[i], vz[i]};
      Object.push_back_datapoint(c, v, &unitvec);
    }
  }
  Object.setchiantifile(table);
  Object.setabundfile("/empty");
  Object.setrendermethod(method);
  Object.setobservationtype(FoMo::Spectroscopic);
  Object.setresolution(nx, ny, nz, nl, 0.2 * 2.99792458e8 / 171.0);
  Object.setoutfile(out);
  Object.setwriteoptions(std::bitset<FoMo::noptions>(options));
  Object.render(lvec, bvec);
  FoMo::RenderCube rc = Object.readrendering();
  int rx, ry, rz, rl;
  double w;
  rc.readresolution(rx, ry, rz, rl, w);
  std::cout << "rendered " << rx << "x" << ry << "x" << rz << " nl " << rl << " method " << rc.readrendermethod()
            << " ng " << rc.readngrid() << " nvars " << rc.readnvars() << std::endl;
  FoMo::FoMoObject copy = Object;
  print_goft(copy.readgoftcube());
  print_goft(Object.readgoftcube());
  return 0;
}
