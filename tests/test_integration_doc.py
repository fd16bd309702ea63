"""The RenderWith* binding printed in INTEGRATION.md section 2 is real code: it is taken from the
document, compiled and linked against include/FoMo.h, libFoMo.so and libfomo_hip.so here (CPU),
and on the GPU run on a GoftCube made by FoMoObject::render (emissionfromdatacube's and
thomsonfromdatacube's output, read with readgoftcube): each view's image equals fomo_render_nn /
fomo_render_thomson of the same view, bit for bit."""
import os
import re
import subprocess

import numpy as np
import pytest

from tests.fomo_reader import read_dat

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "fomo_amd")
C = 2.99792458e8

# Drives the binding: a FoMoObject renders once (writing nothing) to produce its GoftCube, which
# is then handed to RenderWithNearestNeighbour or RenderWithThomson with two views.
MAIN = r'''
#include <cstdlib>
#include <iostream>
#include "FoMo.h"
#include "fomo_hip.h"
namespace FoMo {
RenderCube RenderWithNearestNeighbour(GoftCube, const int, const int, const int, const int, const double,
                                      std::vector<double>, std::vector<double>, const std::string);
RenderCube RenderWithThomson(GoftCube, const int, const int, const int, std::vector<double>, std::vector<double>,
                             const std::string);
}
int main(int argc, char** argv) {
  if (argc < 4) return 2;  // usage: main <table> <outprefix> <nn|thomson>
  const bool th = std::string(argv[3]) == "thomson";
  const int m = th ? 28 : 20;
  std::vector<float> x, y, z, rho, T, vx, vy, vz;
  int64_t cnt = (int64_t)m * m * m;
  if (th && fomo_gen_corona(m, 3, 0.1, &cnt, nullptr, nullptr, nullptr, nullptr) != 0) return 3;
  x.resize(cnt), y.resize(cnt), z.resize(cnt), rho.resize(cnt), T.resize(cnt), vx.resize(cnt), vy.resize(cnt),
      vz.resize(cnt);
  if (th ? fomo_gen_corona(m, 3, 0.1, &cnt, x.data(), y.data(), z.data(), rho.data()) != 0
         : fomo_gen_loop_cube(m, 1, 1e6, 2e6, 0.05, 2e4, 0.0, 0.0, x.data(), y.data(), z.data(), rho.data(), T.data(),
                              vx.data(), vy.data(), vz.data()) != 0)
    return 3;
  FoMo::FoMoObject obj;
  for (int64_t i = 0; i < cnt; ++i) {
    std::vector<double> c{x[i], y[i], z[i]};
    std::vector<double> v{rho[i], T[i], vx[i], vy[i], vz[i]};
    if (th) v.resize(1);
    obj.push_back_datapoint(c, v);
  }
  obj.setchiantifile(argv[1]);
  obj.setrendermethod(th ? "Thomson" : "NearestNeighbour");
  obj.setresolution(th ? 40 : 32, th ? 36 : 24, 30, th ? 1 : 30, 0.2 * 2.99792458e8 / 171.0);
  obj.setwriteoptions(std::bitset<FoMo::noptions>(std::string("0000")));
  obj.render(0.0, 0.0);
  FoMo::GoftCube g = obj.readgoftcube();
  g.setwriteoptions(std::bitset<FoMo::noptions>(std::string("0001")));
  std::vector<double> ls{0.3, 1.1}, bs{1.0};
  FoMo::RenderCube rc = th ? FoMo::RenderWithThomson(g, 40, 36, 30, ls, bs, argv[2])
                           : FoMo::RenderWithNearestNeighbour(g, 32, 24, 30, 30, 0.2 * 2.99792458e8 / 171.0, ls, bs,
                                                             argv[2]);
  std::cout << "rendered " << rc.readngrid() << std::endl;
  return 0;
}
'''


def _binding_source():
    doc = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    sec = doc[doc.index("## 2."):]
    blocks = re.findall(r"```cpp\n(.*?)```", sec, flags=re.S)
    assert blocks and "RenderWithNearestNeighbour" in blocks[0]
    return blocks[0]


@pytest.fixture(scope="module")
def binding_exe(tmp_path_factory):
    d = tmp_path_factory.mktemp("binding")
    (d / "binding.cpp").write_text(_binding_source())
    (d / "main.cpp").write_text(MAIN)
    exe = str(d / "binding_main")
    subprocess.check_call(["g++", "-std=c++17", "-O1", "-I" + os.path.join(ROOT, "include"), str(d / "binding.cpp"),
                           str(d / "main.cpp"), "-o", exe, "-L" + PKG, "-lFoMo", "-lfomo_hip", "-Wl,-rpath," + PKG,
                           "-pthread"])
    return exe


def test_binding_compiles_and_links(binding_exe):
    assert os.path.exists(binding_exe)


@pytest.mark.gpu
@pytest.mark.parametrize("kind", ["nn", "thomson"])
def test_binding_renders_like_the_abi(engine, binding_exe, fe9_table, tmp_path, kind):
    import fomo_amd as F
    path, tab = fe9_table
    prefix = str(tmp_path / "bind.")
    r = subprocess.run([binding_exe, path, prefix, kind], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    if kind == "nn":
        d = F.gen_loop_cube(20, seed=1)
        ds = F.Dataset(d["x"], d["y"], d["z"], [d["rho"], d["T"], d["vx"], d["vy"], d["vz"]])
        T = F.Table.from_file(path)
        em = F.Emission.chianti(ds, T, tab["lambda0"], tab["atweight"])
    else:
        d = F.gen_corona(28, seed=3, sigma=0.1)
        ds = F.Dataset(d["x"], d["y"], d["z"], [d["rho"]])
        T = None
        em = F.Emission.thomson(ds)
    try:
        for l in (0.3, 1.1):
            tag = f"l{round(l / np.pi * 180):03d}b{round(1.0 / np.pi * 180):03d}"
            got = read_dat(prefix + tag + ".dat")
            if kind == "nn":
                img, _ = F.render_nn(em, l, 1.0, 32, 24, 30, 30, 0.2 * C / 171.0)
                assert got["units"][-1] == "erg cm^{-2} s^{-1} \\AA{}^{-1}"
            else:
                img, _ = F.render_thomson(em, l, 1.0, 40, 36, 30)
                assert got["units"] == ["Mm", "Mm", "pB"]
            assert np.count_nonzero(img) > 0
            assert np.array_equal(got["cols"][-1], img), f"{kind} view {tag} differs"
    finally:
        em.close()
        if T is not None:
            T.close()
        ds.close()
