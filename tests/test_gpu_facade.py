"""GPU: the FoMo.h C++ API end to end (push_back_datapoint ... render ... writegoftcube), checked
against the C ABI render and the FP64 oracle, through the binary file it writes."""
import gzip
import os
import subprocess

import numpy as np
import pytest

import fomo_amd as F
from oracle import oracle as O
from tests.fomo_reader import read_dat, read_txt

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "fomo_amd")
C = 2.99792458e8


@pytest.fixture(scope="module")
def api_bin(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("bin") / "api_render")
    subprocess.check_call(["g++", "-std=c++17", "-O1", "-I" + os.path.join(ROOT, "include"),
                           os.path.join(ROOT, "tests", "cpp", "api_render.cpp"), "-o", out, "-L" + PKG, "-lFoMo",
                           "-lfomo_hip", "-Wl,-rpath," + PKG])
    return out


def test_facade_nn_render_file(api_bin, fe9_table, tmp_path):
    path, tab = fe9_table
    m, nx, ny, nz, nl, l, b = 20, 32, 24, 30, 30, 0.3, 1.0
    prefix = str(tmp_path / "run.")
    r = subprocess.run([api_bin, path, prefix, "NearestNeighbour", str(m), str(nx), str(ny), str(nz), str(nl), str(l),
                        str(b)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    f = prefix + "l017b057.dat"
    assert os.path.exists(f), os.listdir(tmp_path)
    got = read_dat(f)
    assert got["dim"] == 3 and got["nvars"] == 1 and got["ng"] == nx * ny * nl
    assert got["units"][:3] == ["Mm", "Mm", "\\AA{}"] and got["chiantifile"] == path
    d = F.gen_loop_cube(m, seed=1)
    pk, fw = O.emission(1, d["rho"], d["T"], tab)
    width = 0.2 * C / 171.0
    ref, _, meta, _ = O.render_nn(1, d["x"], d["y"], d["z"], pk, fw, d["vx"], d["vy"], d["vz"], l, b, nx, ny, nz, nl,
                                  width, 171.0)
    assert O.rel_l2(got["cols"][3], ref) <= 1e-5
    # grid columns: x, y, lambda per element (fomo-nearestneighbour.cpp:209-211)
    j = np.arange(nx)
    x = (j / (nx - 1) * (meta[1] - meta[0]) + meta[0]).astype(np.float32)
    assert np.array_equal(got["cols"][0].reshape(ny, nx, nl)[5, :, 3], x)
    lam = (np.arange(nl) / (nl - 1) * (width * 171.0 / C) - width * 171.0 / C / 2 + 171.0).astype(np.float32)
    assert np.array_equal(got["cols"][2].reshape(ny, nx, nl)[2, 7, :], lam)


def test_facade_unknown_method_exits(api_bin, fe9_table, tmp_path):
    path, _ = fe9_table
    r = subprocess.run([api_bin, path, str(tmp_path / "x."), "NoSuchMethod", "8", "8", "8", "8", "4", "0", "0"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 1 and "unknown rendering method" in r.stderr


def test_facade_projection_render_file(api_bin, fe9_table, tmp_path):
    path, tab = fe9_table
    m, nx, ny, nz, nl, l, b = 20, 28, 22, 30, 24, 0.7, 1.2
    prefix = str(tmp_path / "proj.")
    r = subprocess.run([api_bin, path, prefix, "Projection", str(m), str(nx), str(ny), str(nz), str(nl), str(l),
                        str(b)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    f = prefix + "l040b069.dat"
    assert os.path.exists(f), os.listdir(tmp_path)
    got = read_dat(f)
    assert got["dim"] == 3 and got["nvars"] == 1 and got["ng"] == nx * ny * nl
    d = F.gen_loop_cube(m, seed=1)
    width = 0.2 * C / 171.0
    pk, fw = O.emission(0, d["rho"], d["T"], tab)
    ref, _, meta = O.render_projection(0, d["x"], d["y"], d["z"], pk, fw, d["vx"], d["vy"], d["vz"], l, b, nx, ny,
                                       nz, nl, width, 171.0)
    assert np.count_nonzero(ref) > 0
    assert O.rel_l2(got["cols"][3], ref) <= 1e-6
    lam = (np.arange(nl) / (nl - 1) * (width * 171.0 / C) - width * 171.0 / C / 2 + 171.0).astype(np.float32)
    assert np.array_equal(got["cols"][2].reshape(ny, nx, nl)[3, 4, :], lam)


def test_facade_multiview_batch_text_gzip_and_goftcube(engine, api_bin, fe9_table, tmp_path):
    """render(lvec, bvec) with several NearestNeighbour views runs them as one batch and writes each
    view's files from a background thread: every view's image equals the single-view C ABI render
    of that view bit for bit; the text (the reference's default format), binary and gzip files
    agree; the GoftCube read after the render (also from a copy of the object made before it was
    read) holds the emission computed on the device, with the reference's units."""
    path, tab = fe9_table
    m, nx, ny, nz, nl = 20, 32, 24, 30, 30
    ls, bs = [0.3, 1.1], [1.0, 0.4]
    prefix = str(tmp_path / "multi.")
    r = subprocess.run([api_bin, path, prefix, "NearestNeighbour", str(m), str(nx), str(ny), str(nz), str(nl),
                        ",".join(map(str, ls)), ",".join(map(str, bs)), "0111", "setdata"], capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stderr
    d = F.gen_loop_cube(m, seed=1)
    ds = F.Dataset(d["x"], d["y"], d["z"], [d["rho"], d["T"], d["vx"], d["vy"], d["vz"]])
    T = F.Table.from_file(path)
    em = F.Emission.chianti(ds, T, tab["lambda0"], tab["atweight"])
    width = 0.2 * C / 171.0
    try:
        for l in ls:
            for b in bs:
                tag = f"l{round(l / np.pi * 180):03d}b{round(b / np.pi * 180):03d}"
                dat = read_dat(prefix + tag + ".dat")
                img, _ = F.render_nn(em, l, b, nx, ny, nz, nl, width)
                assert np.array_equal(dat["cols"][3], img), f"view {tag}: batched facade image differs"
                txt = read_txt(prefix + tag + ".txt")
                assert (txt["dim"], txt["ng"], txt["nvars"]) == (dat["dim"], dat["ng"], dat["nvars"])
                assert txt["units"] == dat["units"] and txt["chiantifile"] == dat["chiantifile"]
                np.testing.assert_allclose(txt["data"].T.astype(np.float32), dat["cols"], rtol=1.2e-7, atol=1e-45)
                for ext in (".dat", ".txt"):
                    with gzip.open(prefix + tag + ext + ".gz", "rb") as f:
                        assert f.read() == open(prefix + tag + ext, "rb").read()
        gl = [ln.split() for ln in r.stdout.splitlines() if ln.startswith("goft ")]
        assert len(gl) == 2 and gl[0] == gl[1], r.stdout
        pk, fw = em.read()
        assert int(gl[0][1]) == m ** 3 and int(gl[0][2]) == 5
        assert abs(float(gl[0][3]) / pk.astype(np.float64).sum() - 1) < 1e-12
        assert abs(float(gl[0][4]) / fw.astype(np.float64).sum() - 1) < 1e-12
        # GoftCube units as the reference builds them (fomo-CHIANTI.cpp:301-315): the grid units,
        # the table's emissivity unit, the FWHM unit, then dataunits.at(i) for the velocity
        # columns i = 2..4, which in 3D are the z grid unit and the units of n and T
        ul = [ln.split("|")[1:] for ln in r.stdout.splitlines() if ln.startswith("units")]
        assert len(ul) == 2 and ul[0] == ul[1], r.stdout
        assert ul[0] == ["Mm", "Mm", "Mm", tab["unit"], "\\AA{}", "Mm", "cm^{-3}", "K"], ul[0]
    finally:
        em.close()
        T.close()
        ds.close()
