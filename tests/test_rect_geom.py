"""CPU test of the rectilinear nearest-point query (fomo_amd/csrc/rect_geom.h) against brute force
over product grids with uniform, stretched, random, clustered and offset axes, including exact
distance ties and far targets."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_rect_geometry(tmp_path):
    exe = str(tmp_path / "rect_geom_test")
    subprocess.run(["g++", "-O2", "-std=c++17", "-Wall", "-Werror", "-ffp-contract=off", "-I",
                    os.path.join(ROOT, "fomo_amd", "csrc"), "-o", exe, os.path.join(ROOT, "tests", "cpp",
                                                                                   "rect_geom_test.cpp")], check=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "rect geometry ok" in r.stdout
