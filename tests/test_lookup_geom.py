"""CPU test of the candidate lookup grid's box geometry (fomo_amd/csrc/lookup_geom.h): box
distances and the point and node domination tests that prune candidate lists, against brute
force over sampled targets."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_lookup_geometry(tmp_path):
    exe = str(tmp_path / "lookup_geom_test")
    subprocess.run(["g++", "-O2", "-std=c++17", "-Wall", "-Werror", "-I", os.path.join(ROOT, "fomo_amd", "csrc"),
                    "-o", exe, os.path.join(ROOT, "tests", "cpp", "lookup_geom_test.cpp")], check=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "lookup geometry ok" in r.stdout
