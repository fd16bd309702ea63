"""CPU test of the octree NN query (fomo_amd/csrc/octree.h, the code the device runs) against
brute force, with bounds-checked arrays."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_octree_query_matches_brute_force(tmp_path):
    exe = str(tmp_path / "octree_test")
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-Wall", "-Werror", "-o", exe,
                    os.path.join(ROOT, "tests", "cpp", "octree_test.cpp")], check=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "octree ok" in r.stdout
