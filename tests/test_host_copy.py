"""CPU test of the parallel host copy behind large device-to-host results
(fomo_amd/csrc/host_copy.h): sizes, offsets and helper counts, bytes outside the destination
untouched."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_host_copy_pool(tmp_path):
    exe = str(tmp_path / "host_copy_test")
    subprocess.run(["g++", "-O2", "-std=c++17", "-Wall", "-Werror", "-pthread", "-I",
                    os.path.join(ROOT, "fomo_amd", "csrc"), "-o", exe,
                    os.path.join(ROOT, "tests", "cpp", "host_copy_test.cpp")], check=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "host copy ok" in r.stdout
