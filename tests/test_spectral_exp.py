"""CPU test of the anchor exponentials of the LOS integration (fomo_amd/csrc/spectral_exp.h, the
code the device runs) against long double expl."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_anchor_exp_accuracy(tmp_path):
    exe = str(tmp_path / "spectral_exp_test")
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-Wall", "-Werror", "-o", exe,
                    os.path.join(ROOT, "tests", "cpp", "spectral_exp_test.cpp")], check=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "spectral exp ok" in r.stdout
