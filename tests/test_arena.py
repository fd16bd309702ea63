"""CPU test of the render server's arena allocator (host-only C++ in fomo_amd/csrc/arena.h)."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_arena_randomised(tmp_path):
    exe = str(tmp_path / "arena_test")
    subprocess.run(["g++", "-O2", "-std=c++17", "-Wall", "-Werror", "-o", exe,
                    os.path.join(ROOT, "tests", "cpp", "arena_test.cpp")], check=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "arena ok" in r.stdout
