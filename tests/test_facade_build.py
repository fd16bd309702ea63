"""The FoMo.h C++ API builds and links: an API-usage program, and (where the reference tree is
mounted) the reference's own example program compiled unchanged against this header."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "fomo_amd")
REF_EXAMPLE = "/root/reference/fomo-c/example/example.cpp"


def _link(src, out):
    cmd = ["g++", "-std=c++17", "-O1", "-I" + os.path.join(ROOT, "include"), src, "-o", out, "-L" + PKG, "-lFoMo",
           "-lfomo_hip", "-Wl,-rpath," + PKG]
    subprocess.check_call(cmd)


def test_api_program_links(tmp_path):
    assert os.path.exists(os.path.join(PKG, "libFoMo.so")), "build libFoMo.so first (python -m fomo_amd.build)"
    _link(os.path.join(ROOT, "tests", "cpp", "api_render.cpp"), str(tmp_path / "api_render"))
    assert os.path.exists(tmp_path / "api_render")


@pytest.mark.skipif(not os.path.exists(REF_EXAMPLE), reason="reference example not mounted")
def test_reference_example_compiles_unchanged(tmp_path):
    _link(REF_EXAMPLE, str(tmp_path / "example"))
    out = subprocess.run([str(tmp_path / "example")], capture_output=True, text=True)
    # without arguments the example prints its usage and exits with EXIT_FAILURE before touching the GPU
    assert out.returncode == 1 and "one file name as argument" in out.stdout
