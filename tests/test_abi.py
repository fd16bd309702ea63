"""CPU checks of the C ABI: the library loads, exports every declared entry point, rejects
calls made before initialisation, and its host-only generators are deterministic."""
import ctypes
import os
import re

import numpy as np

import fomo_amd as F

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared(header):
    txt = open(os.path.join(ROOT, "include", header)).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(fomo_\w+)\s*\(", txt)))


def test_every_declared_symbol_is_exported():
    L = ctypes.CDLL(F.LIB_PATH)
    names = _declared("fomo_hip.h")
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    # the ctypes table mirrors the header
    assert set(names) == set(F._SIGS), set(names) ^ set(F._SIGS)


def test_calls_before_init_are_rejected():
    L = F.lib()
    assert L.fomo_dataset_destroy(12345) == -5
    assert L.fomo_batch_run(7) == -5
    assert b"fomo_init" in L.fomo_last_error()


def test_version_string():
    assert F.lib().fomo_version().startswith(b"fomo-mi355x")


def test_loop_cube_generator_is_deterministic():
    a = F.gen_loop_cube(12, seed=3)
    b = F.gen_loop_cube(12, seed=3)
    c = F.gen_loop_cube(12, seed=4)
    for k in a:
        assert np.array_equal(a[k], b[k])
    assert not np.array_equal(a["rho"], c["rho"])
    assert np.array_equal(a["x"], c["x"])  # structured grid: same coordinates
    # loop axis along z: density contrast 10^0.5 inside the 1 Mm radius, lognormal noise ~5 %
    r = np.hypot(a["x"], a["y"])
    inside, outside = a["rho"][r < 0.7], a["rho"][r > 2.0]
    assert abs(np.log10(np.median(inside) / np.median(outside)) - 0.5) < 0.05
    assert 0.03 < np.std(np.log(outside / 1e9)) < 0.07


def test_jitter_moves_points_within_bound():
    a = F.gen_loop_cube(10, seed=1, jitter=0.3)
    b = F.gen_loop_cube(10, seed=1, jitter=0.0)
    h = 10.0 / 9
    d = np.abs(a["x"] - b["x"])
    assert d.max() <= 0.3 * h + 1e-6 and d.max() > 0.2 * h


def test_corona_generator_shell():
    c = F.gen_corona(24, seed=2)
    r = np.sqrt(c["x"].astype(np.float64) ** 2 + c["y"] ** 2 + c["z"] ** 2) / 695.7
    assert r.min() > 1.0 and r.max() <= 3.0 + 1e-6
    assert np.all(c["rho"] > 0)
    assert abs(c["x"].size / 24 ** 3 - 0.5) < 0.1
