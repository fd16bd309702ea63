"""Host-forced failure cases of the reference's error paths, on the CPU (they end the program
before the GPU is used, as the reference's do):
- a G(T) table with fewer density rows than its nrho x nt header (fomo-CHIANTI.cpp:192-197): the
  C ABI reader returns FOMO_ERR_IO, the oracle's reader refuses it, and the FoMo.h facade prints
  the reference's message and exits with EXIT_FAILURE. A last row cut short is not an error in the
  reference: its missing values read as 0 (a failed extraction stores 0, :178-189) and the count
  still matches; the C ABI reader and the oracle both read it so;
- a missing table file (fomo-CHIANTI.cpp:144-149);
- an imaging render (one lambda pixel) with a table that is not an instrument table: its ion field
  is not the channel number (fomo-CHIANTI.cpp:278-282).
"""
import ctypes
import os
import subprocess

import pytest

import fomo_amd as F
from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "fomo_amd")


@pytest.fixture(scope="module")
def api_bin(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("bin") / "api_render")
    subprocess.check_call(["g++", "-std=c++17", "-O1", "-I" + os.path.join(ROOT, "include"),
                           os.path.join(ROOT, "tests", "cpp", "api_render.cpp"), "-o", out, "-L" + PKG, "-lFoMo",
                           "-lfomo_hip", "-Wl,-rpath," + PKG])
    return out


@pytest.fixture(scope="module")
def tables(tmp_path_factory):
    d = tmp_path_factory.mktemp("tables")
    good = str(d / "fe9.dat")
    F.gen_goft_table(good, ion="fe_9", lambda0=171.0, atweight=56.0, nt=21, nrho=9)
    head = open(good).read().splitlines()[:7]  # ion, lambda0, weight, levels, unit, version, "nrho nt"
    t = O.read_table(good)
    g = t["g"].reshape(9, 21)

    def write(path, rows, last_len=21):
        with open(path, "w") as f:
            f.write("\n".join(head) + "\n" + " ".join(f"{v:.9g}" for v in t["logt"]) + "\n")
            for r in range(rows):
                vals = g[r] if r < rows - 1 else g[r][:last_len]
                f.write(f"{t['logn'][r]:.9g} " + " ".join(f"{v:.9e}" for v in vals) + "\n")

    short = str(d / "short_rows.dat")
    write(short, 8)  # one density row fewer than the header says
    ragged = str(d / "ragged_row.dat")
    write(ragged, 9, last_len=18)  # the last row 3 values short
    return good, short, ragged


def _run(api_bin, table, tmp_path, nl):
    return subprocess.run([api_bin, table, str(tmp_path / "out."), "NearestNeighbour", "6", "8", "8", "8", str(nl),
                           "0", "0"], capture_output=True, text=True, timeout=120)


def _read(L, path, arrays=False, nt0=0, nr0=0):
    l0, aw = ctypes.c_double(), ctypes.c_double()
    nt, nr = ctypes.c_int(), ctypes.c_int()
    if not arrays:
        rc = L.fomo_table_read(path.encode(), None, 0, ctypes.byref(l0), ctypes.byref(aw), None, 0, ctypes.byref(nt),
                               ctypes.byref(nr), None, None, None)
        return rc, nt.value, nr.value
    import numpy as np
    lt, ln, g = np.zeros(nt0, np.float32), np.zeros(nr0, np.float32), np.zeros(nt0 * nr0, np.float32)
    fp = ctypes.POINTER(ctypes.c_float)
    rc = L.fomo_table_read(path.encode(), None, 0, ctypes.byref(l0), ctypes.byref(aw), None, 0, ctypes.byref(nt),
                           ctypes.byref(nr), lt.ctypes.data_as(fp), ln.ctypes.data_as(fp), g.ctypes.data_as(fp))
    return rc, lt, ln, g


def test_abi_reader_on_malformed_tables(tables):
    import numpy as np
    good, short, ragged = tables
    L = F.lib()
    assert _read(L, good) == (0, 21, 9)
    for bad in (short, good + ".missing"):
        assert _read(L, bad)[0] == -6, bad
        with pytest.raises(IOError):
            O.read_table(bad)
    # the short last row: accepted, its last three values 0, as the reference (and the oracle) read it
    rc, lt, ln, g = _read(L, ragged, True, 21, 9)
    assert rc == 0
    ref = O.read_table(ragged)
    full = O.read_table(good)
    assert np.array_equal(g, ref["g"]) and np.array_equal(lt, ref["logt"]) and np.array_equal(ln, ref["logn"])
    assert np.all(g[-3:] == 0) and np.array_equal(g[:-3], full["g"][:-3].astype(np.float32))


def test_facade_exits_on_malformed_table(api_bin, tables, tmp_path):
    r = _run(api_bin, tables[1], tmp_path, 30)
    assert r.returncode == 1 and "Error: CHIANTI table not read in properly!" in r.stderr, r.stderr


def test_facade_exits_on_missing_table(api_bin, tables, tmp_path):
    r = _run(api_bin, tables[0] + ".missing", tmp_path, 30)
    assert r.returncode == 1 and "no CHIANTI G(T) file exists" in r.stderr, r.stderr


def test_facade_exits_on_imaging_with_a_line_table(api_bin, tables, tmp_path):
    # nl = 1 selects Imaging; "fe_9" is not the channel 171 of an instrument table
    r = _run(api_bin, tables[0], tmp_path, 1)
    assert r.returncode == 1 and "GOFT table is not correct!" in r.stderr, r.stderr
