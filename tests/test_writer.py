"""The file writer (rows A15 and the asynchronous writer of the next-steps list) on the CPU.

tests/cpp/writer_test.cpp writes cubes with NaN, infinities, signed zeros, subnormals and 8- and
9-digit values. It checks the text file byte for byte against the reference's own text writer
(fomo-io.cpp:26-79, an ostream at setprecision(8) with std::endl after every row, restated there)
and the background writer's files against synchronous writes. Here the same files are read back
with the reader restated from python/readfomo.py: the text and binary formats agree, and each
.gz decompresses to the bytes of the file it was made from (fomo-io.cpp:154-170).
"""
import gzip
import os
import subprocess

import numpy as np

from tests.fomo_reader import read_dat, read_txt

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "fomo_amd")


def test_text_binary_gzip_and_async_writer(tmp_path):
    exe = str(tmp_path / "writer_test")
    subprocess.check_call(["g++", "-std=c++17", "-O1", "-I" + os.path.join(ROOT, "include"),
                           os.path.join(ROOT, "tests", "cpp", "writer_test.cpp"), "-o", exe, "-L" + PKG, "-lFoMo",
                           "-lfomo_hip", "-Wl,-rpath," + PKG, "-pthread"])
    r = subprocess.run([exe, str(tmp_path)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    txt = read_txt(str(tmp_path / "sync.txt"))
    dat = read_dat(str(tmp_path / "sync.dat"))
    for k in ("dim", "ng", "nvars", "chiantifile", "abundfile"):
        assert txt[k] == dat[k], k
    assert txt["units"] == dat["units"] == ["Mm", "Mm", "\\AA{}", "erg cm^{-2} s^{-1} \\AA{}^{-1}", "x"]
    assert txt["dim"] == 3 and txt["nvars"] == 2 and txt["ng"] == 200003
    a = txt["data"].T.astype(np.float32)
    b = dat["cols"]
    assert a.shape == b.shape
    assert np.array_equal(np.isnan(a), np.isnan(b))
    fin = np.isfinite(b)
    assert np.array_equal(a[~fin & ~np.isnan(b)], b[~fin & ~np.isnan(b)])
    # 8 significant digits: within one float32 rounding of the binary value
    np.testing.assert_allclose(a[fin], b[fin], rtol=1.2e-7, atol=0)
    for ext in (".txt", ".dat"):
        raw = open(tmp_path / ("sync" + ext), "rb").read()
        with gzip.open(tmp_path / ("sync" + ext + ".gz"), "rb") as f:
            assert f.read() == raw, ext
    # the gzip of the text file is several members (compressed in parallel); a stream reader
    # (gzip.open) reads them all, as gunzip does
    assert os.path.getsize(tmp_path / "sync.txt.gz") < os.path.getsize(tmp_path / "sync.txt")
