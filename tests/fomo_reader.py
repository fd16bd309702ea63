"""Reader for FoMo output files, restated from the reference's python/readfomo.py:9-119
(text and binary formats, with and without the version marker). Test helper."""
import gzip
import struct

import numpy as np


def read_dat(path):
    op = gzip.open if path.endswith(".gz") else open
    with op(path, "rb") as f:
        raw = f.read()
    pos = raw.find(b"#", 0, 40)
    version = raw[:pos].decode() if pos >= 0 else ""
    off = pos + 1 if "FoMo-v" in version else 0
    dim, ng, nvars = struct.unpack_from("iii", raw, off)
    off += 12
    units = []
    if "FoMo-v" in version:
        for _ in range(dim + nvars):
            (n,) = struct.unpack_from("q", raw, off)
            off += 8
            units.append(raw[off:off + n].decode())
            off += n
    (n,) = struct.unpack_from("q", raw, off)
    off += 8
    chianti = raw[off:off + n].decode()
    off += n
    (n,) = struct.unpack_from("q", raw, off)
    off += 8
    abund = raw[off:off + n].decode()
    off += n
    cols = np.frombuffer(raw, dtype=np.float32, count=(dim + nvars) * ng, offset=off).reshape(dim + nvars, ng)
    return dict(version=version, dim=dim, ng=ng, nvars=nvars, units=units, chiantifile=chianti, abundfile=abund,
                cols=cols)


def read_txt(path):
    op = gzip.open if path.endswith(".gz") else open
    with op(path, "rt") as f:
        lines = f.readlines()
    if "FoMo-v" in lines[0]:
        units = lines[4].rstrip("\n").split("\t")[:-1]
        chianti, abund, body = lines[5].rstrip("\n"), lines[6].rstrip("\n"), lines[7:]
        dim, ng, nvars = int(lines[1]), int(lines[2]), int(lines[3])
    else:
        units, chianti, abund, body = [], lines[3].strip(), lines[4].strip(), lines[5:]
        dim, ng, nvars = int(lines[0]), int(lines[1]), int(lines[2])
    data = np.array([[float(v) for v in ln.split()] for ln in body])
    return dict(dim=dim, ng=ng, nvars=nvars, units=units, chiantifile=chianti, abundfile=abund, data=data)
