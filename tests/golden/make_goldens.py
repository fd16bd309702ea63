"""Writes the golden fixtures of this folder from the oracle. Run from the repository root:

    python tests/golden/make_goldens.py

Each fixture (.npz) holds SHA-256 digests of its inputs (the generated points and fields, the
generated G(T, n) table file), the view parameters, and the oracle's outputs: the FP64 image,
the nearest-neighbour (or tetra vertex) choice of every LOS sample and the view meta. The inputs
come from the library's own seeded generators (fomo_gen_loop_cube, fomo_gen_corona,
fomo_gen_goft_table), so they are rebuilt, not stored. tests/test_golden.py checks the oracle
against these files on the CPU; tests/test_gpu_golden.py checks the HIP engine against them.

Cases (small enough to render in seconds):
  loop16_nn     16^3 loop cube (seed 1), NearestNeighbour, three views, 20x20x24 samples, 12 lambda
  c0_nn         the 64^3 structured loop cube (seed 1), one oblique view, 64x64x64 samples, 8 lambda,
                rows 24..40 (image band and NN choices of the band)
  c0_imaging    the same cube, imaging (one lambda pixel), the whole 64x64 image and the NN choices
                of rows 24..40
  corona40_pb   40^3-grid corona (seed 3), Thomson pB, two views, 48x40x36 samples
  tet12_linear  jittered 12^3 loop cube (seed 9), tetra path: nearest-vertex choice and the
                barycentric ("Linear") image, 32x32x24 samples, 10 lambda
"""
import hashlib
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

C = 2.99792458e8


def digest(*arrays):
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def file_digest(path):
    with open(path, "rb") as f:
        return hashlib.sha256(f.read()).hexdigest()


def table(F, tmpdir):
    """The synthetic Fe IX 171 table of the tests (gen_goft_table defaults)."""
    p = os.path.join(tmpdir, "fe9_171.dat")
    F.gen_goft_table(p, ion="fe_9", lambda0=171.0, atweight=56.0, nt=101, nrho=41, peak_logt=5.9, sigma=0.15)
    return p


def loop_inputs(F, m, seed, **kw):
    d = F.gen_loop_cube(m, seed=seed, **kw)
    return d, digest(*(d[k] for k in ("x", "y", "z", "rho", "T", "vx", "vy", "vz")))


CASES = {
    "loop16_nn": dict(kind="nn", m=16, seed=1, views=[(0.0, np.pi / 2), (0.7, 0.9), (2.3, 0.2)], res=(20, 20, 24),
                      nl=12, rows=None),
    "c0_nn": dict(kind="nn", m=64, seed=1, views=[(0.35, 1.1)], res=(64, 64, 64), nl=8, rows=(24, 40)),
    "c0_imaging": dict(kind="imaging", m=64, seed=1, views=[(1.2, 0.6)], res=(64, 64, 64), nl=1, rows=None,
                       idx_rows=(24, 40)),
    "corona40_pb": dict(kind="thomson", m=40, seed=3, views=[(0.0, np.pi / 2), (0.6, 0.05)], res=(48, 40, 36)),
    "tet12_linear": dict(kind="tet", m=12, seed=9, jitter=0.3, views=[(0.0, np.pi / 2), (0.8, 1.0)],
                         res=(32, 32, 24), nl=10),
}


def make(name, F, O, tmpdir):
    """The case's inputs and the oracle's outputs, as a dict of arrays."""
    c = CASES[name]
    nx, ny, nz = c["res"]
    out = {"views": np.array(c["views"], np.float64), "res": np.array(c["res"], np.int32)}
    if c["kind"] == "thomson":
        d = F.gen_corona(c["m"], seed=c["seed"], sigma=0.1)
        out["input_digest"] = digest(d["x"], d["y"], d["z"], d["rho"])
        rt, rp = O.thomson_emission(1, d["x"], d["y"], d["z"], d["rho"])
        for v, (l, b) in enumerate(c["views"]):
            img, idx, meta, _ = O.render_thomson(1, d["x"], d["y"], d["z"], rt, rp, l, b, nx, ny, nz, want_idx=True)
            out[f"image{v}"], out[f"idx{v}"], out[f"meta{v}"] = img, idx, meta
        return out
    d, dig = loop_inputs(F, c["m"], c["seed"], **({"jitter": c["jitter"]} if "jitter" in c else {}))
    tp = table(F, tmpdir)
    tab = O.read_table(tp)
    out["input_digest"] = dig
    out["table_digest"] = file_digest(tp)
    nl = c["nl"]
    width = 0.2 * C / tab["lambda0"]
    out["nl"], out["width"] = np.int32(nl), np.float64(width)
    imaging = c["kind"] == "imaging"
    pk, fw = O.emission(1, d["rho"], d["T"], tab, imaging=imaging)
    if c["kind"] == "tet":
        from scipy.spatial import Delaunay
        tri = Delaunay(np.stack([d["x"], d["y"], d["z"]], 1).astype(np.float64))
        for v, (l, b) in enumerate(c["views"]):
            _, vert, meta = O.render_tet(d["x"], d["y"], d["z"], pk, fw, d["vx"], d["vy"], d["vz"], l, b, nx, ny, nz,
                                         1, 0.0, tab["lambda0"], tri=tri)
            img, _, _ = O.render_tet(d["x"], d["y"], d["z"], pk, fw, d["vx"], d["vy"], d["vz"], l, b, nx, ny, nz, nl,
                                     width, tab["lambda0"], linear=True, tri=tri)
            out[f"image{v}"], out[f"vert{v}"], out[f"meta{v}"] = img, vert, meta
        return out
    rows = c["rows"]
    if rows:
        out["rows"] = np.array(rows, np.int32)  # the image holds these rows only
    for v, (l, b) in enumerate(c["views"]):
        img, idx, meta, _ = O.render_nn(1, d["x"], d["y"], d["z"], pk, fw, d["vx"], d["vy"], d["vz"], l, b, nx, ny,
                                        nz, nl, width, tab["lambda0"], want_idx=True, rows=rows)
        if rows:
            img = img.reshape(ny, -1)[rows[0]:rows[1]].ravel()
        ir = rows or c.get("idx_rows")
        if ir:
            out["idx_rows"] = np.array(ir, np.int32)
            idx = idx.reshape(ny, -1)[ir[0]:ir[1]].ravel()
        out[f"image{v}"], out[f"idx{v}"], out[f"meta{v}"] = img, idx, meta
    return out


def main():
    import tempfile

    import fomo_amd as F
    from oracle import oracle as O

    with tempfile.TemporaryDirectory() as tmp:
        for name in CASES:
            arrs = make(name, F, O, tmp)
            path = os.path.join(HERE, name + ".npz")
            np.savez_compressed(path, **arrs)
            print(f"{name}: {os.path.getsize(path)} bytes", flush=True)


if __name__ == "__main__":
    main()
