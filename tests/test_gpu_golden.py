"""The HIP engine against the committed golden fixtures (tests/golden/*.npz): the same bars as the
live-oracle parity tests (NN choices and tetra vertices bit-exact, view meta bit-exact, images
within relative L2 1e-5 of the FP64 oracle), but against outputs fixed in the repository."""
import os
import sys
import tempfile

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
import make_goldens as G  # noqa: E402

pytestmark = pytest.mark.gpu
TOL_IMG = 1e-5


def load(name):
    with np.load(os.path.join(HERE, "golden", name + ".npz")) as z:
        return {k: z[k] for k in z.files}


def rel_l2(a, ref):
    a, ref = np.asarray(a, np.float64), np.asarray(ref, np.float64)
    return float(np.linalg.norm(a - ref) / np.linalg.norm(ref))


@pytest.mark.parametrize("name", sorted(G.CASES))
def test_engine_matches_golden(engine, name):
    F = engine
    c = G.CASES[name]
    want = load(name)
    nx, ny, nz = c["res"]
    print(f"golden {name}: {c['kind']}, {len(c['views'])} views, {nx}x{ny}x{nz}", flush=True)
    if c["kind"] == "thomson":
        d = F.gen_corona(c["m"], seed=c["seed"], sigma=0.1)
        assert G.digest(d["x"], d["y"], d["z"], d["rho"]) == str(want["input_digest"])
        ds = F.Dataset(d["x"], d["y"], d["z"], [d["rho"]])
        em = F.Emission.thomson(ds)
        try:
            for v, (l, b) in enumerate(c["views"]):
                img, meta = F.render_thomson(em, l, b, nx, ny, nz)
                idx, _ = F.nn_indices(ds, l, b, nx, ny, nz, thomson=True)
                assert np.array_equal(idx, want[f"idx{v}"]), f"view {v}: {(idx != want[f'idx{v}']).sum()} NN differ"
                assert np.array_equal(meta[:7], want[f"meta{v}"][:7])
                assert rel_l2(img, want[f"image{v}"]) <= TOL_IMG
        finally:
            em.close(), ds.close()
        return
    d, dig = G.loop_inputs(F, c["m"], c["seed"], **({"jitter": c["jitter"]} if "jitter" in c else {}))
    assert dig == str(want["input_digest"])
    with tempfile.TemporaryDirectory() as tmp:
        tp = G.table(F, tmp)
        assert G.file_digest(tp) == str(want["table_digest"])
        T = F.Table.from_file(tp)
    ds = F.Dataset(d["x"], d["y"], d["z"], [d["rho"], d["T"], d["vx"], d["vy"], d["vz"]])
    if c["kind"] == "tet":
        ds.triangulate()
    em = F.Emission.chianti(ds, T, 171.0, 56.0, imaging=c["kind"] == "imaging")
    nl, width = int(want["nl"]), float(want["width"])
    try:
        for v, (l, b) in enumerate(c["views"]):
            if c["kind"] == "tet":
                vert, meta = F.tet_vertices(ds, l, b, nx, ny, nz)
                img, _ = F.render_tet(em, l, b, nx, ny, nz, nl, width, linear=True)
                assert np.array_equal(vert, want[f"vert{v}"]), f"view {v}: {(vert != want[f'vert{v}']).sum()} differ"
                assert np.array_equal(meta[:6], want[f"meta{v}"][:6])
                assert rel_l2(img, want[f"image{v}"]) <= TOL_IMG
                continue
            img, meta = F.render_nn(em, l, b, nx, ny, nz, nl, width)
            idx, _ = F.nn_indices(ds, l, b, nx, ny, nz)
            if "rows" in want:
                r0, r1 = want["rows"]
                img = img.reshape(ny, -1)[r0:r1].ravel()
            if "idx_rows" in want:
                r0, r1 = want["idx_rows"]
                idx = idx.reshape(ny, -1)[r0:r1].ravel()
            assert np.array_equal(idx, want[f"idx{v}"]), f"view {v}: {(idx != want[f'idx{v}']).sum()} NN differ"
            assert np.array_equal(meta[:7], want[f"meta{v}"][:7])
            err = rel_l2(img, want[f"image{v}"])
            print(f"  view {v}: rel L2 {err:.2e}", flush=True)
            assert err <= TOL_IMG
    finally:
        em.close(), T.close(), ds.close()
