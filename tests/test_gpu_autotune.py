"""GPU test of the autotuner (fomo_autotune): it changes speed only. Renders and NN choices are
identical before and after it re-indexes the dataset with the configuration it picks."""
import numpy as np
import pytest

import fomo_amd as F

pytestmark = pytest.mark.gpu

C = 2.99792458e8


def test_autotune_keeps_results(engine, fe9_table):
    path, tab = fe9_table
    d = F.gen_loop_cube(40, seed=17, jitter=0.25)
    ds = F.Dataset(d["x"], d["y"], d["z"], [d["rho"], d["T"], d["vx"], d["vy"], d["vz"]])
    T = F.Table.from_file(path)
    em = F.Emission.chianti(ds, T, tab["lambda0"], tab["atweight"])
    width = 0.2 * C / tab["lambda0"]
    views = [(0.3, 1.2, 48, 40, 50, 30), (2.0, 0.4, 33, 37, 44, 12)]
    before = [F.render_nn(em, l, b, nx, ny, nz, nl, width) for (l, b, nx, ny, nz, nl) in views]
    idx_before = F.nn_indices(ds, 0.3, 1.2, 48, 40, 50)[0]
    r = F.autotune(em, 0.3, 1.2, 48, 40, 50, reps=2)
    assert r["cell_scale"] in (0.7, 1.0, 1.4) and r["subdiv"] in (1, 2, 3)
    assert 0 < r["ms"] <= r["ms_before"]
    if (r["cell_scale"], r["subdiv"]) != (1.0, 2):
        assert r["ms"] < 0.97 * r["ms_before"]
    after = [F.render_nn(em, l, b, nx, ny, nz, nl, width) for (l, b, nx, ny, nz, nl) in views]
    for (a, ma), (c, mc) in zip(before, after):
        assert np.array_equal(a, c) and np.array_equal(ma, mc)
    assert np.array_equal(F.nn_indices(ds, 0.3, 1.2, 48, 40, 50)[0], idx_before)
    em.close(), T.close(), ds.close()
