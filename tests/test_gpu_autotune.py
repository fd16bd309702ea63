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
    assert r["cell_scale"] in (0.7, 1.0, 1.4) and r["subdiv"] in (1, 2, 3) and r["tile"] in (8, 16, 32, 64)
    assert 0 < r["ms"] <= r["ms_before"]
    if (r["cell_scale"], r["subdiv"], r["tile"]) != (1.0, 2, 8):
        assert r["ms"] < 0.97 * r["ms_before"]
    after = [F.render_nn(em, l, b, nx, ny, nz, nl, width) for (l, b, nx, ny, nz, nl) in views]
    for (a, ma), (c, mc) in zip(before, after):
        assert np.array_equal(a, c) and np.array_equal(ma, mc)
    assert np.array_equal(F.nn_indices(ds, 0.3, 1.2, 48, 40, 50)[0], idx_before)
    em.close(), T.close(), ds.close()


@pytest.mark.parametrize("thomson", [False, True])
def test_ray_tiles_keep_nn_choices(engine, thomson):
    """Every ray tile of the NN passes (including the shape of the waves' shared octree walks in
    the exact pass) gives the same NN choices, on a corona whose far field and disk hole reach the
    exact pass."""
    d = F.gen_corona(48, seed=6) if thomson else F.gen_loop_cube(32, seed=18, jitter=0.3)
    ds = F.Dataset(d["x"], d["y"], d["z"], [d["rho"]])
    ref = F.nn_indices(ds, 0.7, 1.1, 70, 53, 41, thomson=thomson)[0]
    for w in (8, 32, 64):
        ds.set_ray_tile(w)
        assert np.array_equal(F.nn_indices(ds, 0.7, 1.1, 70, 53, 41, thomson=thomson)[0], ref), w
    with pytest.raises(F.FomoError):
        ds.set_ray_tile(12)
    ds.close()
