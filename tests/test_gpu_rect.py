"""GPU tests of the rectilinear index: datasets whose points form a product grid are indexed by
their three sorted axes instead of the cell list and candidate lookup grid.

Bars: NN choices bit-exact with the oracle, with the same dataset indexed as scattered points
(cell list + lookup grid), and with the exact octree query alone; images and batches identical
across the index kinds; datasets that are not product grids are not detected as one.
"""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

C = 2.99792458e8
TOL_IMG = 1e-5
VIEWS = [(0.0, np.pi / 2), (0.3, 1.1), (2.5, 0.2), (np.pi / 4, np.pi / 4), (0.0, 0.0), (1.7, -0.6)]


def _product_grid(xs, ys, zs, seed):
    """Points of xs x ys x zs in a random original order, with loop-like fields."""
    rng = np.random.default_rng(seed)
    X, Y, Z = np.meshgrid(np.asarray(xs, np.float32), np.asarray(ys, np.float32), np.asarray(zs, np.float32),
                          indexing="ij")
    order = rng.permutation(X.size)
    d = {"x": X.ravel()[order], "y": Y.ravel()[order], "z": Z.ravel()[order]}
    n = X.size
    d["rho"] = (10 ** rng.uniform(8.8, 9.6, n)).astype(np.float32)
    d["T"] = (10 ** rng.uniform(5.7, 6.3, n)).astype(np.float32)
    for k in ("vx", "vy", "vz"):
        d[k] = rng.normal(0.0, 2.0e4, n).astype(np.float32)
    return d


def _dataset(F, d):
    return F.Dataset(d["x"], d["y"], d["z"], [d["rho"], d["T"], d["vx"], d["vy"], d["vz"]])


def _oracle_idx(d, l, b, nx, ny, nz):
    one = np.ones(d["x"].size)
    return O.render_nn(1, d["x"], d["y"], d["z"], one, one, d["vx"], d["vy"], d["vz"], l, b, nx, ny, nz, 1, 1e5,
                       171.0, want_idx=True)[1]


def _grids():
    rng = np.random.default_rng(3)
    stretched = 0.1 * (1.25 ** np.arange(18) - 1.0)
    random_axis = np.unique(rng.uniform(-2.0, 3.0, 40).astype(np.float32))[:21]
    return {
        "uniform": (np.linspace(0.0, 10.0, 24), np.linspace(0.0, 10.0, 22), np.linspace(0.0, 10.0, 20)),
        "stretched": (stretched, random_axis, np.linspace(-1.0, 1.0, 15)),
        "flat": (np.linspace(0.0, 5.0, 30), np.linspace(0.0, 5.0, 28), np.array([0.0])),
    }


def test_loop_cube_is_rectilinear(engine):
    F = engine
    d = F.gen_loop_cube(40, seed=2)
    ds = _dataset(F, d)
    assert ds.rect_axes() == (40, 40, 40)
    info = ds.lookup_info()
    assert info["subdiv"] == 0 and info["cells"] == 0, info  # no candidate grid is built
    ds.set_rectilinear(False)
    assert ds.rect_axes() is None and ds.lookup_info()["subdiv"] == 2
    ds.set_rectilinear(True)
    assert ds.rect_axes() == (40, 40, 40)
    ds.close()


@pytest.mark.parametrize("kind", ["uniform", "stretched", "flat"])
def test_rect_matches_scattered_exact_and_oracle(engine, kind):
    F = engine
    xs, ys, zs = _grids()[kind]
    d = _product_grid(xs, ys, zs, seed=5)
    ds = _dataset(F, d)
    assert ds.rect_axes() == (len(xs), len(ys), len(zs))
    nx, ny, nz = 41, 37, 44
    rect = [F.nn_indices(ds, l, b, nx, ny, nz)[0] for (l, b) in VIEWS]
    rect_t = [F.nn_indices(ds, l, b, nx, ny, nz, thomson=True)[0] for (l, b) in VIEWS[:3]]
    ds.set_rectilinear(False)
    scat = [F.nn_indices(ds, l, b, nx, ny, nz)[0] for (l, b) in VIEWS]
    ds.set_lookup(0)
    exact = [F.nn_indices(ds, l, b, nx, ny, nz)[0] for (l, b) in VIEWS]
    exact_t = [F.nn_indices(ds, l, b, nx, ny, nz, thomson=True)[0] for (l, b) in VIEWS[:3]]
    for k, (l, b) in enumerate(VIEWS):
        assert np.array_equal(rect[k], scat[k]), f"view {(l, b)}: {(rect[k] != scat[k]).sum()} differ from scattered"
        assert np.array_equal(rect[k], exact[k]), f"view {(l, b)}: {(rect[k] != exact[k]).sum()} differ from exact"
        ridx = _oracle_idx(d, l, b, nx, ny, nz)
        assert np.array_equal(rect[k], ridx), f"view {(l, b)}: {(rect[k] != ridx).sum()} differ from the oracle"
    for k in range(3):
        assert np.array_equal(rect_t[k], exact_t[k])
    ds.close()


def test_rect_midpoint_ties(engine):
    """An integer grid 0..16 seen along z (l = b = 0: the rotation is the identity) with 33 pixels
    and samples per axis: every other sample lies exactly halfway between grid values, so up to 8
    points tie at the nearest distance and the lowest original index decides."""
    F = engine
    v = np.arange(17, dtype=np.float64)
    d = _product_grid(v, v, v, seed=9)
    ds = _dataset(F, d)
    assert ds.rect_axes() == (17, 17, 17)
    nx = ny = nz = 33
    idx, meta = F.nn_indices(ds, 0.0, 0.0, nx, ny, nz)
    ridx = _oracle_idx(d, 0.0, 0.0, nx, ny, nz)
    assert np.array_equal(idx, ridx), f"{(idx != ridx).sum()} NN choices differ"
    # the samples on midpoints: their nearest distance is attained by several points
    k, i, j = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    halves = (i % 2) + (j % 2) + (k % 2)
    found = idx.reshape(nz, ny, nx)
    tied = int((found[halves > 0] >= 0).sum())
    print(f"{tied} samples with a point in their box lie on midpoints of 1 to 3 axes; all resolved as the oracle "
          f"resolves them")
    assert tied > 10000
    ds.set_rectilinear(False)
    assert np.array_equal(F.nn_indices(ds, 0.0, 0.0, nx, ny, nz)[0], idx)
    ds.close()


def test_not_product_grids_are_scattered(engine):
    F = engine
    d = F.gen_loop_cube(12, seed=1)
    keys = ("x", "y", "z", "rho", "T", "vx", "vy", "vz")
    # one point missing
    d1 = {k: d[k][1:] for k in keys}
    # one point duplicated (n is no longer the product of the axis lengths)
    d2 = {k: np.concatenate([d[k], d[k][:1]]) for k in keys}
    # the product of the axis lengths is n, but one grid position is taken twice and one is empty
    d3 = {k: d[k].copy() for k in keys}
    same = np.flatnonzero((d3["x"] == d3["x"][0]) & (d3["y"] == d3["y"][0]))
    d3["z"][same[1]] = d3["z"][same[0]]
    for dd in (d1, d2, d3):
        ds = _dataset(F, dd)
        assert ds.rect_axes() is None
        assert ds.lookup_info()["subdiv"] == 2
        l, b = 0.4, 0.9
        assert np.array_equal(F.nn_indices(ds, l, b, 20, 18, 22)[0], _oracle_idx(dd, l, b, 20, 18, 22))
        ds.close()


def test_rect_batch_and_renders_match_scattered(engine, fe9_table):
    """The fused pass (batch of 4 views x 2 lines, and a single-line view above the fused-pass
    threshold) and the plain render give the same images on both index kinds; one line against
    the FP64 oracle. Switching the index kind re-orders the emission cubes already made."""
    F = engine
    path, tab = fe9_table
    d = F.gen_loop_cube(32, seed=6)
    ds = _dataset(F, d)
    assert ds.rect_axes() == (32, 32, 32)
    T = F.Table.from_file(path)
    tabs2 = O.read_table(path)
    ems = [F.Emission.chianti(ds, T, tab["lambda0"], tab["atweight"]),
           F.Emission.chianti(ds, T, tab["lambda0"], tab["atweight"], imaging=True)]
    width = 0.2 * C / tab["lambda0"]
    ls, bs = [0.0, 0.5, 1.3, 2.9], [np.pi / 2, np.pi / 3, 0.4, 1.0]
    nx, ny, nz, nl = 48, 44, 40, 24
    bt = F.Batch(ems[:1] * 2, ls, bs, nx, ny, nz, nl, width)
    bt.run()
    imgs_rect, meta_rect = bt.read()
    bt.close()
    big_rect, _ = F.render_nn(ems[0], 0.7, 0.8, 260, 256, 24, 30, width)
    plain_rect, _ = F.render_nn(ems[0], 1.1, 0.5, 40, 36, 30, 20, width)
    ds.set_rectilinear(False)
    bt = F.Batch(ems[:1] * 2, ls, bs, nx, ny, nz, nl, width)
    bt.run()
    imgs_scat, meta_scat = bt.read()
    bt.close()
    assert np.array_equal(imgs_rect, imgs_scat)
    assert np.array_equal(meta_rect, meta_scat)
    assert np.array_equal(big_rect, F.render_nn(ems[0], 0.7, 0.8, 260, 256, 24, 30, width)[0])
    assert np.array_equal(plain_rect, F.render_nn(ems[0], 1.1, 0.5, 40, 36, 30, 20, width)[0])
    pk, fw = O.emission(1, d["rho"], d["T"], tabs2)
    ref = O.render_nn(1, d["x"], d["y"], d["z"], pk, fw, d["vx"], d["vy"], d["vz"], ls[1], bs[1], nx, ny, nz, nl,
                      width, tab["lambda0"])[0]
    err = O.rel_l2(imgs_rect[1, 0].ravel(), ref)
    print(f"rectilinear batch view 1 rel L2 vs FP64 oracle {err:.2e}")
    assert err <= TOL_IMG
    for e in ems:
        e.close()
    T.close(), ds.close()


def _masked(F, keep, xs, ys, zs, seed):
    d = _product_grid(xs, ys, zs, seed)
    X, Y, Z = d["x"].astype(np.float64), d["y"].astype(np.float64), d["z"].astype(np.float64)
    m = keep(X, Y, Z)
    return {k: v[m] for k, v in d.items()}


@pytest.mark.parametrize("kind", ["shell", "random"])
def test_masked_grid_matches_scattered_exact_and_oracle(engine, kind):
    """A product grid with holes (a spherical shell cut from a Cartesian mesh, as the corona; or a
    random half of the nodes) is indexed as a masked grid: the unfused NN pass tries the grid's
    nearest node first and takes the candidate grid when it holds no point. NN choices equal the
    same data indexed as scattered points only, the exact query alone and the oracle, in NN and
    Thomson views."""
    F = engine
    v = np.linspace(-3.0, 3.0, 31)
    rng = np.random.default_rng(17)
    if kind == "shell":
        keep = lambda X, Y, Z: ((X * X + Y * Y + Z * Z) > 1.0) & ((X * X + Y * Y + Z * Z) <= 9.0)
    else:
        mask = rng.random(31 ** 3) < 0.5
        keep = lambda X, Y, Z: mask[np.argsort(np.argsort(X + 1e3 * Y + 1e6 * Z))]
    d = _masked(F, keep, v, v, v, seed=4)
    ds = _dataset(F, d)
    assert ds.grid_info() == ("masked", (31, 31, 31))
    assert ds.rect_axes() is None
    nx, ny, nz = 40, 36, 42
    views = VIEWS[:4]
    got = [F.nn_indices(ds, l, b, nx, ny, nz)[0] for (l, b) in views]
    got_t = [F.nn_indices(ds, l, b, nx, ny, nz, thomson=True)[0] for (l, b) in views[:2]]
    ds.set_rectilinear(False)
    assert ds.grid_info()[0] == "scattered"
    scat = [F.nn_indices(ds, l, b, nx, ny, nz)[0] for (l, b) in views]
    ds.set_lookup(0)
    exact = [F.nn_indices(ds, l, b, nx, ny, nz)[0] for (l, b) in views]
    exact_t = [F.nn_indices(ds, l, b, nx, ny, nz, thomson=True)[0] for (l, b) in views[:2]]
    for k, (l, b) in enumerate(views):
        assert np.array_equal(got[k], scat[k]), f"view {(l, b)}: {(got[k] != scat[k]).sum()} differ from scattered"
        assert np.array_equal(got[k], exact[k]), f"view {(l, b)}: {(got[k] != exact[k]).sum()} differ from exact"
        ridx = _oracle_idx(d, l, b, nx, ny, nz)
        assert np.array_equal(got[k], ridx), f"view {(l, b)}: {(got[k] != ridx).sum()} differ from the oracle"
    for k in range(2):
        assert np.array_equal(got_t[k], exact_t[k])
    ds.close()


def test_corona_is_masked_grid(engine):
    """The corona generator's points are a Cartesian mesh cut to 1 < r <= 3: a masked grid. Its
    Thomson renders equal those with the dataset indexed as scattered points, bit for bit."""
    F = engine
    c = F.gen_corona(48, seed=5)
    ds = F.Dataset(c["x"], c["y"], c["z"], [c["rho"]])
    kind, axes = ds.grid_info()
    # the mesh's outermost planes hold no point of the shell when no node lies on an axis
    assert kind == "masked" and all(44 <= a <= 48 for a in axes), (kind, axes)
    em = F.Emission.thomson(ds)
    views = [(0.0, np.pi / 2), (0.9, 0.3), (np.pi / 4, np.pi / 2)]
    imgs = [F.render_thomson(em, l, b, 64, 60, 48)[0] for (l, b) in views]
    ds.set_rectilinear(False)
    for (l, b), img in zip(views, imgs):
        assert np.array_equal(F.render_thomson(em, l, b, 64, 60, 48)[0], img)
    em.close(), ds.close()
