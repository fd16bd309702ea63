"""GPU edge cases of the nearest-neighbour render against the oracle: 2D datacubes (z = 0, as the
reference sets it, fomo-nearestneighbour.cpp:80), face-on views of them (zero z extent: path length
1, all samples of a ray at one position), datasets of 1 and 2 points (degenerate extents: an empty
search box), duplicated points (exact distance ties: lowest original index) and zero densities.
Bars: NN indices bit-exact; images within relative L2 1e-5 of the FP64 oracle."""
import numpy as np
import pytest

import fomo_amd as F
from oracle import oracle as O

pytestmark = pytest.mark.gpu

C = 2.99792458e8
TOL_IMG = 1e-5


def _check(ds, d, tab, em, l, b, nx, ny, nz, nl, imaging=False):
    width = 0.2 * C / tab["lambda0"]
    idx, meta = F.nn_indices(ds, l, b, nx, ny, nz)
    pk, fw = O.emission(1, d["rho"], d["T"], tab, imaging=imaging)
    ref, ridx, rmeta, _ = O.render_nn(1, d["x"], d["y"], d["z"], pk, fw, d["vx"], d["vy"], d["vz"], l, b, nx, ny, nz,
                                      nl, width, tab["lambda0"], want_idx=True)
    assert np.array_equal(meta[:6], rmeta[:6])
    assert np.array_equal(idx, ridx), f"{(idx != ridx).sum()} NN choices differ"
    img, m2 = F.render_nn(em, l, b, nx, ny, nz, nl, width)
    np.testing.assert_allclose(m2[7:10], rmeta[7:10], rtol=1e-12)
    if np.count_nonzero(ref) == 0:
        assert np.count_nonzero(img) == 0
    else:
        assert O.rel_l2(img, ref) <= TOL_IMG
    return idx, img, ref


def _fields(n, rng):
    return {"rho": (1e9 * (1 + 0.3 * rng.random(n))).astype(np.float32),
            "T": (1e6 * (1 + rng.random(n))).astype(np.float32),
            "vx": (2e4 * rng.standard_normal(n)).astype(np.float32),
            "vy": (2e4 * rng.standard_normal(n)).astype(np.float32),
            "vz": (2e4 * rng.standard_normal(n)).astype(np.float32)}


def _make(d, dim=3):
    return F.Dataset(d["x"], d["y"], d["z"] if dim == 3 else None,
                     [d["rho"], d["T"], d["vx"], d["vy"], d["vz"]], dim=dim)


@pytest.mark.parametrize("l,b,nl", [(0.0, 0.0, 12), (0.0, np.pi / 2, 12), (0.7, 0.4, 1)])
def test_2d_datacube(engine, fe9_table, l, b, nl):
    path, tab = fe9_table
    rng = np.random.default_rng(11)
    m = 30
    g = np.arange(m, dtype=np.float64) / (m - 1) * 10 - 5
    yy, xx = np.meshgrid(g, g, indexing="ij")
    d = {"x": (xx.ravel() + 0.1 * rng.uniform(-1, 1, m * m)).astype(np.float32),
         "y": (yy.ravel() + 0.1 * rng.uniform(-1, 1, m * m)).astype(np.float32)}
    d["z"] = np.zeros(m * m, np.float32)
    d.update(_fields(m * m, rng))
    ds = _make(d, dim=2)
    T = F.Table.from_file(path)
    em = F.Emission.chianti(ds, T, tab["lambda0"], tab["atweight"], imaging=(nl == 1))
    try:
        idx, img, ref = _check(ds, d, tab, em, l, b, 24, 20, 9, nl, imaging=(nl == 1))
        assert (idx >= 0).any()
    finally:
        em.close(), T.close(), ds.close()


@pytest.mark.parametrize("n", [1, 2, 5])
def test_tiny_datasets(engine, fe9_table, n):
    path, tab = fe9_table
    rng = np.random.default_rng(20 + n)
    d = {"x": rng.uniform(-1, 1, n).astype(np.float32), "y": rng.uniform(-1, 1, n).astype(np.float32),
         "z": rng.uniform(-1, 1, n).astype(np.float32)}
    d.update(_fields(n, rng))
    ds = _make(d)
    T = F.Table.from_file(path)
    em = F.Emission.chianti(ds, T, tab["lambda0"], tab["atweight"])
    try:
        for (l, b) in [(0.0, np.pi / 2), (1.1, 0.6)]:
            _check(ds, d, tab, em, l, b, 8, 6, 7, 10)
    finally:
        em.close(), T.close(), ds.close()


def test_duplicated_points_tie_to_lowest_index(engine, fe9_table):
    path, tab = fe9_table
    rng = np.random.default_rng(31)
    m = 12
    g = (np.arange(m) / (m - 1) * 10 - 5).astype(np.float32)
    zz, yy, xx = np.meshgrid(g, g, g, indexing="ij")
    base = {"x": xx.ravel(), "y": yy.ravel(), "z": zz.ravel()}
    n = base["x"].size
    # every point twice, the copy first in index order half of the time
    perm = rng.permutation(2 * n)
    d = {k: np.concatenate([v, v])[perm] for k, v in base.items()}
    d.update(_fields(2 * n, rng))
    ds = _make(d)
    T = F.Table.from_file(path)
    em = F.Emission.chianti(ds, T, tab["lambda0"], tab["atweight"])
    try:
        idx, _, _ = _check(ds, d, tab, em, 0.0, np.pi / 2, m, m, m, 10)
        # grid samples sit on grid points: each has two equidistant candidates
        assert (idx >= 0).all()
        twin = {}
        for i, p in enumerate(perm):
            twin.setdefault(p % n, []).append(i)
        lowest = {min(v) for v in twin.values()}
        assert set(np.unique(idx).tolist()) <= lowest
    finally:
        em.close(), T.close(), ds.close()


def test_zero_density_points(engine, fe9_table):
    path, tab = fe9_table
    rng = np.random.default_rng(41)
    n = 3000
    d = {"x": rng.uniform(-4, 4, n).astype(np.float32), "y": rng.uniform(-4, 4, n).astype(np.float32),
         "z": rng.uniform(-4, 4, n).astype(np.float32)}
    d.update(_fields(n, rng))
    d["rho"][::3] = 0.0  # log10 of zero density: the reference warns and the emission is 0
    ds = _make(d)
    T = F.Table.from_file(path)
    em = F.Emission.chianti(ds, T, tab["lambda0"], tab["atweight"])
    try:
        pk_dev, _ = em.read()
        assert (pk_dev[::3] == 0).all()
        _check(ds, d, tab, em, 0.5, 1.0, 20, 18, 16, 12)
    finally:
        em.close(), T.close(), ds.close()


@pytest.mark.parametrize("nl", [2, 64, 65, 128])
def test_single_line_fused_lambda_limits(engine, fe9_table, tmp_path, nl):
    """The single-line fused pass at its lambda limits (2 and 64 pixels in one pass, 65 and 128 in
    two) on a 256 x 256 view: equal to the same line in a two-line batch (one line per lane, or
    k_nn + the integration passes beyond 64 pixels) bit for bit, and within 1e-5 of the FP64
    oracle."""
    path, tab = fe9_table
    rng = np.random.default_rng(50 + nl)
    n = 1500
    d = {"x": rng.uniform(-3, 3, n).astype(np.float32), "y": rng.uniform(-3, 3, n).astype(np.float32),
         "z": rng.uniform(-3, 3, n).astype(np.float32)}
    d.update(_fields(n, rng))
    ds = _make(d)
    T = F.Table.from_file(path)
    T2 = F.Table.from_file(F.gen_goft_table(str(tmp_path / "b.dat"), lambda0=193.5, atweight=56.0, peak_logt=6.2))
    em = F.Emission.chianti(ds, T, tab["lambda0"], tab["atweight"])
    em2 = F.Emission.chianti(ds, T2, 193.5, 56.0)
    try:
        l, b = 0.9, 0.7
        _, _, ref = _check(ds, d, tab, em, l, b, 256, 256, 6, nl)
        img, meta = F.render_nn(em, l, b, 256, 256, 6, nl, 0.2 * C / tab["lambda0"])
        bt = F.Batch([em, em2], [l], [b], 256, 256, 6, nl, 0.2 * C / tab["lambda0"])
        bt.run()
        out, _ = bt.read()
        bt.close()
        assert np.array_equal(out[0, 0].ravel(), img)
        assert np.count_nonzero(ref) > 0
    finally:
        em.close(), em2.close(), T.close(), T2.close(), ds.close()


def test_single_line_fused_2d_datacube(engine, fe9_table):
    """A 2D datacube (z = 0) seen obliquely on a 256 x 256 view: the single-line fused pass against
    the FP64 oracle (NN indices bit-exact through the separate NN pass)."""
    path, tab = fe9_table
    rng = np.random.default_rng(61)
    m = 40
    g = np.arange(m, dtype=np.float64) / (m - 1) * 10 - 5
    yy, xx = np.meshgrid(g, g, indexing="ij")
    d = {"x": (xx.ravel() + 0.1 * rng.uniform(-1, 1, m * m)).astype(np.float32),
         "y": (yy.ravel() + 0.1 * rng.uniform(-1, 1, m * m)).astype(np.float32)}
    d["z"] = np.zeros(m * m, np.float32)
    d.update(_fields(m * m, rng))
    ds = _make(d, dim=2)
    T = F.Table.from_file(path)
    em = F.Emission.chianti(ds, T, tab["lambda0"], tab["atweight"])
    try:
        _check(ds, d, tab, em, 0.7, 0.4, 256, 256, 5, 30)
    finally:
        em.close(), T.close(), ds.close()


def test_streamed_ingest_equals_bulk_upload(engine, fe9_table):
    """The loaders' streaming ingest (fomo_dataset_begin / append / commit) gives the dataset that
    one bulk upload gives: chunks of uneven sizes appended out of order render bit-identical
    images and meta; incomplete or overlapping chunks are refused at commit, out-of-range
    chunks at append, and an uncommitted dataset cannot be used."""
    import ctypes
    F = engine
    path, tab = fe9_table
    d = F.gen_loop_cube(30, seed=13, jitter=0.2)
    names = ("rho", "T", "vx", "vy", "vz")
    n = d["x"].size
    rng = np.random.default_rng(3)
    cuts = np.sort(rng.choice(np.arange(1, n), 9, replace=False))
    bounds = list(zip(np.r_[0, cuts], np.r_[cuts, n]))
    rng.shuffle(bounds)
    chunks = [(a, d["x"][a:b], d["y"][a:b], d["z"][a:b], [d[k][a:b] for k in names]) for a, b in bounds]
    ds_s = F.Dataset.streamed(n, chunks, nvars=5)
    ds_b = F.Dataset(d["x"], d["y"], d["z"], [d[k] for k in names])
    T = F.Table.from_file(path)
    ems = [F.Emission.chianti(ds, T, tab["lambda0"], tab["atweight"]) for ds in (ds_s, ds_b)]
    w = 0.2 * 2.99792458e8 / tab["lambda0"]
    for l, b in [(0.3, 1.0), (2.0, 0.4)]:
        a, ma = F.render_nn(ems[0], l, b, 40, 36, 44, 24, w)
        c, mc = F.render_nn(ems[1], l, b, 40, 36, 44, 24, w)
        assert np.array_equal(a, c) and np.array_equal(ma, mc)
    for e in ems:
        e.close()
    L = F.lib()
    out = ctypes.c_int()
    assert L.fomo_dataset_begin(100, 3, 0, ctypes.byref(out)) == 0
    x = np.zeros(60, np.float32)
    fp = x.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
    assert L.fomo_dataset_append(out.value, 50, 60, fp, fp, fp, None) == -1  # past n
    assert L.fomo_dataset_append(out.value, 0, 60, fp, fp, fp, None) == 0
    assert L.fomo_dataset_commit(out.value, 0.0) == -5  # 40 points missing
    assert L.fomo_dataset_append(out.value, 50, 50, fp, fp, fp, None) == 0
    assert L.fomo_dataset_commit(out.value, 0.0) == -5  # points 50..59 written twice
    assert L.fomo_dataset_info(out.value, None, None, None) == -3  # not usable before its commit
    assert L.fomo_dataset_destroy(out.value) == 0
    T.close()
    ds_s.close()
    ds_b.close()
