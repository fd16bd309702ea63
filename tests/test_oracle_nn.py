"""Pin the oracle's nearest-neighbour query (row A10) against brute force and scipy.

Semantics restated from fomo-nearestneighbour.cpp:171-193 (Boost rtree query
nearest(target,1) && within(box)): the nearest point strictly inside the float box
[float(p-md), float(p+md)] by double squared distance of float coordinates to float(p);
ties -> lowest index (Boost's tie order is implementation-defined; parity unpinned there).
"""
import numpy as np
import pytest
from scipy.spatial import cKDTree

from oracle import oracle as O


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_kdtree_equals_bruteforce_random(seed):
    rng = np.random.default_rng(seed)
    n = 3000
    x, y, z = (rng.uniform(-2, 2, n).astype(np.float32) for _ in range(3))
    m = 4000
    px, py, pz = (rng.uniform(-2.5, 2.5, m) for _ in range(3))
    for md in (0.05, 0.2, 1.0, 10.0):
        a = O.nn_query(x, y, z, px, py, pz, md, "kdtree")
        b = O.nn_query(x, y, z, px, py, pz, md, "brute")
        assert np.array_equal(a, b)
    # with a small box many targets find nothing
    a = O.nn_query(x, y, z, px, py, pz, 0.02, "kdtree")
    assert (a < 0).sum() > 100


def test_ties_go_to_lowest_index():
    # 4 points symmetric around the origin, duplicated; targets at the exact centre
    x = np.array([1, -1, 0, 0, 1, -1], np.float32)
    y = np.array([0, 0, 1, -1, 0, 0], np.float32)
    z = np.zeros(6, np.float32)
    px, py, pz = np.zeros(5), np.zeros(5), np.zeros(5)
    a = O.nn_query(x, y, z, px, py, pz, 5.0, "kdtree")
    b = O.nn_query(x, y, z, px, py, pz, 5.0, "brute")
    assert np.all(a == 0) and np.all(b == 0)


def test_box_is_strict():
    x = np.array([0.5], np.float32)
    y = np.array([0.0], np.float32)
    z = np.array([0.0], np.float32)
    # box [-0.5, 0.5] in x: a point exactly on the face is outside
    a = O.nn_query(x, y, z, np.array([0.0]), np.array([0.0]), np.array([0.0]), 0.5, "kdtree")
    assert a[0] == -1
    a = O.nn_query(x, y, z, np.array([0.0]), np.array([0.0]), np.array([0.0]), 0.5000001, "kdtree")
    assert a[0] == 0


def test_matches_scipy_ckdtree_when_box_is_large():
    rng = np.random.default_rng(5)
    n = 20000
    pts = rng.normal(size=(n, 3)).astype(np.float32)
    q = rng.normal(size=(5000, 3))
    got = O.nn_query(pts[:, 0], pts[:, 1], pts[:, 2], q[:, 0], q[:, 1], q[:, 2], 1e6, "kdtree")
    tree = cKDTree(pts.astype(np.float64))
    qf = q.astype(np.float32).astype(np.float64)  # the reference queries the float target
    dist, idx = tree.query(qf, k=2)
    d_got = np.sum((pts[got].astype(np.float64) - qf) ** 2, axis=1)
    np.testing.assert_allclose(d_got, dist[:, 0] ** 2, rtol=1e-12)
    unique = dist[:, 1] > dist[:, 0] * (1 + 1e-9)
    assert np.array_equal(got[unique], idx[unique, 0])


def test_multi_line_render_equals_single_line_renders_and_flags_ties():
    """render_nn_multi shares one query per sample between lines: each line's image equals its own
    render_nn, and the tie flags mark exactly the samples where a second point in the box lies at
    the nearest distance (a structured grid seen along an axis has many)."""
    C = 2.99792458e8
    m = 9
    a = np.linspace(-4, 4, m).astype(np.float32)
    zz, yy, xx = np.meshgrid(a, a, a, indexing="ij")
    x, y, z = xx.ravel(), yy.ravel(), zz.ravel()
    n = x.size
    rng = np.random.default_rng(3)
    peaks = [rng.uniform(0.5, 2, n) for _ in range(3)]
    fwhms = [rng.uniform(0.01, 0.03, n) for _ in range(3)]
    v = [rng.normal(0, 2e4, n).astype(np.float32) for _ in range(3)]
    width = 0.2 * C / 171.0
    for l, b, mode in ((0.0, np.pi / 2, 1), (0.7, 1.1, 0)):
        nx, ny, nz, nl = 13, 11, 17, 12
        imgs, idx, tie, meta, _ = O.render_nn_multi(mode, x, y, z, peaks, fwhms, *v, l, b, nx, ny, nz, nl, width,
                                                    171.0, rows=(2, 9))
        for g in range(3):
            ref, ridx, rmeta, _ = O.render_nn(mode, x, y, z, peaks[g], fwhms[g], *v, l, b, nx, ny, nz, nl, width, 171.0,
                                              want_idx=True, rows=(2, 9))
            assert np.array_equal(imgs[g], ref)
            band = ridx.reshape(ny, nx * nz)[2:9]
            assert np.array_equal(idx.reshape(ny, nx * nz)[2:9], band)
            assert np.array_equal(meta, rmeta)
        # tie flags by brute force over the band's samples
        p = O.view_samples(meta[:6], l, b, nx, ny, nz)
        sel = np.zeros((ny, nx, nz), bool)
        sel[2:9] = True
        pts = np.stack([x, y, z], 1).astype(np.float64)
        md = meta[6]
        want = []
        for q in zip(p[0][sel], p[1][sel], p[2][sel]):
            t = np.array(q, np.float32).astype(np.float64)
            lo = np.array([np.float32(c - md) for c in q], np.float64)
            hi = np.array([np.float32(c + md) for c in q], np.float64)
            inb = np.all((pts > lo) & (pts < hi), 1)
            d = (((pts[:, 0] - t[0]) ** 2 + (pts[:, 1] - t[1]) ** 2) + (pts[:, 2] - t[2]) ** 2)[inb]
            want.append(inb.any() and (d == d.min()).sum() > 1)
        got = tie.reshape(ny, nx, nz)[sel].astype(bool)
        assert np.array_equal(got, np.array(want))
        if l == 0.0:
            assert got.sum() > 0
