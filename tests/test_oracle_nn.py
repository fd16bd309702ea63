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
