"""CPU tests of the host Delaunay builder (fomo_amd/csrc/delaunay.h, predicates.h), the structure
of the tetra render path (replacing CGAL's Delaunay_triangulation_3, fomo-CGAL.cpp:33-73).

* predicates: signs equal exact rational arithmetic on near-degenerate inputs;
* general position: the tets equal scipy.spatial.Delaunay's (Qhull, an independent
  implementation) as sets;
* exact grids (every cube cospherical): positively oriented tets, volumes summing to the hull
  volume, no point strictly inside any circumsphere;
* duplicates: one vertex per location, carrying the last input index.
"""
import os
import subprocess
from fractions import Fraction

import numpy as np
import pytest
from scipy.spatial import ConvexHull, Delaunay, cKDTree

import fomo_amd as F

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _exact_orient(p):
    a, b, c, d = [[Fraction(v) for v in q] for q in p]
    m = [[b[i] - a[i] for i in range(3)], [c[i] - a[i] for i in range(3)], [d[i] - a[i] for i in range(3)]]
    det = (m[0][0] * (m[1][1] * m[2][2] - m[1][2] * m[2][1]) - m[0][1] * (m[1][0] * m[2][2] - m[1][2] * m[2][0])
           + m[0][2] * (m[1][0] * m[2][1] - m[1][1] * m[2][0]))
    return (det > 0) - (det < 0)


def _det(m):
    n = len(m)
    if n == 1:
        return m[0][0]
    return sum(((-1) ** j) * m[0][j] * _det([r[:j] + r[j + 1:] for r in m[1:]]) for j in range(n))


def _exact_insphere5(p):
    rows = []
    for q in p:
        x, y, z = (Fraction(v) for v in q)
        rows.append([x, y, z, x * x + y * y + z * z, Fraction(1)])
    d = _det(rows)
    return (d > 0) - (d < 0)


def test_predicates_match_exact_rationals(tmp_path):
    exe = str(tmp_path / "pred_test")
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-Wall", "-Werror", "-o", exe,
                    os.path.join(ROOT, "tests", "cpp", "pred_test.cpp")], check=True)
    rng = np.random.default_rng(5)
    cases, want = [], []
    for t in range(300):
        if t % 2 == 0:
            # nearly coplanar: d on the plane of a,b,c, then nudged by a few ulps (or exactly on it)
            a, b, c = rng.normal(size=(3, 3)) * 10
            u, v = rng.random(2)
            d = a + u * (b - a) + v * (c - a)
            if t % 4 == 0:
                d = d + np.spacing(d) * rng.integers(-3, 4, 3)
            pts = [a, b, c, d]
            cases.append("O " + " ".join(repr(float(x)) for q in pts for x in q))
            want.append(_exact_orient(pts))
        else:
            # nearly cospherical: 5 points on a sphere (rounded to double), some exactly on a grid
            if t % 6 == 1:
                pts = [np.array(q, float) for q in [(0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1), (1, 1, 1)]]
            else:
                cen, r = rng.normal(size=3), 1 + rng.random()
                dirs = rng.normal(size=(5, 3))
                pts = list(cen + r * dirs / np.linalg.norm(dirs, axis=1)[:, None])
            cases.append("S " + " ".join(repr(float(x)) for q in pts for x in q))
            want.append(_exact_insphere5(pts))
    r = subprocess.run([exe], input="\n".join(cases) + "\n", capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    got = [int(v) for v in r.stdout.split()]
    assert got == want
    assert 0 in want and 1 in want and -1 in want  # the degenerate and both signs are exercised


def test_tetra_walk_exact_orient_matches_rationals(tmp_path):
    """The tetra walk's exact orientation test (float vertices, double sample point; the fallback
    for samples the filtered FP64 test cannot place) against exact rational arithmetic: points
    exactly on a face plane, a few ulps off it, with coordinates of mixed magnitudes (differences
    of floats that do not fit one double), and random ones."""
    exe = str(tmp_path / "pred_test")
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-Wall", "-Werror", "-o", exe,
                    os.path.join(ROOT, "tests", "cpp", "pred_test.cpp")], check=True)
    rng = np.random.default_rng(11)
    cases, want = [], []
    for t in range(600):
        scale = 10.0 ** rng.integers(-6, 7) if t % 3 == 0 else 10.0
        a, b, c = (rng.normal(size=(3, 3)) * scale).astype(np.float32).astype(np.float64)
        if t % 5 == 0:
            a[0] = np.float32(rng.normal() * 1e-12)  # coordinates whose float differences overflow a double
        if t % 4 == 0:
            # exactly on the plane: a lattice plane z = const with float vertices and a double point
            zc = float(np.float32(rng.normal() * scale))
            a[2] = b[2] = c[2] = zc
            d = np.array([rng.normal() * scale, rng.normal() * scale, zc])
        else:
            u, v = rng.random(2)
            d = a + u * (b - a) + v * (c - a)
            if t % 4 == 1:
                d = d + np.spacing(d) * rng.integers(-3, 4, 3)
            elif t % 4 == 3:
                d = rng.normal(size=3) * scale
        pts = [a, b, c, d]
        cases.append("F " + " ".join(repr(float(x)) for q in pts for x in q))
        want.append(_exact_orient(pts))
    r = subprocess.run([exe], input="\n".join(cases) + "\n", capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    got = [int(v) for v in r.stdout.split()]
    assert got == want
    assert want.count(0) > 50 and 1 in want and -1 in want


@pytest.mark.parametrize("n,seed", [(10, 1), (300, 2), (20000, 3)])
def test_general_position_equals_qhull(n, seed):
    p = np.random.default_rng(seed).random((n, 3)).astype(np.float32)
    tets = F.delaunay(p[:, 0], p[:, 1], p[:, 2])
    ref = Delaunay(p.astype(np.float64)).simplices
    assert set(map(tuple, np.sort(tets, 1))) == set(map(tuple, np.sort(ref, 1)))


def _valid(p, tets):
    P = p.astype(np.float64)
    a, b, c, d = (P[tets[:, i]] for i in range(4))
    vol = np.einsum("ij,ij->i", b - a, np.cross(c - a, d - a)) / 6
    A = np.stack([b - a, c - a, d - a], 1)
    rhs = 0.5 * np.stack([np.sum((b - a) ** 2, 1), np.sum((c - a) ** 2, 1), np.sum((d - a) ** 2, 1)], 1)
    cen = np.linalg.solve(A, rhs[..., None])[..., 0] + a
    rad = np.linalg.norm(cen - a, axis=1)
    tree = cKDTree(P)
    inside = sum(len(tree.query_ball_point(cen[i], rad[i] * (1 - 1e-9))) for i in range(len(tets)))
    return vol.min(), vol.sum() / ConvexHull(P).volume, inside


def test_jittered_grid_equals_qhull_and_is_valid():
    g = np.stack(np.meshgrid(*[np.arange(9.0)] * 3, indexing="ij"), -1).reshape(-1, 3)
    j = (g + np.random.default_rng(4).uniform(-0.3, 0.3, g.shape)).astype(np.float32)
    tets = F.delaunay(j[:, 0], j[:, 1], j[:, 2])
    ref = Delaunay(j.astype(np.float64)).simplices
    assert set(map(tuple, np.sort(tets, 1))) == set(map(tuple, np.sort(ref, 1)))
    vmin, vfrac, inside = _valid(j, tets)
    assert vmin > 0 and abs(vfrac - 1) < 1e-12 and inside == 0


def test_exact_grid_is_a_valid_delaunay_triangulation():
    """Every unit cube is cospherical: the symbolic perturbation must still give non-flat,
    positively oriented tets that tile the hull, with empty circumspheres."""
    g = np.stack(np.meshgrid(np.arange(7.0), np.arange(6.0), np.arange(5.0), indexing="ij"), -1).reshape(-1, 3)
    g = g.astype(np.float32)
    tets = F.delaunay(g[:, 0], g[:, 1], g[:, 2])
    vmin, vfrac, inside = _valid(g, tets)
    assert vmin > 0 and abs(vfrac - 1) < 1e-12 and inside == 0
    # insertion-order independence: a permuted input gives the same tets (as point sets)
    perm = np.random.default_rng(0).permutation(len(g))
    t2 = F.delaunay(g[perm, 0], g[perm, 1], g[perm, 2])
    s1 = set(map(tuple, np.sort(tets, 1)))
    s2 = set(map(tuple, np.sort(perm[t2], 1)))
    assert s1 == s2


def test_duplicates_keep_the_last_index():
    p = np.random.default_rng(6).random((400, 3)).astype(np.float32)
    d = np.concatenate([p, p[:50], p[:50]])  # points 0..49 appear three times
    tets = F.delaunay(d[:, 0], d[:, 1], d[:, 2])
    used = np.unique(tets)
    assert not np.isin(np.arange(50), used).any() and not np.isin(np.arange(400, 450), used).any()
    assert np.isin(np.arange(450, 500), used).all()
    ref = F.delaunay(p[:, 0], p[:, 1], p[:, 2])
    remap = np.arange(len(p))
    remap[:50] = np.arange(450, 500)
    assert set(map(tuple, np.sort(remap[ref], 1))) == set(map(tuple, np.sort(tets, 1)))


def _shell(n, seed):
    r = np.random.default_rng(seed)
    d = r.normal(size=(n, 3))
    return (d / np.linalg.norm(d, axis=1)[:, None] * (1 + 0.1 * r.random((n, 1)))).astype(np.float32)


def _clusters(n, seed):
    r = np.random.default_rng(seed)
    c = r.random((12, 3))
    return (c[r.integers(0, 12, n)] + 0.02 * r.normal(size=(n, 3))).astype(np.float32)


@pytest.mark.parametrize("kind", ["uniform", "jittered", "exact_grid", "shell", "clusters", "duplicates"])
def test_parallel_build_equals_serial(kind):
    """The parallel build (levels of a randomized insertion order, each inserted by worker threads
    that own Morton chunks of the points and defer what crosses a chunk border) against the serial
    one: the same tets in the same order (the canonical order), for several thread counts. Sizes
    reach the parallel levels (over 8192 distinct points), so deferral, the shifted passes and the
    serial remainder all run; the exact grid is all cospherical (the symbolic perturbation decides
    every cube)."""
    r = np.random.default_rng(21)
    if kind == "uniform":
        p = r.random((30000, 3)).astype(np.float32)
    elif kind == "jittered":
        g = np.stack(np.meshgrid(*[np.arange(32.0)] * 3, indexing="ij"), -1).reshape(-1, 3)
        p = (g + r.uniform(-0.3, 0.3, g.shape)).astype(np.float32)
    elif kind == "exact_grid":
        p = np.stack(np.meshgrid(np.arange(26.0), np.arange(24.0), np.arange(22.0), indexing="ij"),
                     -1).reshape(-1, 3).astype(np.float32)
    elif kind == "shell":
        p = _shell(25000, 22)
    elif kind == "clusters":
        p = _clusters(25000, 23)
    else:
        q = r.random((20000, 3)).astype(np.float32)
        p = np.concatenate([q, q[::3], q[:1000]])[r.permutation(20000 + 6667 + 1000)]
    try:
        F.set_delaunay_threads(1)
        ref = F.delaunay(p[:, 0], p[:, 1], p[:, 2])
        for nt in (3, 8):
            F.set_delaunay_threads(nt)
            got = F.delaunay(p[:, 0], p[:, 1], p[:, 2])
            assert got.shape == ref.shape and np.array_equal(got, ref), (kind, nt)
    finally:
        F.set_delaunay_threads(0)
    if kind in ("uniform", "shell"):
        qh = Delaunay(p.astype(np.float64)).simplices
        assert set(map(tuple, np.sort(ref, 1))) == set(map(tuple, np.sort(qh, 1)))


def test_parallel_build_is_valid_on_a_jittered_cube():
    """A larger parallel build checked on its own: positively oriented tets tiling the hull with
    empty circumspheres, equal as a set to Qhull's."""
    g = np.stack(np.meshgrid(*[np.arange(40.0)] * 3, indexing="ij"), -1).reshape(-1, 3)
    p = (g + np.random.default_rng(24).uniform(-0.3, 0.3, g.shape)).astype(np.float32)
    try:
        F.set_delaunay_threads(6)
        tets = F.delaunay(p[:, 0], p[:, 1], p[:, 2])
    finally:
        F.set_delaunay_threads(0)
    ref = Delaunay(p.astype(np.float64)).simplices
    assert set(map(tuple, np.sort(tets, 1))) == set(map(tuple, np.sort(ref, 1)))
    vmin, vfrac, inside = _valid(p, tets)
    assert vmin > 0 and abs(vfrac - 1) < 1e-12 and inside == 0


def test_parallel_build_under_thread_sanitizer(tmp_path):
    """The parallel build on the host, compiled with ThreadSanitizer: identical to the serial build
    on jittered, uniform and exact-grid inputs, with no data race reported between the workers
    (they share the tet array without locks, under ownership rules)."""
    exe = str(tmp_path / "delaunay_par_test")
    subprocess.run(["g++", "-O1", "-g", "-std=c++17", "-ffp-contract=off", "-fsanitize=thread", "-pthread",
                    "-I", os.path.join(ROOT, "fomo_amd", "csrc"), "-o", exe,
                    os.path.join(ROOT, "tests", "cpp", "delaunay_par_test.cpp")], check=True)
    env = dict(os.environ, TSAN_OPTIONS="halt_on_error=1 exitcode=66")
    r = subprocess.run([exe], capture_output=True, text=True, timeout=600, env=env)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr[-4000:]
    assert "delaunay parallel ok" in r.stdout and "ThreadSanitizer" not in r.stderr


def test_delaunay_thread_setting():
    """The builder's thread count: set and read back, 0 restores the default, negative is refused."""
    try:
        F.set_delaunay_threads(5)
        assert F.lib().fomo_delaunay_threads() == 5
        with pytest.raises(F.FomoError):
            F.set_delaunay_threads(-1)
        assert F.lib().fomo_delaunay_threads() == 5
    finally:
        F.set_delaunay_threads(0)
    assert F.lib().fomo_delaunay_threads() == 0
