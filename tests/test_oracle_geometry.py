"""Independent checks of the oracle's view geometry (rows A7, A9, A10, A12).

The oracle restates the reference's rotation (fomo-nearestneighbour.cpp:83-85), LOS unit vector
(:87-90), LOS sample de-rotation (:165; fomo-thomson.cpp:295; fomo-CGAL.cpp:187-192), search
half-widths (fomo-nearestneighbour.cpp:138; fomo-thomson.cpp:249-260) and Thomson's blended
rotation and sin^2 chi (fomo-thomson.cpp:136-199). The GPU is compared with the oracle bit for
bit, so a slip shared by both would pass every GPU test. These checks pin the oracle instead with
what the mathematics fixes, built without its formulas:

- the rotation is the product of two elementary rotations, R_y(-b) R_z(l), taken from
  scipy.spatial.transform (so it preserves lengths);
- the LOS samples are the image-frame sample grid rotated back by the transpose of that matrix,
  and rotating them forward returns the grid;
- the LOS unit vector is the image z axis in data coordinates, and a Doppler-shifted spectrum at an
  oblique view is centred on the rotated velocity's z component;
- a data point placed exactly at a sample position is that sample's nearest point (distance 0);
- Thomson's rotation is the matrix blend (1-s) F + s P with s a logistic function (scipy.special),
  F the same elementary product with eps-clamped cos b / sin b, P the rotation about z; sin^2 chi
  is 0 on the line of sight through the origin and 1 perpendicular to it at b = 0;
- both search half-widths, including the sigmoid blend, at its midpoint and at its limits.

test_checks_catch_planted_mistakes rebuilds the oracle from edited copies of its source (a sign
flipped in a sin(l) term, in the LOS unit vector, in the blend, ...) and shows each edit fails at
least one of these checks.
"""
import concurrent.futures
import math
import os
import subprocess

import numpy as np
import pytest
from scipy.special import expit
from scipy.spatial.transform import Rotation

from oracle import oracle as O

C = 2.99792458e8
LN2 = math.log(2.0)
ANGLES = [(0.0, 0.0), (0.0, math.pi / 2), (0.7, 1.1), (2.3, -0.4), (-1.2, 2.6), (math.pi, math.pi / 3),
          (0.31, 1.309), (5.9, -1.45)]


def rz(t):
    return Rotation.from_rotvec([0.0, 0.0, t]).as_matrix()


def ry(t):
    return Rotation.from_rotvec([0.0, t, 0.0]).as_matrix()


def view_matrix(l, b):
    """Rotation by l about z, then by -b about y."""
    return ry(-b) @ rz(l)


def cloud(n, seed, scale=5.0):
    r = np.random.default_rng(seed)
    return [(scale * r.uniform(-1, 1, n)).astype(np.float32) for _ in range(3)]


def close(a, b, tol, what):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    err = np.max(np.abs(a - b)) if a.size else 0.0
    assert err <= tol, f"{what}: max |diff| {err:.3e} > {tol:.1e}"


# ---- checks (each raises AssertionError on failure; the mutation test reuses them) ----------------


def check_rotation_is_elementary_product():
    x, y, z = cloud(500, 1)
    pts = np.stack([x, y, z]).astype(np.float64)
    for l, b in ANGLES:
        xr, yr, zr = O.rotate(x, y, z, l, b)
        want = view_matrix(l, b) @ pts
        close(np.stack([xr, yr, zr]), want, 1e-13, f"rotation at l={l}, b={b}")
        close(np.sqrt(xr ** 2 + yr ** 2 + zr ** 2), np.linalg.norm(pts, axis=0), 1e-13, "|r'| = |r|")


def check_extents_are_rotated_bounds():
    x, y, z = cloud(800, 2)
    pts = np.stack([x, y, z]).astype(np.float64)
    for l, b in ANGLES:
        r = view_matrix(l, b) @ pts
        want = [r[0].min(), r[0].max(), r[1].min(), r[1].max(), r[2].min(), r[2].max()]
        close(O.extents(x, y, z, l, b), want, 1e-13, f"extents at l={l}, b={b}")


def image_grid(ext, nx, ny, nz, cgal=False):
    jj = np.arange(nx)[None, :, None]
    ii = np.arange(ny)[:, None, None]
    kk = np.arange(nz)[None, None, :]
    xs = ext[0] + (ext[1] - ext[0]) * jj / (nx - 1)
    ys = ext[2] + (ext[3] - ext[2]) * ii / (ny - 1)
    if nz == 1:
        zs = np.full(kk.shape, float(ext[4]))
    else:
        zs = ext[4] + (ext[5] - ext[4]) * kk / (nz - 1)
    shape = (ny, nx, nz)
    return [np.broadcast_to(a, shape).astype(np.float64) for a in (xs, ys, zs)]


def check_samples_rotate_back_to_image_grid():
    x, y, z = cloud(300, 3)
    for l, b in ANGLES:
        ext = O.extents(x, y, z, l, b)
        for cgal in (False, True):
            for nx, ny, nz in ((7, 5, 6), (4, 3, 1)):
                p = O.view_samples(ext, l, b, nx, ny, nz, cgal=cgal)
                g = image_grid(ext, nx, ny, nz, cgal)
                mt = view_matrix(l, b).T
                want = np.einsum("ij,jk->ik", mt, np.stack([a.ravel() for a in g]))
                close(np.stack([a.ravel() for a in p]), want, 1e-13, f"samples at l={l}, b={b}, cgal={cgal}")
                back = O.rotate_d(*(a.ravel() for a in p), l, b)
                close(np.stack(back), np.stack([a.ravel() for a in g]), 1e-13, "rotated samples = image grid")


def check_los_velocity_is_rotated_z():
    r = np.random.default_rng(4)
    v = [r.normal(0, 3e4, 200).astype(np.float32) for _ in range(3)]
    vv = np.stack(v).astype(np.float64)
    for l, b in ANGLES:
        want = (view_matrix(l, b) @ vv)[2]
        close(O.los_velocity(*v, l, b), want, 1e-9, f"v_los at l={l}, b={b}")  # m/s, |v| ~ 1e5


def check_point_at_sample_is_its_nearest():
    base = cloud(400, 5)
    for l, b in ((0.7, 1.1), (2.3, -0.4), (-1.2, 2.6)):
        nx, ny, nz = 9, 8, 10
        ext = O.extents(*base, l, b)
        p = O.view_samples(ext, l, b, nx, ny, nz)
        r = np.random.default_rng(6)
        pick = [(int(r.integers(1, ny - 1)), int(r.integers(1, nx - 1)), int(r.integers(1, nz - 1)))
                for _ in range(25)]
        pick = sorted(set(pick))
        extra = [np.array([p[a][q] for q in pick], np.float32) for a in range(3)]
        x, y, z = [np.concatenate([base[a], extra[a]]) for a in range(3)]
        assert np.array_equal(O.extents(x, y, z, l, b), ext), "interior points moved the extents"
        one = np.ones(x.size)
        zero = np.zeros(x.size, np.float32)
        _, idx, _, _ = O.render_nn(1, x, y, z, one, one, zero, zero, zero, l, b, nx, ny, nz, 1, 1e5, 171.0,
                                   want_idx=True)
        idx = idx.reshape(ny, nx, nz)
        got = [int(idx[q]) for q in pick]
        assert got == list(range(base[0].size, x.size)), f"samples at l={l}, b={b} missed the point placed on them"


def check_doppler_centroid_at_oblique_views():
    m, nl = 6, 601
    a = np.linspace(-5, 5, m).astype(np.float32)
    zz, yy, xx = np.meshgrid(a, a, a, indexing="ij")
    x, y, z = xx.ravel(), yy.ravel(), zz.ravel()
    n = x.size
    v = np.array([2.1e4, -1.3e4, 0.9e4], np.float32)
    peak, w = 1e-3, 0.02
    width = 0.5 * C / 171.0
    lam = np.arange(nl) / (nl - 1) * 0.5 - 0.25
    for l, b in ((0.7, 1.1), (2.3, -0.4), (-1.2, 2.6)):
        img, _, _, _ = O.render_nn(1, x, y, z, np.full(n, peak), np.full(n, w), np.full(n, v[0]), np.full(n, v[1]),
                                   np.full(n, v[2]), l, b, m, m, m, nl, width, 171.0)
        spec = img.reshape(m * m, nl)
        lit = spec.sum(1) > 0
        assert lit.sum() >= m * m // 2
        cen = (spec[lit] * lam).sum(1) / spec[lit].sum(1)
        want = (view_matrix(l, b) @ v.astype(np.float64))[2] / C * 171.0
        close(cen, np.full(cen.shape, want), 1e-9, f"Doppler centroid at l={l}, b={b}")


def thomson_matrices(l, b):
    eps = 1e-6
    cb, sb = math.cos(b), math.sin(b)
    cbe = eps if abs(cb) < eps else cb
    sbe = eps if abs(sb) < eps else sb
    tilt = np.array([[cbe, 0.0, -sbe], [0.0, 1.0, 0.0], [sbe, 0.0, cbe]])
    if cbe == cb and sbe == sb:
        assert np.allclose(tilt, ry(-b), atol=1e-15)
    full = tilt @ rz(l)
    s = expit(20.0 * (abs(b) - 1.309))
    blend = (1 - s) * full + s * rz(l)
    # The LOS direction in the data frame, F^T e_z. The reference dots it with the blended,
    # rotated position (fomo-thomson.cpp:193-198), so these checks do the same.
    unit = full.T @ np.array([0.0, 0.0, 1.0])
    return blend, unit


def check_thomson_rotation_and_sin2chi():
    x, y, z = cloud(500, 7, scale=2000.0)
    pts = np.stack([x, y, z]).astype(np.float64)
    for l, b in ANGLES + [(0.4, 0.0), (1.0, math.pi / 2)]:
        xa, ya, za, s2 = O.rotate(x, y, z, l, b, thomson=True)
        bm, u = thomson_matrices(l, b)
        r = bm @ pts
        close(np.stack([xa, ya, za]), r, 1e-9, f"Thomson rotation at l={l}, b={b}")  # Mm, |r| ~ 3e3
        want = 1 - (u @ r) ** 2 / np.sum(r * r, 0)
        close(s2, want, 1e-12, f"sin^2 chi at l={l}, b={b}")


def check_thomson_sin2chi_limits():
    for l in (0.0, 0.9, -2.2):
        # b = 0: the line of sight is the z axis; points on it scatter nothing, points in the
        # perpendicular plane give sin^2 chi = 1 (to the 1e-6 clamp of sin b, squared).
        on = [np.zeros(5, np.float32), np.zeros(5, np.float32), np.linspace(700, 2000, 5).astype(np.float32)]
        _, _, _, s_on = O.rotate(*on, l, 0.0, thomson=True)
        assert np.all(np.abs(s_on) < 1e-10), f"sin^2 chi on the LOS at l={l}: {s_on}"
        t = np.linspace(0, 2 * np.pi, 9)[:-1]
        perp = [(1500 * np.cos(t)).astype(np.float32), (1500 * np.sin(t)).astype(np.float32), np.zeros(8, np.float32)]
        _, _, _, s_perp = O.rotate(*perp, l, 0.0, thomson=True)
        assert np.all(np.abs(s_perp - 1) < 1e-10), f"sin^2 chi perpendicular at l={l}: {s_perp}"
    # b = pi/2, l = 0: the blend is 0.9947 of the pole rotation; points on the x axis (the LOS)
    # keep only the small remainder of the full rotation.
    _, _, _, s = O.rotate(np.array([900.0, -1700.0], np.float32), np.zeros(2, np.float32), np.zeros(2, np.float32),
                          0.0, math.pi / 2, thomson=True)
    assert np.all(s < 1e-4) and np.all(s >= 0)


def check_maxdistance():
    for ext, nx, ny in (([-3.0, 5.0, -1.0, 2.0, 0, 1], 17, 9), ([-10.0, 10.0, -10.0, 10.0, -2, 2], 64, 64),
                        ([0.0, 1.0, 0.0, 7.0, 0, 1], 5, 33)):
        dx = (ext[1] - ext[0]) / (nx - 1)
        dy = (ext[3] - ext[2]) / (ny - 1)
        close(O.maxdistance(ext, nx, ny), max(dx, dy) / 0.3, 1e-14 * max(dx, dy), "NN maxdistance")
        eq = max(0.5 * math.hypot(dx, dy), 0.1 * dx)
        away = max(dx, dy) / 0.005
        for b in (0.0, 0.1745, -0.1745, 0.6, math.pi / 2, -math.pi / 2, 3.0):
            s = expit(10.0 * (abs(b) - 0.1745))
            got = O.maxdistance(ext, nx, ny, b, thomson=True)
            close(got, (1 - s) * eq + s * away, 1e-12 * away, f"Thomson maxdistance at b={b}")
        close(O.maxdistance(ext, nx, ny, 0.1745, thomson=True), 0.5 * (eq + away), 1e-12 * away, "sigmoid midpoint")
        assert abs(O.maxdistance(ext, nx, ny, math.pi / 2, thomson=True) / away - 1) < 1e-5


def check_thomson_render_geometry():
    """Extents, search half-width, NN choices and the pB sum of a Thomson render, from the checked
    parts: the blended extents from the matrix blend, the samples from the transposed rotation,
    the choices from the brute-force box query, pB = scale * sum max(pol * sin^2 chi, 0)."""
    x, y, z = cloud(600, 8, scale=2000.0)
    pts = np.stack([x, y, z]).astype(np.float64)
    r = np.random.default_rng(9)
    tot = r.uniform(0.5, 2.0, x.size)
    pol = r.uniform(0.1, 1.0, x.size)
    nx, ny, nz = 12, 10, 14
    for l, b in ((0.7, 1.1), (2.3, 0.2), (0.0, math.pi / 2)):
        img, idx, meta, _ = O.render_thomson(1, x, y, z, tot, pol, l, b, nx, ny, nz, want_idx=True)
        bm, u = thomson_matrices(l, b)
        rr = bm @ pts
        ext = [rr[0].min(), rr[0].max(), rr[1].min(), rr[1].max(), rr[2].min(), rr[2].max()]
        close(meta[:6], ext, 1e-9, f"Thomson extents at l={l}, b={b}")
        md = O.maxdistance(meta[:6], nx, ny, b, thomson=True)
        assert meta[6] == md
        p = O.view_samples(meta[:6], l, b, nx, ny, nz)
        want_p = view_matrix(l, b).T @ np.stack([a.ravel() for a in image_grid(meta[:6], nx, ny, nz)])
        close(np.stack([a.ravel() for a in p]), want_p, 1e-9, "Thomson samples")
        bf = O.nn_query(x, y, z, *(a.ravel() for a in p), md, method="brute")
        assert np.array_equal(idx, bf)
        s2 = 1 - (u @ rr) ** 2 / np.sum(rr * rr, 0)
        contrib = np.where(idx >= 0, np.maximum(pol[np.maximum(idx, 0)] * s2[np.maximum(idx, 0)], 0.0), 0.0)
        want_img = meta[8] * contrib.reshape(ny * nx, nz).sum(1)
        np.testing.assert_allclose(img, want_img, rtol=1e-12, atol=1e-12 * np.abs(want_img).max())


CHECKS = [check_rotation_is_elementary_product, check_extents_are_rotated_bounds,
          check_samples_rotate_back_to_image_grid, check_los_velocity_is_rotated_z,
          check_point_at_sample_is_its_nearest, check_doppler_centroid_at_oblique_views,
          check_thomson_rotation_and_sin2chi, check_thomson_sin2chi_limits, check_maxdistance,
          check_thomson_render_geometry]


@pytest.mark.parametrize("check", CHECKS, ids=[c.__name__[6:] for c in CHECKS])
def test_geometry(check):
    check()


# Planted mistakes: (exact source text, replacement). Each text occurs once in the oracle source.
# (The 0.1 * dx term of Thomson's equator formula is never the larger one, since
# 0.5 * hypot(dx, dy) >= 0.5 * dx, so a change there cannot show in any output.)
MUTANTS = [
    ("*yr = g0 * sin(l) + g1 * cos(l);", "*yr = -g0 * sin(l) + g1 * cos(l);"),
    ("- g1 * cos(b) * sin(l) - g2 * sin(b);", "+ g1 * cos(b) * sin(l) - g2 * sin(b);"),
    ("*zr = g0 * sin(b) * cos(l) - g1 * sin(b) * sin(l)", "*zr = g0 * sin(b) * cos(l) + g1 * sin(b) * sin(l)"),
    ("u[1] = -sin(b) * sin(l);", "u[1] = sin(b) * sin(l);"),
    ("*p0 = x * cos(b) * cos(l) + y * sin(l)", "*p0 = x * cos(b) * cos(l) - y * sin(l)"),
    ("*p1 = -x * cos(b) * sin(l)", "*p1 = x * cos(b) * sin(l)"),
    ("+ y * cos(l) - z * sin(b) * sin(l);", "+ y * cos(l) + z * sin(b) * sin(l);"),
    ("*p2 = -x * sin(b) + z * cos(b);", "*p2 = x * sin(b) + z * cos(b);"),
    ("double full_y = g0 * sin_l + g1 * cos_l;", "double full_y = -g0 * sin_l + g1 * cos_l;"),
    ("double full_x = g0 * cos_b * cos_l - g1 * cos_b * sin_l", "double full_x = g0 * cos_b * cos_l + g1 * cos_b * sin_l"),
    ("double pole_x = g0 * cos_l - g1 * sin_l;", "double pole_x = g0 * cos_l + g1 * sin_l;"),
    ("double unit[3] = {sin_b * cos(l), -sin_b * sin(l), cos_b};", "double unit[3] = {sin_b * cos(l), sin_b * sin(l), cos_b};"),
    ("exp(-steepness * (std::fabs(b) - threshold))", "exp(steepness * (std::fabs(b) - threshold))"),
    ("if (std::abs(sin_b) < epsilon) sin_b = epsilon;", ""),
    ("exp(-10 * (std::fabs(b) - 0.1745))", "exp(10 * (std::fabs(b) - 0.1745))"),
    ("std::max(0.5 * std::sqrt(", "std::max(0.25 * std::sqrt("),
    ("return std::max((e[1] - e[0]) / (nx - 1), (e[3] - e[2]) / (ny - 1)) / .3;",
     "return std::min((e[1] - e[0]) / (nx - 1), (e[3] - e[2]) / (ny - 1)) / .3;"),
    ("*z = nz > 1 ? double(k) / (nz - 1) * (e[5] - e[4]) + e[4] : e[4];",
     "*z = nz > 1 ? double(k) / nz * (e[5] - e[4]) + e[4] : e[4];"),
    ("if (nz != 1) deltaz /= (nz - 1);\n    *z", "if (nz != 1) deltaz /= nz;\n    *z"),
]


def _build_mutant(src, k, old, new, outdir):
    assert src.count(old) == 1, f"mutant {k}: the text to edit occurs {src.count(old)} times"
    cpp = os.path.join(outdir, f"mutant{k}.cpp")
    so = os.path.join(outdir, f"liboracle_mutant{k}.so")
    with open(cpp, "w") as f:
        f.write(src.replace(old, new))
    subprocess.check_call(["g++", "-O1", "-fopenmp", "-ffp-contract=off", "-fPIC", "-std=c++17", "-shared", "-o", so,
                           cpp])
    return so


def test_checks_catch_planted_mistakes(tmp_path):
    src = open(os.path.join(os.path.dirname(O.__file__), "fomo_oracle.cpp")).read()
    with concurrent.futures.ThreadPoolExecutor(8) as ex:
        libs = list(ex.map(lambda a: _build_mutant(src, a[0], a[1][0], a[1][1], str(tmp_path)),
                           enumerate(MUTANTS)))
    saved = O._LIB
    survivors = []
    try:
        for k, so in enumerate(libs):
            O._LIB = O.load(so)
            caught = []
            for check in CHECKS:
                try:
                    check()
                except AssertionError:
                    caught.append(check.__name__)
            if not caught:
                survivors.append(MUTANTS[k])
    finally:
        O._LIB = saved
    assert not survivors, f"planted mistakes no check catches: {survivors}"
