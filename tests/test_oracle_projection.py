"""Known-answer tests for the oracle's projection method (fomo-projection.cpp:17-214).

Inputs are built by hand so that pixels and spectra follow from the geometry and closed forms:
grid points seen along an axis land on known columns, rounding ties go away from zero
(std::round), imaging sums are float sums in point order, and a single moving point gives a
Gaussian at its Doppler shift.
"""
import numpy as np
import pytest

from oracle import oracle as O

C = 2.99792458e8
LN2 = np.log(2.0)


def _grid(m, L=10.0):
    a = (-L / 2 + L * np.arange(m) / (m - 1)).astype(np.float32)
    zz, yy, xx = np.meshgrid(a, a, a, indexing="ij")
    return xx.ravel(), yy.ravel(), zz.ravel()


def test_axis_view_pixels_are_grid_columns():
    # view (0, 0): x' = x, y' = y, so point (a, b, c) of an m^3 grid lands on pixel (b, a)
    m = 9
    x, y, z = _grid(m)
    n = x.size
    one = np.ones(n)
    zero = np.zeros(n, np.float32)
    _, pix, meta = O.render_projection(0, x, y, z, one, one, zero, zero, zero, 0.0, 0.0, m, m, m, 1, 1e5, 171.0,
                                       want_pix=True)
    iz, iy, jx = np.meshgrid(np.arange(m), np.arange(m), np.arange(m), indexing="ij")
    assert np.array_equal(pix, (iy * m + jx).ravel())
    # twice the resolution: every grid column falls on an even pixel
    _, pix2, _ = O.render_projection(0, x, y, z, one, one, zero, zero, zero, 0.0, 0.0, 2 * m - 1, 2 * m - 1, m, 1,
                                     1e5, 171.0, want_pix=True)
    assert np.array_equal(pix2, (2 * iy * (2 * m - 1) + 2 * jx).ravel())
    assert np.allclose(meta[:6], [-5, 5, -5, 5, -5, 5])
    assert meta[7] == 10.0 / (m - 1)


def test_rounding_ties_go_away_from_zero():
    # view (0, 0), x extent [0, 4] on 5 pixels: (x - 0) * 4 / 4 = x, ties at .5 round up (std::round);
    # round-half-to-even would give 0, 2, 2 for 0.5, 1.5, 2.5
    x = np.array([0.0, 0.5, 1.5, 2.5, 4.0], np.float32)
    y = np.array([0.0, 4.0, 0.0, 0.0, 0.0], np.float32)
    z = np.zeros(5, np.float32)
    one = np.ones(5)
    zero = np.zeros(5, np.float32)
    _, pix, _ = O.render_projection(0, x, y, z, one, one, zero, zero, zero, 0.0, 0.0, 5, 5, 2, 1, 1e5, 171.0,
                                    want_pix=True)
    assert pix.tolist() == [0, 4 * 5 + 1, 2, 3, 4]


def test_imaging_is_the_float_sum_in_point_order():
    m = 7
    x, y, z = _grid(m)
    n = x.size
    rng = np.random.default_rng(3)
    peak = rng.lognormal(0.0, 2.0, n).astype(np.float32)  # wide dynamic range: order matters in float
    zero = np.zeros(n, np.float32)
    img0, _, meta = O.render_projection(0, x, y, z, peak.astype(np.float64), np.ones(n), zero, zero, zero, 0.0, 0.0,
                                        m, m, m, 1, 1e5, 171.0)
    img1, _, _ = O.render_projection(1, x, y, z, peak.astype(np.float64), np.ones(n), zero, zero, zero, 0.0, 0.0,
                                     m, m, m, 1, 1e5, 171.0)
    # points of one pixel differ only in z, which is the slowest index: point order = z order
    pk = peak.reshape(m, m, m)
    acc = np.zeros((m, m), np.float32)
    for k in range(m):
        acc = (acc + pk[k]).astype(np.float32)
    sc = (10.0 / (m - 1)) * 1e8
    expect = (sc * acc.astype(np.float64)).astype(np.float32).astype(np.float64)
    assert np.array_equal(img0, expect.ravel())
    np.testing.assert_allclose(img1, sc * pk.astype(np.float64).sum(axis=0).ravel(), rtol=1e-13)
    assert meta[8] == sc


@pytest.mark.parametrize("mode,rtol", [(1, 1e-12), (0, 2e-6)])
def test_single_moving_point_gives_its_doppler_profile(mode, rtol):
    # view (0, pi/2): LOS unit vector (1, -0, cos(pi/2)), so v_los = vx for vz = 0.
    # Point A at the origin moves; point B at (4, 4, 4) has zero peak. x' = -z: A has the larger
    # x' (column nx-1); y' = y: A has the smaller y' (row 0).
    x = np.array([0.0, 4.0], np.float32)
    y = np.array([0.0, 4.0], np.float32)
    z = np.array([0.0, 4.0], np.float32)
    peak = np.array([2.0e-3, 0.0])
    w = np.array([0.03, 1.0])
    v = 2.0e4
    vx = np.array([v, 0.0], np.float32)
    zero = np.zeros(2, np.float32)
    nx, ny, nz, nl = 6, 5, 3, 201
    lam0 = 171.0
    width = 0.3 * C / lam0
    img, _, meta = O.render_projection(mode, x, y, z, peak, w, vx, zero, zero, 0.0, np.pi / 2, nx, ny, nz, nl, width,
                                       lam0)
    cube = img.reshape(ny, nx, nl)
    lit = np.argwhere(np.abs(cube).sum(axis=2) > 0)
    assert lit.tolist() == [[0, nx - 1]]
    wl = 0.3
    lam = np.arange(nl) / (nl - 1) * wl - wl / 2
    shift = v / C * lam0
    pw = float(np.float32(w[0])) if mode == 0 else w[0]
    pp = float(np.float32(peak[0])) if mode == 0 else peak[0]
    sc = meta[8]
    expect = sc * pp * np.exp(-4 * LN2 * (lam - shift) ** 2 / pw ** 2)
    np.testing.assert_allclose(cube[0, nx - 1], expect, rtol=rtol, atol=expect.max() * 1e-12)
    # the line centre moves to +v/c lambda0 (a red shift for a receding point)
    assert abs(lam[np.argmax(cube[0, nx - 1])] - shift) <= wl / (nl - 1)
    # path length from the z' extent: z' = x sin(b) cos(l) + z cos(b) -> [0, 4] over nz - 1 steps
    assert meta[7] == pytest.approx(4.0 / (nz - 1), rel=1e-15)


def test_flat_slab_and_dn_scaling():
    # all points at z = 0 seen face-on: z' extent is 0, so the path length is 1 Mm (:166-170)
    m = 6
    a = (np.arange(m) * 2.0).astype(np.float32)
    yy, xx = np.meshgrid(a, a, indexing="ij")
    x, y = xx.ravel(), yy.ravel()
    z = np.zeros_like(x)
    n = x.size
    zero = np.zeros(n, np.float32)
    peak = np.full(n, 3.0)
    img, _, meta = O.render_projection(0, x, y, z, peak, np.ones(n), zero, zero, zero, 0.0, 0.0, m, m, 4, 1, 1e5,
                                       171.0, dn=True)
    assert meta[7] == 1.0
    dx = np.float32(10.0 / (m - 1))
    apix = (float(dx) / 0.715) * (float(dx) / 0.715) * (np.pi / 180 / 3600) ** 2
    assert meta[9] == pytest.approx(apix, rel=1e-15)
    np.testing.assert_allclose(img, np.float32(1e8 * apix * 3.0), rtol=1e-7)


def test_degenerate_x_extent_puts_points_in_column_zero():
    # at view (0, 0) x' = x = 1 for every point: the x' extent is 0 and all points go to column 0
    y = np.array([0.0, 1.0, 2.0, 3.0], np.float32)
    x = np.ones(4, np.float32)
    z = np.array([0.0, 1.0, 0.0, 1.0], np.float32)
    one = np.ones(4)
    zero = np.zeros(4, np.float32)
    _, pix, _ = O.render_projection(0, x, y, z, one, one, zero, zero, zero, 0.0, 0.0, 4, 4, 2, 1, 1e5, 171.0,
                                    want_pix=True)
    assert pix.tolist() == [0, 4, 8, 12]
