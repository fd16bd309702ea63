"""GPU tests of the projection method (the reference's "Projection", fomo-projection.cpp:17-214).

Bars: the pixel of every point equal to the oracle's (integer work, bit-exact); given the same
per-point emission (the device's, read back), the imaging image (nl == 1, float sums of peaks)
bit-identical to the oracle's point-order float sums, and the spectroscopic image within
relative L2 1e-6 of them (the device takes each term's Gaussian from a ratio recurrence rather
than one exp per term, so a term may differ in its last bits); and
from the inputs alone, within relative L2 1e-5 of the FP64 oracle (FP64 emission, double sums).
"""
import numpy as np
import pytest

import fomo_amd as F
from oracle import oracle as O

pytestmark = pytest.mark.gpu

C = 2.99792458e8
TOL_FLOAT = 1e-6
TOL_IMG = 1e-5


@pytest.fixture(scope="module")
def cube(engine, fe9_table):
    path, tab = fe9_table
    d = F.gen_loop_cube(24, seed=17, jitter=0.3, sausage=0.1)
    ds = F.Dataset(d["x"], d["y"], d["z"], [d["rho"], d["T"], d["vx"], d["vy"], d["vz"]])
    T = F.Table.from_file(path)
    em = F.Emission.chianti(ds, T, tab["lambda0"], tab["atweight"])
    emi = F.Emission.chianti(ds, T, tab["lambda0"], tab["atweight"], imaging=True)
    # mode 0: the device's own per-point emission (float); mode 1: the FP64 oracle's
    e = {0: (em.read(), emi.read()),
         1: (O.emission(1, d["rho"], d["T"], tab), O.emission(1, d["rho"], d["T"], tab, imaging=True))}
    yield {"d": d, "tab": tab, "ds": ds, "em": em, "emi": emi, "e": e}
    em.close(), emi.close(), T.close(), ds.close()


def _oracle(c, mode, l, b, nx, ny, nz, nl, width, dn=False, want_pix=False):
    d = c["d"]
    (pk, fw), (pki, fwi) = c["e"][mode]
    if nl == 1:
        pk, fw = pki, fwi
    return O.render_projection(mode, d["x"], d["y"], d["z"], pk, fw, d["vx"], d["vy"], d["vz"], l, b, nx, ny, nz,
                               nl, width, c["tab"]["lambda0"], dn=dn, want_pix=want_pix)


VIEWS = [(0.4, 1.1), (0.0, np.pi / 2), (2.3, 0.35), (1.0, 0.0), (np.pi / 4, np.pi / 4), (5.5, 2.9)]


@pytest.mark.parametrize("l,b", VIEWS)
@pytest.mark.parametrize("nx,ny", [(31, 27), (8, 12)])
def test_pixels_match_oracle(cube, l, b, nx, ny):
    pix, meta = F.projection_pixels(cube["ds"], l, b, nx, ny, 16)
    _, rpix, rmeta = _oracle(cube, 0, l, b, nx, ny, 16, 1, 0.0, want_pix=True)
    assert np.array_equal(meta[:6], rmeta[:6])
    assert pix.min() >= 0 and pix.max() < nx * ny
    assert np.array_equal(pix, rpix), f"{(pix != rpix).sum()} of {pix.size} points differ"


@pytest.mark.parametrize("l,b,nx,ny,dn", [(0.4, 1.1, 40, 36, False), (2.3, 0.35, 9, 7, False),
                                          (1.2, 0.6, 64, 64, True), (0.0, np.pi / 2, 24, 20, False)])
def test_imaging_is_bit_identical_to_point_order_sums(cube, l, b, nx, ny, dn):
    img, meta = F.render_projection(cube["emi"], l, b, nx, ny, 30, 1, 0.0, dn=dn)
    ref, _, rmeta = _oracle(cube, 0, l, b, nx, ny, 30, 1, 0.0, dn=dn)
    ref1, _, _ = _oracle(cube, 1, l, b, nx, ny, 30, 1, 0.0, dn=dn)
    np.testing.assert_allclose(meta[7:10], rmeta[7:10], rtol=1e-15)
    assert np.count_nonzero(ref) > 0
    same = (img.astype(np.float64) == ref).mean()
    e1 = O.rel_l2(img, ref1)
    print(f"projection imaging {nx}x{ny} dn={dn}: {same:.4f} of pixels bit-identical, rel L2 {e1:.2e} vs FP64")
    assert np.array_equal(img.astype(np.float64), ref)
    assert e1 <= TOL_IMG


@pytest.mark.parametrize("l,b,nx,ny,nl", [(0.4, 1.1, 32, 28, 30), (2.3, 0.35, 10, 9, 61), (1.0, 0.0, 20, 24, 7)])
def test_spectral_matches_oracle(cube, l, b, nx, ny, nl):
    width = 0.2 * C / cube["tab"]["lambda0"]
    img, meta = F.render_projection(cube["em"], l, b, nx, ny, 30, nl, width)
    ref0, _, rmeta = _oracle(cube, 0, l, b, nx, ny, 30, nl, width)
    ref1, _, _ = _oracle(cube, 1, l, b, nx, ny, 30, nl, width)
    assert np.array_equal(meta[:6], rmeta[:6])
    e0, e1 = O.rel_l2(img, ref0), O.rel_l2(img, ref1)
    same = (img.astype(np.float64) == ref0).mean()
    print(f"projection nl={nl}: rel L2 {e0:.2e} vs float oracle ({same:.4f} bit-identical), {e1:.2e} vs FP64")
    assert np.count_nonzero(ref0) > 0
    assert e0 <= TOL_FLOAT
    assert e1 <= TOL_IMG


def test_degenerate_extent_matches_oracle(engine, fe9_table):
    # every point on the plane x = 2: at view (0, 0) the x' extent is 0 and all points fall in column 0
    path, tab = fe9_table
    rng = np.random.default_rng(5)
    n = 2000
    x = np.full(n, 2.0, np.float32)
    y = rng.uniform(-3, 3, n).astype(np.float32)
    z = rng.uniform(-3, 3, n).astype(np.float32)
    rho = np.full(n, 1e9, np.float32)
    T = rng.uniform(0.8e6, 2.0e6, n).astype(np.float32)
    ds = F.Dataset(x, y, z, [rho, T])
    Tb = F.Table.from_file(path)
    emi = F.Emission.chianti(ds, Tb, tab["lambda0"], tab["atweight"], imaging=True)
    try:
        img, _ = F.render_projection(emi, 0.0, 0.0, 6, 5, 4, 1, 0.0)
        pk, fw = emi.read()
        zero = np.zeros(n, np.float32)
        ref, _, _ = O.render_projection(0, x, y, z, pk, fw, zero, zero, zero, 0.0, 0.0, 6, 5, 4, 1, 0.0,
                                        tab["lambda0"])
        assert np.count_nonzero(img.reshape(5, 6)[:, 1:]) == 0
        assert np.array_equal(img.astype(np.float64), ref)
    finally:
        emi.close(), Tb.close(), ds.close()


def test_projection_server_equals_sync(cube):
    width = 0.2 * C / cube["tab"]["lambda0"]
    reqs = [(cube["em"], F.METHOD_PROJECTION, 0.4, 1.1, 32, 30, 40, 30, width),
            (cube["emi"], F.METHOD_PROJECTION, 1.7, 0.5, 24, 24, 24, 1, 0.0),
            (cube["em"], F.METHOD_NN, 0.4, 1.1, 32, 30, 40, 30, width),
            (cube["em"], F.METHOD_PROJECTION, 2.2, 0.8, 20, 28, 33, 12, width)]
    sync = []
    for r in reqs:
        fn = F.render_projection if r[1] == F.METHOD_PROJECTION else F.render_nn
        sync.append(fn(r[0], *r[2:8], r[8]))
    sv = F.Server(device_bytes=1 << 30, host_bytes=256 << 20, max_inflight=4, nstreams=2)
    try:
        ts = [sv.submit(r[0], r[1], *r[2:8], lambda_width=r[8]) for r in reqs]
        for t, (img, meta) in zip(ts, sync):
            simg, smeta = sv.result(t)
            assert np.array_equal(simg, img)
            assert np.array_equal(smeta, meta)
    finally:
        sv.close()
