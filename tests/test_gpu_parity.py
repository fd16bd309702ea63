"""GPU parity: the HIP engine (through the C ABI) against the CPU oracle.

Bars: nearest-neighbour choices and view extents bit-exact; images within relative L2 1e-5 of
the FP64 oracle (the target tolerance); emissivity within float32 rounding of FP64.
"""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

C = 2.99792458e8
TOL_IMG = 1e-5


def _dataset(F, d):
    return F.Dataset(d["x"], d["y"], d["z"], [d["rho"], d["T"], d["vx"], d["vy"], d["vz"]])


def _oracle_render(d, tab, l, b, nx, ny, nz, nl, width, mode=1, imaging=False, dn=False):
    pk, fw = O.emission(mode, d["rho"], d["T"], tab, imaging=imaging)
    return O.render_nn(mode, d["x"], d["y"], d["z"], pk, fw, d["vx"], d["vy"], d["vz"], l, b, nx, ny, nz, nl, width,
                       tab["lambda0"], dn=dn, want_idx=True)


def test_c0_emission_matches_fp64(engine, c0_data, fe9_table):
    F = engine
    path, tab = fe9_table
    ds = _dataset(F, c0_data)
    T = F.Table.from_file(path)
    em = F.Emission.chianti(ds, T, tab["lambda0"], tab["atweight"])
    peak, fwhm = em.read()
    pk, fw = O.emission(1, c0_data["rho"], c0_data["T"], tab)
    # stored float32 of an FP64 computation: within 1 float ulp (plus libm last-bit slack)
    np.testing.assert_allclose(fwhm, fw, rtol=2e-7, atol=0)
    np.testing.assert_allclose(peak, pk, rtol=4e-7, atol=0)
    assert np.array_equal(peak == 0, pk == 0)
    em.close(), T.close(), ds.close()


def test_emission_at_half_bin_density_edges(engine, fe9_table):
    """The density row is the nearest one (fomo-CHIANTI.cpp:97), so a density whose index lies at a
    half-bin edge sits where G jumps by the difference of two rows. The reference's float pipeline
    takes the index from float-rounded logarithms (FoMo::log10 returns float), this engine from
    FP64 ones, as the FP64 oracle does. Densities within +-60 float ulps of four half-bin edges:
    the engine equals the FP64 oracle at every one; the float pipeline takes the other row at some
    of them (the recorded difference between the two)."""
    F = engine
    path, tab = fe9_table
    n0, dn = float(tab["logn"][0]), float(tab["logn"][1]) - float(tab["logn"][0])
    rho = []
    for k in (5, 12, 20, 33):
        c = np.float32(10.0 ** (n0 + (k + 0.5) * dn))
        up, dw = c, c
        rho.append(c)
        for _ in range(60):
            up, dw = np.nextafter(up, np.float32(np.inf)), np.nextafter(dw, np.float32(0))
            rho += [up, dw]
    rho = np.array(rho, np.float32)
    T = np.full(rho.size, np.float32(10.0 ** 5.9))
    n = rho.size
    xyz = np.random.default_rng(8).random((3, n)).astype(np.float32)
    zeros = np.zeros(n, np.float32)
    ds = F.Dataset(xyz[0], xyz[1], xyz[2], [rho, T, zeros, zeros, zeros])
    tb = F.Table.from_file(path)
    em = F.Emission.chianti(ds, tb, tab["lambda0"], tab["atweight"])
    try:
        peak, _ = em.read()
        pk64, _ = O.emission(1, rho, T, tab)
        pk32, _ = O.emission(0, rho, T, tab)
        assert (pk64 > 0).all()
        np.testing.assert_allclose(peak, pk64, rtol=4e-7, atol=0)
        other_row = np.abs(pk32 / pk64 - 1) > 1e-3
        print(f"half-bin edges: {int(other_row.sum())} of {n} densities take the other row in the float pipeline")
        assert other_row.any()
    finally:
        em.close(), tb.close(), ds.close()


def test_c0_full_render(engine, c0_data, fe9_table):
    """Config 0 end to end: 64^3 loop, view (0, pi/2), 64x64 image, nz 64, nl 30, 0.2 A."""
    F = engine
    path, tab = fe9_table
    d = c0_data
    width = 0.2 * C / tab["lambda0"]
    ds = _dataset(F, d)
    T = F.Table.from_file(path)
    em = F.Emission.chianti(ds, T, tab["lambda0"], tab["atweight"])
    l, b, nx, ny, nz, nl = 0.0, np.pi / 2, 64, 64, 64, 30
    img, meta = F.render_nn(em, l, b, nx, ny, nz, nl, width)
    idx, imeta = F.nn_indices(ds, l, b, nx, ny, nz)
    ref, ridx, rmeta, _ = _oracle_render(d, tab, l, b, nx, ny, nz, nl, width)
    assert np.array_equal(meta[:7], rmeta[:7])
    assert np.array_equal(idx, ridx), f"{(idx != ridx).sum()} NN choices differ"
    err = O.rel_l2(img, ref)
    print(f"C0 image rel L2 vs FP64 oracle: {err:.3e}")
    assert err <= TOL_IMG
    ref0, _, _, _ = _oracle_render(d, tab, l, b, nx, ny, nz, nl, width, mode=0)
    print(f"C0 image rel L2 vs reference-float oracle: {O.rel_l2(img, ref0):.3e}")
    em.close(), T.close(), ds.close()


@pytest.mark.parametrize("l,b", [(0.3, 1.1), (1.0, np.pi / 3), (2.5, 0.2), (0.0, 0.0), (np.pi / 4, np.pi / 4)])
def test_oblique_views(engine, fe9_table, l, b):
    F = engine
    path, tab = fe9_table
    d = F.gen_loop_cube(28, seed=11)
    width = 0.2 * C / tab["lambda0"]
    ds = _dataset(F, d)
    T = F.Table.from_file(path)
    em = F.Emission.chianti(ds, T, tab["lambda0"], tab["atweight"])
    nx, ny, nz, nl = 40, 36, 45, 30
    img, meta = F.render_nn(em, l, b, nx, ny, nz, nl, width)
    idx, _ = F.nn_indices(ds, l, b, nx, ny, nz)
    ref, ridx, rmeta, _ = _oracle_render(d, tab, l, b, nx, ny, nz, nl, width)
    assert np.array_equal(meta[:7], rmeta[:7])
    assert np.array_equal(idx, ridx), f"{(idx != ridx).sum()} NN choices differ"
    assert O.rel_l2(img, ref) <= TOL_IMG
    em.close(), T.close(), ds.close()


def test_jittered_points(engine, fe9_table):
    """Unstructured (jittered) points, as in the large-scale configs."""
    F = engine
    path, tab = fe9_table
    d = F.gen_loop_cube(30, seed=5, t_out=2.5e6, jitter=0.3, sausage=0.1)
    width = 0.2 * C / tab["lambda0"]
    ds = _dataset(F, d)
    T = F.Table.from_file(path)
    em = F.Emission.chianti(ds, T, tab["lambda0"], tab["atweight"])
    for (l, b) in [(0.0, np.pi / 2), (0.7, 0.9)]:
        nx, ny, nz, nl = 50, 50, 40, 20
        img, _ = F.render_nn(em, l, b, nx, ny, nz, nl, width)
        idx, _ = F.nn_indices(ds, l, b, nx, ny, nz)
        ref, ridx, _, _ = _oracle_render(d, tab, l, b, nx, ny, nz, nl, width)
        assert np.array_equal(idx, ridx)
        assert O.rel_l2(img, ref) <= TOL_IMG
    em.close(), T.close(), ds.close()


def test_imaging_nl1(engine, fe9_table):
    """lambda_pixel == 1 (Imaging): fwhm=1, intensity = sum of peaks (fomo-nearestneighbour.cpp:217-225)."""
    F = engine
    path, tab = fe9_table
    d = F.gen_loop_cube(24, seed=3)
    ds = _dataset(F, d)
    T = F.Table.from_file(path)
    em = F.Emission.chianti(ds, T, tab["lambda0"], tab["atweight"], imaging=True)
    l, b, nx, ny, nz = 0.5, 1.2, 30, 30, 30
    img, meta = F.render_nn(em, l, b, nx, ny, nz, 1, 1e5)
    ref, ridx, rmeta, _ = _oracle_render(d, tab, l, b, nx, ny, nz, 1, 1e5, imaging=True)
    assert O.rel_l2(img, ref) <= TOL_IMG
    em.close(), T.close(), ds.close()


def test_dn_scaling(engine, fe9_table):
    F = engine
    path, tab = fe9_table
    d = F.gen_loop_cube(20, seed=4)
    ds = _dataset(F, d)
    T = F.Table.from_file(path)
    em = F.Emission.chianti(ds, T, tab["lambda0"], tab["atweight"], imaging=True)
    img, meta = F.render_nn(em, 0.2, 0.4, 24, 20, 16, 1, 1e5, dn=True)
    ref, _, rmeta, _ = _oracle_render(d, tab, 0.2, 0.4, 24, 20, 16, 1, 1e5, imaging=True, dn=True)
    assert meta[9] == rmeta[9] and meta[9] < 1e-5
    assert O.rel_l2(img, ref) <= TOL_IMG
    em.close(), T.close(), ds.close()


def test_narrow_lines_direct_path(engine, fe9_table):
    """Lines much narrower than a lambda pixel take the direct-exp pass; results still match."""
    F = engine
    path, tab = fe9_table
    d = F.gen_loop_cube(16, seed=9)
    ds = _dataset(F, d)
    T = F.Table.from_file(path)
    heavy = dict(tab)
    heavy["atweight"] = 5.6e5  # thermal width / 100
    em = F.Emission.chianti(ds, T, tab["lambda0"], heavy["atweight"])
    width = 0.2 * C / tab["lambda0"]
    img, _ = F.render_nn(em, 0.3, 0.8, 20, 20, 20, 40, width)
    pk, fw = O.emission(1, d["rho"], d["T"], heavy)
    ref, _, _, _ = O.render_nn(1, d["x"], d["y"], d["z"], pk, fw, d["vx"], d["vy"], d["vz"], 0.3, 0.8, 20, 20, 20, 40,
                               width, tab["lambda0"])
    assert O.rel_l2(img, ref) <= TOL_IMG
    em.close(), T.close(), ds.close()


def test_fused_batch_direct_lines(engine, fe9_table):
    """A batch whose lines include one much narrower than a lambda pixel: the fused pass flags that
    line's rays and its DIRECT instantiation redoes them; every line equals its single-line render
    bit for bit, and the narrow line matches the FP64 oracle."""
    F = engine
    path, tab = fe9_table
    d = F.gen_loop_cube(16, seed=9)
    ds = _dataset(F, d)
    T = F.Table.from_file(path)
    atws = [tab["atweight"], 5.6e5, tab["atweight"]]  # the middle line: thermal width / 100
    ems = [F.Emission.chianti(ds, T, tab["lambda0"], a) for a in atws]
    width = 0.2 * C / tab["lambda0"]
    ls, bs = [0.3, 1.9], [0.8, np.pi / 2]
    bt = F.Batch(ems, ls, bs, 20, 20, 20, 40, width)
    bt.run()
    out, _ = bt.read()
    for v in range(2):
        for g, e in enumerate(ems):
            img, _ = F.render_nn(e, ls[v], bs[v], 20, 20, 20, 40, width)
            assert np.array_equal(out[v, g].ravel(), img), f"view {v} line {g}"
    heavy = dict(tab)
    heavy["atweight"] = atws[1]
    pk, fw = O.emission(1, d["rho"], d["T"], heavy)
    ref, _, _, _ = O.render_nn(1, d["x"], d["y"], d["z"], pk, fw, d["vx"], d["vy"], d["vz"], ls[0], bs[0], 20, 20, 20,
                               40, width, tab["lambda0"])
    assert O.rel_l2(out[0, 1].ravel(), ref) <= TOL_IMG
    bt.close()
    for o in ems + [T, ds]:
        o.close()


@pytest.mark.parametrize("nlines,nl", [(3, 60), (4, 7), (5, 33), (8, 64)])
def test_fused_batch_lines_equal_plain(engine, fe9_table, tmp_path, nlines, nl):
    """Batches of 3-8 lines run the fused NN + integration pass (2, 4 or 8 lanes per ray); lines
    with different lambda0 and atomic weight must each equal their single-line render bit for bit."""
    F = engine
    path, tab = fe9_table
    lam = [171.0, 195.1, 211.3, 193.5, 284.2, 335.4, 94.0, 131.2][:nlines]
    atw = [56.0, 56.0, 28.0, 56.0, 56.0, 56.0, 40.0, 56.0][:nlines]
    d = F.gen_loop_cube(20, seed=4, jitter=0.2)
    ds = _dataset(F, d)
    tabs = [F.Table.from_file(F.gen_goft_table(str(tmp_path / f"l{i}.dat"), lambda0=lam[i], atweight=atw[i],
                                               peak_logt=5.8 + 0.15 * i)) for i in range(nlines)]
    ems = [F.Emission.chianti(ds, t, lam[i], atw[i]) for i, t in enumerate(tabs)]
    ls, bs = [0.3, 2.1], [1.1, np.pi / 2]
    nx, ny, nz = 26, 22, 33
    width = 0.25 * C / 171.0
    bt = F.Batch(ems, ls, bs, nx, ny, nz, nl, width)
    bt.run()
    out, meta = bt.read()
    for v in range(len(ls)):
        for g, e in enumerate(ems):
            img, m = F.render_nn(e, ls[v], bs[v], nx, ny, nz, nl, width)
            assert np.count_nonzero(img) > 0
            assert np.array_equal(out[v, g].ravel(), img), f"view {v} line {g}"
            assert np.array_equal(meta[v], m)
    bt.close()
    for o in ems + tabs + [ds]:
        o.close()


@pytest.mark.parametrize("nlines,nl", [(2, 60), (3, 17), (4, 60), (4, 64), (4, 5)])
def test_shared_profile_batch_equals_plain(engine, fe9_table, tmp_path, nlines, nl):
    """Lines with one lambda0 and atomic weight (G(T) tables peaked at different temperatures, as
    in a multi-table batch of one line) share their Gaussian profile: the fused pass evaluates it
    once per sample for all lines, one stride-4 chain per lane. Each line must equal its
    single-line render bit for bit, and match the FP64 oracle."""
    F = engine
    path, tab = fe9_table
    d = F.gen_loop_cube(20, seed=12, jitter=0.2)
    ds = _dataset(F, d)
    peaks = [5.8, 6.2, 6.4, 6.9][:nlines]
    tabs = [F.Table.from_file(F.gen_goft_table(str(tmp_path / f"s{i}.dat"), peak_logt=p)) for i, p in enumerate(peaks)]
    ems = [F.Emission.chianti(ds, t, tab["lambda0"], tab["atweight"]) for t in tabs]
    ls, bs = [0.3, 2.1, 1.0], [1.1, np.pi / 2, np.pi / 3]
    nx, ny, nz = 26, 22, 35
    width = 0.2 * C / tab["lambda0"]
    bt = F.Batch(ems, ls, bs, nx, ny, nz, nl, width)
    bt.run()
    out, meta = bt.read()
    for v in range(len(ls)):
        for g, e in enumerate(ems):
            img, m = F.render_nn(e, ls[v], bs[v], nx, ny, nz, nl, width)
            assert np.count_nonzero(img) > 0
            assert np.array_equal(out[v, g].ravel(), img), f"view {v} line {g}"
            assert np.array_equal(meta[v], m)
    last = O.read_table(str(tmp_path / f"s{nlines - 1}.dat"))
    pk, fw = O.emission(1, d["rho"], d["T"], last)
    ref, _, _, _ = O.render_nn(1, d["x"], d["y"], d["z"], pk, fw, d["vx"], d["vy"], d["vz"], ls[0], bs[0], nx, ny, nz,
                               nl, width, tab["lambda0"])
    assert O.rel_l2(out[0, nlines - 1].ravel(), ref) <= TOL_IMG
    bt.close()
    for o in ems + tabs + [ds]:
        o.close()


def test_shared_profile_direct_pass(engine, fe9_table, tmp_path):
    """A shared profile much narrower than a lambda pixel: every flagged (ray, line) is redone by
    the shared-profile DIRECT instantiation; lines still equal their single-line renders bit for
    bit and match the FP64 oracle."""
    F = engine
    path, tab = fe9_table
    d = F.gen_loop_cube(16, seed=9)
    ds = _dataset(F, d)
    tabs = [F.Table.from_file(path), F.Table.from_file(F.gen_goft_table(str(tmp_path / "n.dat"), peak_logt=6.3)),
            F.Table.from_file(path)]
    heavy = 5.6e5  # thermal width / 100
    ems = [F.Emission.chianti(ds, t, tab["lambda0"], heavy) for t in tabs]
    width = 0.2 * C / tab["lambda0"]
    ls, bs = [0.3, 1.9], [0.8, np.pi / 2]
    bt = F.Batch(ems, ls, bs, 20, 20, 20, 40, width)
    bt.run()
    out, _ = bt.read()
    # the direct pass's list of flagged waves is cleared by every replay of the graph: replays
    # give the same images
    for _ in range(3):
        bt.run()
        again, _ = bt.read()
        assert np.array_equal(again, out)
    for v in range(2):
        for g, e in enumerate(ems):
            img, _ = F.render_nn(e, ls[v], bs[v], 20, 20, 20, 40, width)
            assert np.array_equal(out[v, g].ravel(), img), f"view {v} line {g}"
    assert np.array_equal(out[:, 0], out[:, 2])
    hv = dict(tab)
    hv["atweight"] = heavy
    pk, fw = O.emission(1, d["rho"], d["T"], hv)
    ref, _, _, _ = O.render_nn(1, d["x"], d["y"], d["z"], pk, fw, d["vx"], d["vy"], d["vz"], ls[1], bs[1], 20, 20, 20,
                               40, width, tab["lambda0"])
    assert O.rel_l2(out[1, 0].ravel(), ref) <= TOL_IMG
    bt.close()
    for o in ems + tabs + [ds]:
        o.close()


@pytest.mark.parametrize("atw,nl", [(56.0, 60), (56.0, 7), (5.6e5, 40), (56.0, 100), (5.6e5, 77)])
def test_single_line_fused_view(engine, fe9_table, tmp_path, atw, nl):
    """A single-line view of 256 x 256 rays (the fused pass's threshold) runs the fused pass with
    one line: the 4 lanes of a ray run the profile's 4 chains, of one lambda pass, or of two in one
    launch for 65 <= nl <= 128. It must equal the same line rendered in a batch with a second line
    (one line per lane, the shared profile, or, beyond 64 pixels, k_nn + two k_los_spectral passes)
    bit for bit, and match the FP64 oracle; atomic weight 5.6e5 (thermal width / 100) takes the
    direct-exp pass."""
    F = engine
    path, tab = fe9_table
    d = F.gen_loop_cube(14, seed=21, jitter=0.2)
    ds = _dataset(F, d)
    T = F.Table.from_file(path)
    T2 = F.Table.from_file(F.gen_goft_table(str(tmp_path / "o.dat"), lambda0=193.5, atweight=56.0, peak_logt=6.2))
    em = F.Emission.chianti(ds, T, tab["lambda0"], atw)
    em_same = F.Emission.chianti(ds, T, tab["lambda0"], atw)          # shared profile with em
    em_other = F.Emission.chianti(ds, T2, 193.5, 56.0)                 # a different profile
    width = 0.2 * C / tab["lambda0"]
    l, b, nx, ny, nz = 0.6, 1.0, 256, 256, 12
    img, meta = F.render_nn(em, l, b, nx, ny, nz, nl, width)
    for pair in ([em, em_same], [em, em_other]):
        bt = F.Batch(pair, [l], [b], nx, ny, nz, nl, width)
        bt.run()
        out, bmeta = bt.read()
        assert np.array_equal(out[0, 0].ravel(), img)
        assert np.array_equal(bmeta[0], meta)
        bt.close()
    hv = dict(tab)
    hv["atweight"] = atw
    pk, fw = O.emission(1, d["rho"], d["T"], hv)
    ref, _, rmeta, _ = O.render_nn(1, d["x"], d["y"], d["z"], pk, fw, d["vx"], d["vy"], d["vz"], l, b, nx, ny, nz, nl,
                                   width, tab["lambda0"])
    assert np.array_equal(meta[:7], rmeta[:7])
    assert np.count_nonzero(ref) > 0
    assert O.rel_l2(img, ref) <= TOL_IMG
    for o in (em, em_same, em_other, T, T2, ds):
        o.close()


@pytest.mark.parametrize("nl", [1, 100])
def test_unfused_batch_equals_plain(engine, fe9_table, tmp_path, nl):
    """Batches the fused pass does not take (imaging, or more than 64 lambda pixels: two lambda
    passes) run k_nn + the integration kernels on two streams; they too equal single renders."""
    F = engine
    path, tab = fe9_table
    d = F.gen_loop_cube(20, seed=6)
    ds = _dataset(F, d)
    tabs = [F.Table.from_file(path), F.Table.from_file(F.gen_goft_table(str(tmp_path / "b.dat"), peak_logt=6.3))]
    ems = [F.Emission.chianti(ds, t, tab["lambda0"], tab["atweight"], imaging=(nl == 1)) for t in tabs]
    ls, bs = [0.2, 1.4, 2.9], [np.pi / 2, 0.7, 1.2]
    nx, ny, nz = 24, 20, 31
    width = 0.2 * C / tab["lambda0"]
    bt = F.Batch(ems, ls, bs, nx, ny, nz, nl, width)
    bt.run()
    out, meta = bt.read()
    for v in range(len(ls)):
        for g, e in enumerate(ems):
            img, m = F.render_nn(e, ls[v], bs[v], nx, ny, nz, nl, width)
            assert np.count_nonzero(img) > 0
            assert np.array_equal(out[v, g].ravel(), img), f"view {v} line {g}"
            assert np.array_equal(meta[v], m)
    bt.close()
    for o in ems + tabs + [ds]:
        o.close()


def test_batch_equals_plain(engine, fe9_table, tmp_path):
    """A batched graph (views x lines) gives bit-identical images to one render per view and line."""
    F = engine
    path, tab = fe9_table
    p2 = F.gen_goft_table(str(tmp_path / "l2.dat"), peak_logt=6.2)
    d = F.gen_loop_cube(24, seed=2)
    ds = _dataset(F, d)
    T1 = F.Table.from_file(path)
    T2 = F.Table.from_file(p2)
    e1 = F.Emission.chianti(ds, T1, tab["lambda0"], tab["atweight"])
    e2 = F.Emission.chianti(ds, T2, tab["lambda0"], tab["atweight"])
    ls = [0.0, 0.4, 1.3]
    bs = [np.pi / 2, np.pi / 3, 0.5]
    nx, ny, nz, nl = 32, 28, 30, 60
    width = 0.2 * C / tab["lambda0"]
    bt = F.Batch([e1, e2], ls, bs, nx, ny, nz, nl, width)
    bt.run()
    out, meta = bt.read()
    for v in range(3):
        for g, e in enumerate([e1, e2]):
            img, m = F.render_nn(e, ls[v], bs[v], nx, ny, nz, nl, width)
            assert np.array_equal(out[v, g].ravel(), img)
            assert np.array_equal(meta[v], m)
    bt.run()  # replay gives the same bits
    out2, _ = bt.read()
    assert np.array_equal(out, out2)
    bt.close()
    for o in (e1, e2, T1, T2, ds):
        o.close()


def test_thomson_small(engine):
    F = engine
    d = F.gen_corona(40, seed=3)
    ds = F.Dataset(d["x"], d["y"], d["z"], [d["rho"]])
    em = F.Emission.thomson(ds)
    tot, pol = em.read()
    rt, rp = O.thomson_emission(1, d["x"], d["y"], d["z"], d["rho"])
    np.testing.assert_allclose(tot, rt, rtol=4e-7)
    np.testing.assert_allclose(pol, rp, rtol=4e-7)
    for (l, b) in [(0.0, np.pi / 2), (0.6, 0.05), (1.0, 0.5)]:
        nx, ny, nz = 48, 40, 36
        img, meta = F.render_thomson(em, l, b, nx, ny, nz)
        idx, _ = F.nn_indices(ds, l, b, nx, ny, nz, thomson=True)
        ref, ridx, rmeta, _ = O.render_thomson(1, d["x"], d["y"], d["z"], rt, rp, l, b, nx, ny, nz, want_idx=True)
        assert np.array_equal(meta[:7], rmeta[:7])
        assert np.array_equal(idx, ridx), f"{(idx != ridx).sum()} NN choices differ"
        assert O.rel_l2(img, ref) <= TOL_IMG
    em.close(), ds.close()


@pytest.mark.parametrize("jitter", [0.0, 0.3])
def test_lookup_fast_path_matches_ring_search(engine, jitter):
    """The candidate lookup grid changes speed only: NN choices equal the ring search and the oracle.
    The structured cube (jitter 0) would be indexed by its axes; it is indexed as scattered points
    here, so that its lookup grid is tested too."""
    F = engine
    d = F.gen_loop_cube(36, seed=21, jitter=jitter)
    ds = _dataset(F, d)
    ds.set_rectilinear(False)
    assert ds.rect_axes() is None
    info = ds.lookup_info()
    assert info["subdiv"] == 2 and info["cells"] > 0
    assert info["fallback_cells"] < 0.05 * info["cells"], info
    assert info["candidates"] < 8 * info["cells"], info
    views = [(0.0, np.pi / 2), (0.37, 1.2), (1.9, 0.4), (np.pi / 4, np.pi / 3)]
    fast = [F.nn_indices(ds, l, b, 48, 44, 50)[0] for (l, b) in views]
    ds.set_lookup(0)
    assert ds.lookup_info()["subdiv"] == 0
    ring = [F.nn_indices(ds, l, b, 48, 44, 50)[0] for (l, b) in views]
    for (l, b), a, r in zip(views, fast, ring):
        assert np.array_equal(a, r)
        md = O.render_nn(1, d["x"], d["y"], d["z"], np.ones(d["x"].size), np.ones(d["x"].size), d["vx"], d["vy"],
                         d["vz"], l, b, 48, 44, 50, 1, 1e5, 171.0, want_idx=True)[1]
        assert np.array_equal(a, md)
    ds.set_lookup(1)
    assert np.array_equal(F.nn_indices(ds, *views[1], 48, 44, 50)[0], fast[1])
    ds.close()


def test_thomson_solar_disk_samples(engine):
    """Samples inside the solar disk hole and outside the shell resolve through the exact BVH path."""
    F = engine
    d = F.gen_corona(72, seed=8)
    ds = F.Dataset(d["x"], d["y"], d["z"], [d["rho"]])
    em = F.Emission.thomson(ds)
    rt, rp = O.thomson_emission(1, d["x"], d["y"], d["z"], d["rho"])
    for (l, b) in [(0.0, np.pi / 2), (0.9, 0.3)]:
        nx, ny, nz = 64, 60, 48
        img, meta = F.render_thomson(em, l, b, nx, ny, nz)
        idx, _ = F.nn_indices(ds, l, b, nx, ny, nz, thomson=True)
        ref, ridx, rmeta, _ = O.render_thomson(1, d["x"], d["y"], d["z"], rt, rp, l, b, nx, ny, nz, want_idx=True)
        assert np.array_equal(idx, ridx), f"{(idx != ridx).sum()} NN choices differ"
        assert O.rel_l2(img, ref) <= TOL_IMG
    ds.set_lookup(0)  # BVH only
    idx0, _ = F.nn_indices(ds, 0.9, 0.3, 64, 60, 48, thomson=True)
    assert np.array_equal(idx0, ridx)
    em.close(), ds.close()


@pytest.mark.parametrize("budget", [1, 24, 0])
def test_deep_range_walk_matches_oracle(engine, fe9_table, budget):
    """An exact query whose wave walk reaches the far budget moves to the range walk (the tree's
    node ranges walked in parallel from the running best, then merged by distance and original
    index); once its queue is full the other waves finish their own walks. NN choices stay the
    oracle's in all three regimes: every wave queued or finishing in place (budget 1), some
    queued (24), none (0). The lookup grid is off, so every sample takes the exact query."""
    F = engine
    d = F.gen_corona(72, seed=8)
    ds = F.Dataset(d["x"], d["y"], d["z"], [d["rho"]])
    ds.set_lookup(0)
    ds.set_far_budget(budget)
    rt, rp = O.thomson_emission(1, d["x"], d["y"], d["z"], d["rho"])
    nx, ny, nz = 64, 60, 48
    for (l, b) in [(0.0, np.pi / 2), (0.9, 0.3)]:
        idx, meta = F.nn_indices(ds, l, b, nx, ny, nz, thomson=True)
        queued, cap = int(meta[10]), int(meta[11])
        print(f"budget {budget} view ({l}, {b}): {queued} waves reached the budget, queue holds {cap}", flush=True)
        if budget == 0:
            assert queued == 0
        elif budget == 1:
            assert queued > cap  # both the range walk and the walks finished in place
        else:
            assert queued > 0
        ridx = O.render_thomson(1, d["x"], d["y"], d["z"], rt, rp, l, b, nx, ny, nz, want_idx=True)[1]
        assert np.array_equal(idx, ridx), f"{(idx != ridx).sum()} NN choices differ"
    ds.close()
    # an imaging NN render takes k_nn + k_nn_far (not the fused pass): the merge writes the
    # per-sample records, so the image must not depend on the budget
    path, tab = fe9_table
    c = F.gen_loop_cube(24, seed=4, jitter=0.3)
    ds = _dataset(F, c)
    ds.set_lookup(0)
    T = F.Table.from_file(path)
    em = F.Emission.chianti(ds, T, tab["lambda0"], tab["atweight"], imaging=True)
    l, b, nx, ny, nz = 0.7, 0.9, 40, 36, 44
    ds.set_far_budget(0)
    ref_img, _ = F.render_nn(em, l, b, nx, ny, nz, 1, 1.0)
    ds.set_far_budget(budget)
    img, _ = F.render_nn(em, l, b, nx, ny, nz, 1, 1.0)
    assert np.array_equal(img, ref_img)
    oref, _, _, _ = _oracle_render(c, tab, l, b, nx, ny, nz, 1, 1.0, imaging=True)
    assert O.rel_l2(img, oref) <= TOL_IMG
    em.close(), T.close(), ds.close()


def test_far_group_lists_match_exact_query(engine):
    """Lookup cells far from the data (beyond the shell, inside the disk hole) take their group's
    candidate list; NN choices stay those of the exact octree query and of the oracle."""
    F = engine
    d = F.gen_corona(56, seed=5)
    ds = F.Dataset(d["x"], d["y"], d["z"], [d["rho"]])
    info = ds.lookup_info()
    # without group lists every cell more than 8 cell-list cells from a point (the box corners
    # beyond r = 3 and the middle of the hole) is left to the exact query
    assert info["fallback_cells"] < 0.05 * info["cells"], info
    # the group paths this test covers are built: far groups whose cells all read the group's
    # header (no per-cell record), and outer groups for targets beyond the lookup grid
    groups = ds.lookup_groups()
    print("lookup groups:", groups, flush=True)
    assert groups["far_lists"] > 0 and groups["full_groups"] > 0, groups
    assert groups["outer_groups"] > 0 and groups["outer_lists"] > 0, groups
    views = [(0.0, np.pi / 2), (0.9, 0.3), (2.2, 0.0), (np.pi / 4, np.pi / 4)]
    # the rotated views place LOS samples beyond the lookup grid (the data box plus 6 cell-list
    # cells on each side), where the outer groups answer
    n_, dims, _ = ds.info()
    lo = np.array([d["x"].min(), d["y"].min(), d["z"].min()], np.float64)
    hi = np.array([d["x"].max(), d["y"].max(), d["z"].max()], np.float64)
    margin = 7.0 * (hi - lo) / (np.array(dims) - 1)
    ext = O.extents(d["x"], d["y"], d["z"], np.pi / 4, np.pi / 4)
    p = np.stack([a.ravel() for a in O.view_samples(ext, np.pi / 4, np.pi / 4, 56, 52, 60)], 1)
    beyond = np.any((p < lo - margin) | (p > hi + margin), 1)
    assert beyond.sum() > 1000, beyond.sum()
    fast = [F.nn_indices(ds, l, b, 56, 52, 60, thomson=True)[0] for (l, b) in views]
    ds.set_lookup(0)
    exact = [F.nn_indices(ds, l, b, 56, 52, 60, thomson=True)[0] for (l, b) in views]
    for (l, b), a, e in zip(views, fast, exact):
        assert np.array_equal(a, e), f"view {(l, b)}: {(a != e).sum()} NN choices differ"
    rt, rp = O.thomson_emission(1, d["x"], d["y"], d["z"], d["rho"])
    for k in (1, 3):
        ridx = O.render_thomson(1, d["x"], d["y"], d["z"], rt, rp, *views[k], 56, 52, 60, want_idx=True)[1]
        assert np.array_equal(fast[k], ridx)
    ds.close()


def test_bilinear_goft_matches_scipy(engine, fe9_table):
    """The bilinear G(T, n) extension against scipy's RegularGridInterpolator (linear in log T and
    log n on the table's grid, FP64): the peak ratio bilinear / reference equals G_bil / G_ref per
    point (the normalisation, density and width factors cancel), with G_ref the reference's
    linear-in-T, nearest-row value computed here in numpy. Points span the table, with some outside
    it in T and n (zero in both modes), and the imaging branch is checked too."""
    from scipy.interpolate import RegularGridInterpolator
    F = engine
    path, tab = fe9_table
    lt = tab["logt"].astype(np.float64)
    ln = tab["logn"].astype(np.float64)
    G = tab["g"].astype(np.float64).reshape(tab["nrho"], tab["nt"])
    r = np.random.default_rng(12)
    n = 20000
    # temperatures near the table's G peak (no float underflow of the peaks), a tenth outside the
    # table's temperature range on either side
    lT = r.uniform(5.4, 6.4, n)
    side = r.random(n)
    lT = np.where(side < 0.05, r.uniform(lt[0] - 0.3, lt[0], n), np.where(side > 0.95, r.uniform(lt[-1], lt[-1] + 0.3, n), lT))
    lN = r.uniform(ln[0] - 0.5, ln[-1] + 0.5, n)
    rho = (10.0 ** lN).astype(np.float32)
    T = (10.0 ** lT).astype(np.float32)
    lr, lT2 = np.log10(rho.astype(np.float64)), np.log10(T.astype(np.float64))
    dt, dn = lt[1] - lt[0], ln[1] - ln[0]
    ft = np.floor((lT2 - lt[0]) / dt)
    fn = np.floor((lr - ln[0]) / dn)
    rn = np.round((lr - ln[0]) / dn)
    in_t = (ft >= 0) & (ft + 1 <= tab["nt"] - 1)
    bil_ok = in_t & (fn >= 0) & (fn + 1 <= tab["nrho"] - 1)
    ref_ok = in_t & (rn >= 0) & (rn <= tab["nrho"] - 1)
    gb = np.zeros(n)
    gb[bil_ok] = RegularGridInterpolator((ln, lt), G)(np.stack([lr[bil_ok], lT2[bil_ok]], 1))
    gr = np.zeros(n)
    i = np.flatnonzero(ref_ok)
    ri, t0 = rn[i].astype(int), ft[i].astype(int)
    gr[i] = (lT2[i] - lt[t0 + 1]) * (G[ri, t0] - G[ri, t0 + 1]) / (lt[t0] - lt[t0 + 1]) + G[ri, t0 + 1]
    xyz = r.random((3, n)).astype(np.float32)
    zeros = np.zeros(n, np.float32)
    ds = F.Dataset(xyz[0], xyz[1], xyz[2], [rho, T, zeros, zeros, zeros])
    tb = F.Table.from_file(path)
    try:
        for imaging in (False, True):
            e0 = F.Emission.chianti(ds, tb, tab["lambda0"], tab["atweight"], imaging=imaging)
            e1 = F.Emission.chianti(ds, tb, tab["lambda0"], tab["atweight"], imaging=imaging, bilinear=True)
            p0, w0 = e0.read()
            p1, w1 = e1.read()
            e0.close(), e1.close()
            assert np.array_equal(w0, w1)
            assert np.array_equal(p1 != 0, gb != 0) and np.array_equal(p0 != 0, gr != 0)
            both = (gb != 0) & (gr != 0)
            assert both.sum() > n // 3
            np.testing.assert_allclose(p1[both].astype(np.float64) / p0[both], gb[both] / gr[both], rtol=1e-6)
            assert not np.array_equal(p0, p1)
    finally:
        tb.close(), ds.close()
