"""GPU tests of the render server (continuous batching over resident data).

Bars: a request's image and geometry are bit-identical whether it runs alone, in a mix of
concurrently admitted requests of other shapes and methods, or through the synchronous
renderer (the same kernels on private buffers; no cross-request reductions). One request of the
mix is also checked against the FP64 oracle at the target tolerance (rel L2 1e-5).
Admission never exceeds the memory budget or the in-flight limit.
"""
import numpy as np
import pytest

import fomo_amd as F
from oracle import oracle as O

pytestmark = pytest.mark.gpu

C = 2.99792458e8
TOL_IMG = 1e-5


@pytest.fixture(scope="module")
def scene(engine, fe9_table):
    path, tab = fe9_table
    d = F.gen_loop_cube(40, seed=5, jitter=0.2)
    ds = F.Dataset(d["x"], d["y"], d["z"], [d["rho"], d["T"], d["vx"], d["vy"], d["vz"]])
    T = F.Table.from_file(path)
    em = F.Emission.chianti(ds, T, tab["lambda0"], tab["atweight"])
    emi = F.Emission.chianti(ds, T, tab["lambda0"], tab["atweight"], imaging=True)
    c = F.gen_corona(36, seed=4)
    dsc = F.Dataset(c["x"], c["y"], c["z"], [c["rho"]])
    emt = F.Emission.thomson(dsc)
    yield {"d": d, "tab": tab, "ds": ds, "T": T, "em": em, "emi": emi, "c": c, "dsc": dsc, "emt": emt}
    for k in ("em", "emi", "emt", "T", "ds", "dsc"):
        {"em": em, "emi": emi, "emt": emt, "T": T, "ds": ds, "dsc": dsc}[k].close()


def _requests(s):
    w = 0.2 * C / s["tab"]["lambda0"]
    M, TH = F.METHOD_NN, F.METHOD_THOMSON
    return [
        (s["em"], M, 0.0, np.pi / 2, 48, 40, 44, 30, w, False),
        (s["em"], M, 0.7, 1.1, 33, 29, 20, 12, w, False),
        (s["emi"], M, 0.2, 0.4, 64, 64, 64, 1, 0.0, False),
        (s["emt"], TH, 0.0, np.pi / 2, 40, 40, 36, 1, 0.0, False),
        (s["em"], M, 2.1, 0.3, 20, 24, 96, 60, w, True),
        (s["emt"], TH, 0.9, 0.3, 52, 44, 30, 1, 0.0, False),
        (s["em"], M, 1.3, np.pi / 3, 80, 72, 32, 4, w, False),
        (s["emi"], M, 3.0, 1.4, 16, 16, 16, 1, 0.0, True),
    ]


def _sync_render(r):
    em, method, l, b, nx, ny, nz, nl, w, dn = r
    if method == F.METHOD_THOMSON:
        return F.render_thomson(em, l, b, nx, ny, nz)
    return F.render_nn(em, l, b, nx, ny, nz, nl, w, dn=dn)


def test_alone_equals_mixed_equals_sync(scene):
    reqs = _requests(scene)
    sync = [_sync_render(r) for r in reqs]
    sv = F.Server(device_bytes=512 << 20, host_bytes=64 << 20, max_inflight=6, nstreams=4)
    try:
        alone = []
        for r in reqs:
            t = sv.submit(*r[:8], lambda_width=r[8], dn=r[9])
            alone.append(sv.result(t, timeout=60))
        # the mix: every request three times, interleaved, all submitted before any completes
        order = [i for _ in range(3) for i in range(len(reqs))]
        tickets = [sv.submit(*reqs[i][:8], lambda_width=reqs[i][8], dn=reqs[i][9]) for i in order]
        mixed = [sv.result(t, timeout=120) for t in tickets]
        st = sv.stats()
    finally:
        sv.close()
    for i, r in enumerate(reqs):
        assert np.array_equal(alone[i][0], sync[i][0]), f"request {i}: server alone != synchronous render"
        assert np.array_equal(alone[i][1], sync[i][1])
        assert np.count_nonzero(alone[i][0]) > 0
    for k, i in enumerate(order):
        assert np.array_equal(mixed[k][0], alone[i][0]), f"request {i} (copy {k // len(reqs)}) differs in the mix"
        assert np.array_equal(mixed[k][1], alone[i][1])
    assert st["completed"] == len(reqs) + len(order) and st["failed"] == 0
    assert 1 < st["peak_running"] <= 6, st
    assert st["device_peak"] <= st["device_budget"]
    assert st["device_in_use"] == 0 and st["host_in_use"] == 0


def test_server_result_matches_oracle(scene):
    d, tab = scene["d"], scene["tab"]
    w = 0.2 * C / tab["lambda0"]
    l, b, nx, ny, nz, nl = 0.4, 1.2, 36, 30, 40, 30
    sv = F.Server(device_bytes=256 << 20, host_bytes=16 << 20, max_inflight=4, nstreams=2)
    try:
        t = sv.submit(scene["em"], F.METHOD_NN, l, b, nx, ny, nz, nl, lambda_width=w)
        img, meta = sv.result(t, timeout=60)
        tt = sv.times(t)
    finally:
        sv.close()
    pk, fw = O.emission(1, d["rho"], d["T"], tab)
    ref, _, rmeta, _ = O.render_nn(1, d["x"], d["y"], d["z"], pk, fw, d["vx"], d["vy"], d["vz"], l, b, nx, ny, nz,
                                   nl, w, tab["lambda0"])
    assert np.array_equal(meta[:7], rmeta[:7])
    assert O.rel_l2(img, ref) <= TOL_IMG
    assert 0 <= tt[0] <= tt[1] <= tt[2]


def test_admission_by_footprint(scene):
    """A budget that holds one large request at a time: the large ones serialise, small ones pass
    them (bounded), and every result is still exact."""
    em = scene["em"]
    w = 0.2 * C / scene["tab"]["lambda0"]
    big = (128, 128, 64, 30)
    small = (32, 32, 32, 30)
    fb, hb = F.Server.footprint(em, F.METHOD_NN, *big)
    fs, hs = F.Server.footprint(em, F.METHOD_NN, *small)
    assert fb > 8 * fs
    budget = fb + 2 * fs + (1 << 20)
    sv = F.Server(device_bytes=budget, host_bytes=8 << 20, max_inflight=8, nstreams=4)
    try:
        with pytest.raises(F.FomoError):  # larger than the whole budget: rejected at submit
            sv.submit(em, F.METHOD_NN, 0.1, 0.2, 512, 512, 128, 30, lambda_width=w)
        shapes = [big, small, big, small, small, big, small]
        tickets = [sv.submit(em, F.METHOD_NN, 0.3 * k, 1.0, *s[:3], s[3], lambda_width=w)
                   for k, s in enumerate(shapes)]
        res = [sv.result(t, timeout=120) for t in tickets]
        st = sv.stats()
    finally:
        sv.close()
    assert st["device_peak"] <= budget and st["failed"] == 0
    for k, s in enumerate(shapes):
        ref, _ = F.render_nn(em, 0.3 * k, 1.0, *s[:3], s[3], w)
        assert np.array_equal(res[k][0], ref)


def test_objects_in_use_cannot_be_destroyed(engine, fe9_table):
    path, tab = fe9_table
    d = F.gen_loop_cube(24, seed=2)
    ds = F.Dataset(d["x"], d["y"], d["z"], [d["rho"], d["T"], d["vx"], d["vy"], d["vz"]])
    T = F.Table.from_file(path)
    em = F.Emission.chianti(ds, T, tab["lambda0"], tab["atweight"])
    sv = F.Server(device_bytes=64 << 20, host_bytes=4 << 20, max_inflight=2, nstreams=1)
    lib = F.lib()
    try:
        tickets = [sv.submit(em, F.METHOD_NN, 0.1 * k, 0.5, 24, 24, 24, 8, lambda_width=5e4) for k in range(12)]
        # queued requests hold the dataset and the cube (12 requests through 2 slots take far
        # longer than these two calls)
        assert lib.fomo_goft_destroy(em.id) == -5
        assert lib.fomo_dataset_reindex(ds.id, 0.0) == -5
        for t in tickets:
            sv.result(t, timeout=60)
        assert lib.fomo_goft_destroy(em.id) == 0  # nothing in flight any more
        em.id = None
    finally:
        sv.close()
        T.close(), ds.close()


def test_cancel_queued_requests(scene):
    """Queued requests can be cancelled: they never run, their tickets fail with
    FOMO_ERR_CANCELLED and their output buffers stay untouched; requests already admitted or
    finished cannot be cancelled, and the others still equal the synchronous render."""
    w = 0.2 * C / scene["tab"]["lambda0"]
    req = (scene["em"], F.METHOD_NN, 0.7, 1.1, 512, 512, 128, 30)
    ref, _ = F.render_nn(scene["em"], 0.7, 1.1, 512, 512, 128, 30, w)
    sv = F.Server(device_bytes=2 << 30, host_bytes=1 << 30, max_inflight=1, nstreams=1)
    try:
        outs = [np.full(512 * 512 * 30, -7.0, np.float32) for _ in range(24)]
        tks = [sv.submit(*req, lambda_width=w, out=o) for o in outs]
        cancelled = [sv.cancel(t) for t in tks[12:]]
        # one request runs at a time, each for milliseconds: the last twelve are still queued
        assert all(cancelled), cancelled
        assert sv.cancel(tks[12]) is False  # already cancelled
        for t in tks[12:]:
            with pytest.raises(F.FomoError) as ei:
                sv.result(t, 60)
            assert ei.value.code == -7
            assert sv.poll(t) == -7
        for t, o in zip(tks[:12], outs[:12]):
            img, _ = sv.result(t, 60)
            assert np.array_equal(img, ref)
        assert sv.cancel(tks[0]) is False  # finished
        for o in outs[12:]:
            assert np.all(o == -7.0)
        st = sv.stats()
        assert st["completed"] == 12 and st["failed"] == 12 and st["queued"] == 0
    finally:
        sv.close()
