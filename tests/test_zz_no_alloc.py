"""The allocation-free NN render (fomo_render_nn_noalloc) against fomo_render_nn.

Both run the same kernels, so images and meta must agree bit for bit. The caller's device
buffers and stream come from the HIP runtime the engine itself links (libamdhip64.so.7, through
ctypes), as a C caller would make them. Calls run in a row on one stream: a spectral view of a
40^3 cube (per-sample buffers in the scratch array), then a smaller imaging view of a 24^3 cube,
then a 256x256 single-line view (fused pass: per-ray buffers only), then the first view again
(the array holds the previous call's contents). A view whose buffers exceed the array is
rejected.
"""
import ctypes
import os
import tempfile

import numpy as np
import pytest

import fomo_amd as F

C = 2.99792458e8


class Hip:
    """The few HIP runtime calls a caller of the allocation-free entry point makes itself."""

    def __init__(self):
        F.lib()  # the engine loads the runtime; this opens the same instance
        self.L = ctypes.CDLL("libamdhip64.so.7")
        self.L.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.L.hipFree.argtypes = [ctypes.c_void_p]
        self.L.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.L.hipStreamCreate.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
        self.L.hipStreamSynchronize.argtypes = [ctypes.c_void_p]
        self.L.hipStreamDestroy.argtypes = [ctypes.c_void_p]

    def ok(self, rc):
        assert rc == 0, f"HIP error {rc}"

    def malloc(self, nbytes):
        p = ctypes.c_void_p()
        self.ok(self.L.hipMalloc(ctypes.byref(p), nbytes))
        return p.value

    def upload(self, dev, a):
        self.ok(self.L.hipMemcpy(dev, a.ctypes.data, a.nbytes, 1))  # hipMemcpyHostToDevice

    def download(self, a, dev):
        self.ok(self.L.hipMemcpy(a.ctypes.data, dev, a.nbytes, 2))  # hipMemcpyDeviceToHost
        return a


@pytest.fixture(scope="module")
def scenes(engine):
    with tempfile.TemporaryDirectory() as td:
        path = F.gen_goft_table(os.path.join(td, "fe9.dat"))
        tab = F.read_table(path)
        T = F.Table.from_file(path)
    out = {}
    for m, seed in ((40, 11), (24, 12)):
        d = F.gen_loop_cube(m, seed=seed)
        ds = F.Dataset(d["x"], d["y"], d["z"], [d["rho"], d["T"], d["vx"], d["vy"], d["vz"]])
        em = F.Emission.chianti(ds, T, tab["lambda0"], tab["atweight"])
        out[m] = (ds, em)
    yield out, tab
    for ds, em in out.values():
        em.close()
        ds.close()
    T.close()


@pytest.mark.gpu
def test_noalloc_equals_render_nn(scenes):
    cubes, tab = scenes
    hip = Hip()
    width = 0.2 * C / tab["lambda0"]
    views = [
        (40, (0.4, 1.1, 48, 40, 50, 24, width)),   # per-sample buffers (k_nn + k_los_spectral)
        (24, (2.0, 0.3, 20, 18, 16, 1, width)),    # smaller input, imaging
        (40, (0.9, 1.3, 256, 256, 64, 30, width)),  # fused pass, per-ray buffers only
        (40, (0.4, 1.1, 48, 40, 50, 24, width)),   # the first view again
    ]
    cap = max(v[2] * v[3] * max(v[5], 1) for _, v in views)
    d_img = hip.malloc(4 * cap)
    d_meta = hip.malloc(8 * 10)
    stream = ctypes.c_void_p()
    hip.ok(hip.L.hipStreamCreate(ctypes.byref(stream)))
    try:
        for m, view in views:
            l, b, nx, ny, nz, nl, w = view
            print(f"noalloc: cube {m}^3, view {view[:6]}", flush=True)
            ds, em = cubes[m]
            npx = nx * ny * max(nl, 1)
            hip.upload(d_img, np.full(npx, -7.0, np.float32))
            hip.upload(d_meta, np.full(10, -7.0))
            F.render_nn_noalloc(em, l, b, nx, ny, nz, nl, w, d_img, d_meta, stream.value)
            hip.ok(hip.L.hipStreamSynchronize(stream))
            got = hip.download(np.zeros(npx, np.float32), d_img)
            meta = hip.download(np.zeros(10), d_meta)
            ref, rmeta = F.render_nn(em, *view)
            assert np.array_equal(got, ref), f"{(got != ref).sum()} of {npx} pixels differ"
            assert np.array_equal(meta, rmeta)
            assert ref.max() > 0
    finally:
        hip.L.hipStreamDestroy(stream)
        hip.L.hipFree(d_img)
        hip.L.hipFree(d_meta)


@pytest.mark.gpu
def test_noalloc_rejects_oversize_views(scenes):
    cubes, _ = scenes
    ds, em = cubes[40]
    hip = Hip()
    d_img = hip.malloc(16)
    try:
        # imaging, 256 x 256 x 300 samples: about 16 bytes per sample, far above the array
        with pytest.raises(F.FomoError) as ei:
            F.render_nn_noalloc(em, 0.3, 1.0, 256, 256, 300, 1, 1e5, d_img, None, None)
        assert ei.value.code == -1 and "scratch" in str(ei.value)
        with pytest.raises(F.FomoError):
            F.render_nn_noalloc(em, 0.3, 1.0, 16, 16, 8, 1, 1e5, None, None, None)
    finally:
        hip.L.hipFree(d_img)
