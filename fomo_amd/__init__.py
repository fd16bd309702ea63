"""fomo_amd — Python binding of the MI355X FoMo render engine (libfomo_hip.so, include/fomo_hip.h).

The engine is native: HIP kernels for gfx950 behind a C ABI. This module is plumbing for tests,
the benchmark and scripting; it loads the in-tree shared library with ctypes and raises if it is
missing. There is no CPU fallback.
"""
import ctypes
import os

import numpy as np

__all__ = [
    "lib", "FomoError", "init", "shutdown", "device_info", "Dataset", "Table", "Emission", "Batch",
    "render_nn", "render_nn_noalloc", "render_thomson", "nn_indices", "Server", "METHOD_NN", "METHOD_THOMSON", "METHOD_TET", "METHOD_LINEAR", "METHOD_PROJECTION", "delaunay", "set_delaunay_threads", "render_tet",
    "autotune", "tet_vertices", "render_projection", "projection_pixels", "gen_loop_cube", "gen_goft_table", "gen_corona",
    "read_table", "profile_enable", "profile_reset", "profile_stats", "LIB_PATH",
]

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libfomo_hip.so")
_LIB = None

_c_int_p = ctypes.POINTER(ctypes.c_int)
_c_i64_p = ctypes.POINTER(ctypes.c_int64)
_c_dbl_p = ctypes.POINTER(ctypes.c_double)
_c_flt_p = ctypes.POINTER(ctypes.c_float)
_c_i32_p = ctypes.POINTER(ctypes.c_int32)

# name -> (restype, argtypes); mirrors include/fomo_hip.h
_SIGS = {
    "fomo_last_error": (ctypes.c_char_p, []),
    "fomo_version": (ctypes.c_char_p, []),
    "fomo_init": (ctypes.c_int, [ctypes.c_int]),
    "fomo_shutdown": (ctypes.c_int, []),
    "fomo_device_info": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int, _c_dbl_p, _c_int_p]),
    "fomo_dataset_create": (ctypes.c_int, [ctypes.c_int64, ctypes.c_int, _c_flt_p, _c_flt_p, _c_flt_p, ctypes.c_int,
                                           ctypes.POINTER(_c_flt_p), ctypes.c_double, _c_int_p]),
    "fomo_dataset_destroy": (ctypes.c_int, [ctypes.c_int]),
    "fomo_dataset_begin": (ctypes.c_int, [ctypes.c_int64, ctypes.c_int, ctypes.c_int, _c_int_p]),
    "fomo_dataset_append": (ctypes.c_int, [ctypes.c_int, ctypes.c_int64, ctypes.c_int64, _c_flt_p, _c_flt_p, _c_flt_p,
                                           ctypes.POINTER(_c_flt_p)]),
    "fomo_dataset_commit": (ctypes.c_int, [ctypes.c_int, ctypes.c_double]),
    "fomo_dataset_info": (ctypes.c_int, [ctypes.c_int, _c_i64_p, _c_int_p, _c_int_p]),
    "fomo_dataset_reindex": (ctypes.c_int, [ctypes.c_int, ctypes.c_double]),
    "fomo_dataset_set_lookup": (ctypes.c_int, [ctypes.c_int, ctypes.c_int]),
    "fomo_dataset_set_ray_tile": (ctypes.c_int, [ctypes.c_int, ctypes.c_int]),
    "fomo_dataset_set_far_budget": (ctypes.c_int, [ctypes.c_int, ctypes.c_int]),
    "fomo_dataset_tet_stats": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(ctypes.c_int64)]),
    "fomo_dataset_lookup_info": (ctypes.c_int, [ctypes.c_int, _c_int_p, _c_i64_p, _c_i64_p, _c_i64_p]),
    "fomo_dataset_lookup_groups": (ctypes.c_int, [ctypes.c_int, _c_i64_p]),
    "fomo_dataset_set_rectilinear": (ctypes.c_int, [ctypes.c_int, ctypes.c_int]),
    "fomo_dataset_rect_info": (ctypes.c_int, [ctypes.c_int, _c_int_p]),
    "fomo_dataset_grid_info": (ctypes.c_int, [ctypes.c_int, _c_int_p, _c_int_p]),
    "fomo_table_read": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int, _c_dbl_p, _c_dbl_p,
                                       ctypes.c_char_p, ctypes.c_int, _c_int_p, _c_int_p, _c_flt_p, _c_flt_p,
                                       _c_flt_p]),
    "fomo_table_create": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, _c_flt_p, _c_flt_p, _c_flt_p, _c_int_p]),
    "fomo_table_destroy": (ctypes.c_int, [ctypes.c_int]),
    "fomo_emission_chianti": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                             ctypes.c_int, ctypes.c_double, _c_int_p]),
    "fomo_emission_chianti_interp": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                                    ctypes.c_int, ctypes.c_double, ctypes.c_int, _c_int_p]),
    "fomo_emission_thomson": (ctypes.c_int, [ctypes.c_int, _c_int_p]),
    "fomo_emission_given": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_double, _c_int_p]),
    "fomo_goft_destroy": (ctypes.c_int, [ctypes.c_int]),
    "fomo_goft_read": (ctypes.c_int, [ctypes.c_int, _c_flt_p, _c_flt_p]),
    "fomo_render_nn": (ctypes.c_int, [ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_int,
                                      ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_int, _c_flt_p, _c_dbl_p]),
    "fomo_render_nn_noalloc": (ctypes.c_int, [ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                              ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_int,
                                              ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    "fomo_render_thomson": (ctypes.c_int, [ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                           ctypes.c_int, ctypes.c_int, _c_flt_p, _c_dbl_p]),
    "fomo_nn_indices": (ctypes.c_int, [ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_int,
                                       ctypes.c_int, ctypes.c_int, _c_i32_p, _c_dbl_p]),
    "fomo_server_create": (ctypes.c_int, [ctypes.c_int64, ctypes.c_int64, ctypes.c_int, ctypes.c_int, _c_int_p]),
    "fomo_server_destroy": (ctypes.c_int, [ctypes.c_int]),
    "fomo_server_footprint": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                             ctypes.c_int, _c_i64_p, _c_i64_p]),
    "fomo_server_submit": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                          ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                          ctypes.c_int, _c_flt_p, _c_dbl_p, _c_i64_p]),
    "fomo_server_poll": (ctypes.c_int, [ctypes.c_int, ctypes.c_int64, _c_int_p]),
    "fomo_server_wait": (ctypes.c_int, [ctypes.c_int, ctypes.c_int64, ctypes.c_double, _c_int_p]),
    "fomo_server_wait_all": (ctypes.c_int, [ctypes.c_int, ctypes.c_double]),
    "fomo_server_cancel": (ctypes.c_int, [ctypes.c_int, ctypes.c_int64]),
    "fomo_server_times": (ctypes.c_int, [ctypes.c_int, ctypes.c_int64, _c_dbl_p]),
    "fomo_server_stats": (ctypes.c_int, [ctypes.c_int, _c_dbl_p]),
    "fomo_set_delaunay_threads": (ctypes.c_int, [ctypes.c_int]),
    "fomo_delaunay_threads": (ctypes.c_int, []),
    "fomo_delaunay": (ctypes.c_int, [ctypes.c_int64, _c_flt_p, _c_flt_p, _c_flt_p, ctypes.c_int64, _c_i64_p,
                                     _c_i32_p]),
    "fomo_render_tet": (ctypes.c_int, [ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_int,
                                       ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_int,
                                       _c_flt_p, _c_dbl_p]),
    "fomo_render_projection": (ctypes.c_int, [ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                              ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_int,
                                              _c_flt_p, _c_dbl_p]),
    "fomo_projection_pixels": (ctypes.c_int, [ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                              ctypes.c_int, ctypes.c_int, _c_i32_p, _c_dbl_p]),
    "fomo_dataset_triangulate": (ctypes.c_int, [ctypes.c_int, _c_i64_p, _c_dbl_p]),
    "fomo_tet_vertices": (ctypes.c_int, [ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_int,
                                         ctypes.c_int, _c_i32_p, _c_dbl_p]),
    "fomo_autotune": (ctypes.c_int, [ctypes.c_int, _c_int_p, ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                     ctypes.c_int, ctypes.c_int, ctypes.c_int, _c_dbl_p]),
    "fomo_batch_create": (ctypes.c_int, [ctypes.c_int, _c_int_p, ctypes.c_int, _c_dbl_p, _c_dbl_p, ctypes.c_int,
                                         ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_int,
                                         _c_int_p]),
    "fomo_batch_run": (ctypes.c_int, [ctypes.c_int]),
    "fomo_batch_launch": (ctypes.c_int, [ctypes.c_int]),
    "fomo_batch_sync": (ctypes.c_int, [ctypes.c_int]),
    "fomo_batch_profile": (ctypes.c_int, [ctypes.c_int]),
    "fomo_batch_read": (ctypes.c_int, [ctypes.c_int, _c_flt_p, _c_dbl_p]),
    "fomo_batch_read_view": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_int, _c_flt_p, _c_dbl_p]),
    "fomo_batch_destroy": (ctypes.c_int, [ctypes.c_int]),
    "fomo_profile_enable": (ctypes.c_int, [ctypes.c_int]),
    "fomo_profile_reset": (ctypes.c_int, []),
    "fomo_profile_stats": (ctypes.c_int, [ctypes.c_char_p, _c_i64_p, _c_dbl_p]),
    "fomo_copy_probe": (ctypes.c_int, [ctypes.c_int64, ctypes.c_int, _c_dbl_p]),
    "fomo_gather_probe": (ctypes.c_int, [ctypes.c_int64, ctypes.c_int64, ctypes.c_int, _c_dbl_p]),
    "fomo_gen_loop_cube": (ctypes.c_int, [ctypes.c_int, ctypes.c_uint64, ctypes.c_double, ctypes.c_double,
                                          ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                          _c_flt_p, _c_flt_p, _c_flt_p, _c_flt_p, _c_flt_p, _c_flt_p, _c_flt_p,
                                          _c_flt_p]),
    "fomo_gen_goft_table": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_double, ctypes.c_double,
                                           ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double]),
    "fomo_gen_corona": (ctypes.c_int, [ctypes.c_int, ctypes.c_uint64, ctypes.c_double, _c_i64_p, _c_flt_p,
                                       _c_flt_p, _c_flt_p, _c_flt_p]),
}


class FomoError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"fomo error {code}: {msg}")
        self.code = code


def lib():
    """Load libfomo_hip.so (in-tree). Raises if it has not been built."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m fomo_amd.build`")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = L
    return _LIB


def _check(rc):
    if rc != 0:
        raise FomoError(rc, lib().fomo_last_error().decode())


def _f32(a):
    a = np.ascontiguousarray(a, dtype=np.float32)
    return a, a.ctypes.data_as(_c_flt_p)


def _ptr(a, ct):
    return a.ctypes.data_as(ctypes.POINTER(ct))


def init(device=0):
    _check(lib().fomo_init(device))


def shutdown():
    _check(lib().fomo_shutdown())


def device_info():
    buf = ctypes.create_string_buffer(256)
    hbm = ctypes.c_double()
    cus = ctypes.c_int()
    _check(lib().fomo_device_info(buf, 256, ctypes.byref(hbm), ctypes.byref(cus)))
    return buf.value.decode(), hbm.value, cus.value


def profile_enable(on=True):
    _check(lib().fomo_profile_enable(1 if on else 0))


def profile_reset():
    _check(lib().fomo_profile_reset())


def copy_probe(nbytes, reps=10):
    """Mean ms of a streaming device copy of nbytes (reads nbytes, writes nbytes)."""
    ms = ctypes.c_double()
    _check(lib().fomo_copy_probe(int(nbytes), reps, ctypes.byref(ms)))
    return ms.value


def gather_probe(nbytes, lines, reps=10):
    """Mean ms of reading `lines` random 128-byte lines of an nbytes buffer."""
    ms = ctypes.c_double()
    _check(lib().fomo_gather_probe(int(nbytes), int(lines), reps, ctypes.byref(ms)))
    return ms.value


def profile_stats(kernel):
    c = ctypes.c_int64()
    ms = ctypes.c_double()
    _check(lib().fomo_profile_stats(kernel.encode(), ctypes.byref(c), ctypes.byref(ms)))
    return c.value, ms.value


class Dataset:
    """Device-resident point cloud + cell-list index (DataCube on the device)."""

    def __init__(self, x, y, z, variables, dim=3, cell_scale=0.0):
        self._keep = []
        xs, xp = _f32(x)
        ys, yp = _f32(y)
        if dim == 3:
            zs, zp = _f32(z)
        else:
            zs, zp = None, None
        self._keep += [xs, ys, zs]
        n = xs.size
        vars_c = []
        for v in variables:
            vs, vp = _f32(v)
            if vs.size != n:
                raise ValueError("variable length differs from the coordinate length")
            self._keep.append(vs)
            vars_c.append(vp)
        arr = (_c_flt_p * max(1, len(vars_c)))(*vars_c)
        out = ctypes.c_int()
        _check(lib().fomo_dataset_create(n, dim, xp, yp, zp, len(vars_c), arr, cell_scale, ctypes.byref(out)))
        self.id = out.value
        self.n = n
        self._keep = None

    @classmethod
    def streamed(cls, n, chunks, dim=3, nvars=0, cell_scale=0.0):
        """A dataset filled by the streaming ingest: chunks yields (offset, x, y, z, [vars]) with
        float32 arrays of one length; every point must be covered once."""
        out = ctypes.c_int()
        _check(lib().fomo_dataset_begin(int(n), dim, nvars, ctypes.byref(out)))
        self = cls.__new__(cls)
        self.id, self.n, self._keep = out.value, int(n), None
        for off, x, y, z, variables in chunks:
            xs, xp = _f32(x)
            ys, yp = _f32(y)
            zs, zp = _f32(z) if dim == 3 else (None, None)
            vs = [_f32(v) for v in variables]
            arr = (_c_flt_p * max(1, len(vs)))(*[v[1] for v in vs])
            _check(lib().fomo_dataset_append(self.id, int(off), xs.size, xp, yp, zp, arr))
        _check(lib().fomo_dataset_commit(self.id, cell_scale))
        return self

    def info(self):
        n = ctypes.c_int64()
        dims = (ctypes.c_int * 3)()
        mx = ctypes.c_int()
        _check(lib().fomo_dataset_info(self.id, ctypes.byref(n), dims, ctypes.byref(mx)))
        return n.value, tuple(dims), mx.value

    def set_ray_tile(self, width):
        """Ray tile width (8, 16, 32 or 64 rays) of the NN passes' blocks. Results do not change."""
        _check(lib().fomo_dataset_set_ray_tile(self.id, width))

    def tet_stats(self):
        """Tetra path counts since the triangulation was built: samples walked again with exact
        orientation tests, and samples left unresolved."""
        out = (ctypes.c_int64 * 2)()
        _check(lib().fomo_dataset_tet_stats(self.id, out))
        return {"exact_rewalks": int(out[0]), "unresolved": int(out[1])}

    def set_far_budget(self, steps):
        """Node steps of an exact query's walk before it moves to the parallel range walk (0 = never).
        Results do not change."""
        _check(lib().fomo_dataset_set_far_budget(self.id, steps))

    def set_lookup(self, subdiv):
        """Candidate lookup grid subdivision (0 = ring search only). Results do not change."""
        _check(lib().fomo_dataset_set_lookup(self.id, subdiv))

    def lookup_info(self):
        """Candidate lookup grid: active subdivision (0 = off), cells, candidates summed over the
        cells (a far group's shared list once per cell using it), cells left to the exact query."""
        sd = ctypes.c_int()
        cells, cand, fb = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
        _check(lib().fomo_dataset_lookup_info(self.id, ctypes.byref(sd), ctypes.byref(cells), ctypes.byref(cand),
                                              ctypes.byref(fb)))
        return dict(subdiv=sd.value, cells=cells.value, candidates=cand.value, fallback_cells=fb.value)

    def set_rectilinear(self, on):
        """Index a product-grid dataset by its axes (on, the default) or as scattered points (off).
        Results do not change."""
        _check(lib().fomo_dataset_set_rectilinear(self.id, 1 if on else 0))

    def rect_axes(self):
        """Values per axis of a rectilinear dataset's grid, or None for a scattered one."""
        a = (ctypes.c_int * 3)()
        _check(lib().fomo_dataset_rect_info(self.id, a))
        return tuple(a) if a[0] else None

    def grid_info(self):
        """Index kind ("scattered", "rectilinear" or "masked") and the grid's values per axis."""
        k = ctypes.c_int()
        a = (ctypes.c_int * 3)()
        _check(lib().fomo_dataset_grid_info(self.id, ctypes.byref(k), a))
        return ("scattered", "rectilinear", "masked")[k.value], (tuple(a) if a[0] else None)

    def lookup_groups(self):
        """Far and outer groups of the lookup grid and how many hold a shared candidate list."""
        c = np.zeros(5, np.int64)
        _check(lib().fomo_dataset_lookup_groups(self.id, _ptr(c, ctypes.c_int64)))
        return dict(far_groups=int(c[0]), far_lists=int(c[1]), full_groups=int(c[2]), outer_groups=int(c[3]),
                    outer_lists=int(c[4]))

    def triangulate(self):
        """Build the Delaunay triangulation for the tetra path; returns (tets, seconds)."""
        nt, sec = ctypes.c_int64(), ctypes.c_double()
        _check(lib().fomo_dataset_triangulate(self.id, ctypes.byref(nt), ctypes.byref(sec)))
        return nt.value, sec.value

    def reindex(self, cell_scale=0.0):
        _check(lib().fomo_dataset_reindex(self.id, cell_scale))

    def close(self):
        if self.id is not None and _LIB is not None:
            lib().fomo_dataset_destroy(self.id)
            self.id = None


def read_table(path):
    """Parse a CHIANTI G(T,n) table (returns a dict with ion, lambda0, atweight, unit, logt, logn, g)."""
    ion = ctypes.create_string_buffer(256)
    unit = ctypes.create_string_buffer(1024)
    l0 = ctypes.c_double()
    aw = ctypes.c_double()
    nt = ctypes.c_int()
    nr = ctypes.c_int()
    _check(lib().fomo_table_read(path.encode(), ion, 256, ctypes.byref(l0), ctypes.byref(aw), unit, 1024,
                                 ctypes.byref(nt), ctypes.byref(nr), None, None, None))
    logt = np.zeros(nt.value, np.float32)
    logn = np.zeros(nr.value, np.float32)
    g = np.zeros(nt.value * nr.value, np.float32)
    _check(lib().fomo_table_read(path.encode(), ion, 256, ctypes.byref(l0), ctypes.byref(aw), unit, 1024,
                                 ctypes.byref(nt), ctypes.byref(nr), _ptr(logt, ctypes.c_float),
                                 _ptr(logn, ctypes.c_float), _ptr(g, ctypes.c_float)))
    return dict(ion=ion.value.decode(), lambda0=l0.value, atweight=aw.value, unit=unit.value.decode(),
                nt=nt.value, nrho=nr.value, logt=logt, logn=logn, g=g)


class Table:
    """Device-resident G(T,n) table."""

    def __init__(self, logt, logn, g):
        lt, ltp = _f32(logt)
        ln, lnp = _f32(logn)
        gg, gp = _f32(g)
        out = ctypes.c_int()
        _check(lib().fomo_table_create(lt.size, ln.size, ltp, lnp, gp, ctypes.byref(out)))
        self.id = out.value

    @classmethod
    def from_file(cls, path):
        t = read_table(path)
        obj = cls(t["logt"], t["logn"], t["g"])
        obj.meta = t
        return obj

    def close(self):
        if self.id is not None and _LIB is not None:
            lib().fomo_table_destroy(self.id)
            self.id = None


class Emission:
    """Device-resident emission cube (GoftCube): chianti (peak, fwhm) or thomson (n tot, n pol)."""

    def __init__(self, gid, n):
        self.id = gid
        self.n = n

    @classmethod
    def chianti(cls, ds, table, lambda0, atweight, imaging=False, abundratio=1.0, bilinear=False):
        """bilinear=True (an extension) interpolates G(T, n) linearly in log n too, where the
        reference takes the nearest density row."""
        out = ctypes.c_int()
        _check(lib().fomo_emission_chianti_interp(ds.id, table.id, lambda0, atweight, 1 if imaging else 0,
                                                  abundratio, 1 if bilinear else 0, ctypes.byref(out)))
        return cls(out.value, ds.n)

    @classmethod
    def thomson(cls, ds):
        out = ctypes.c_int()
        _check(lib().fomo_emission_thomson(ds.id, ctypes.byref(out)))
        return cls(out.value, ds.n)

    @classmethod
    def given(cls, ds, lambda0=0.0, thomson=False):
        """The emission of a dataset holding a computed GoftCube: variables (peak, fwhm, vx, vy, vz)
        of a line at lambda0, or (n tot, n pol) with thomson=True."""
        out = ctypes.c_int()
        _check(lib().fomo_emission_given(ds.id, 1 if thomson else 0, lambda0, ctypes.byref(out)))
        return cls(out.value, ds.n)

    def read(self):
        a = np.zeros(self.n, np.float32)
        b = np.zeros(self.n, np.float32)
        _check(lib().fomo_goft_read(self.id, _ptr(a, ctypes.c_float), _ptr(b, ctypes.c_float)))
        return a, b

    def close(self):
        if self.id is not None and _LIB is not None:
            lib().fomo_goft_destroy(self.id)
            self.id = None


def render_nn(em, l, b, nx, ny, nz, nl, lambda_width, dn=False, out=True):
    """One view; returns (image[ny*nx*max(nl,1)] float32 or None, meta[10])."""
    meta = np.zeros(10, np.float64)
    img = np.zeros(nx * ny * max(nl, 1), np.float32) if out else None
    _check(lib().fomo_render_nn(em.id, l, b, nx, ny, nz, nl, lambda_width, 1 if dn else 0,
                                _ptr(img, ctypes.c_float) if out else None, _ptr(meta, ctypes.c_double)))
    return img, meta


def render_nn_noalloc(em, l, b, nx, ny, nz, nl, lambda_width, out_dev, meta_dev=None, stream=None, dn=False):
    """Allocation-free render_nn: enqueues on `stream` (a hipStream_t as an int, None = the
    library's stream) into device buffers given as integer addresses, and returns at once."""
    _check(lib().fomo_render_nn_noalloc(em.id, l, b, nx, ny, nz, nl, lambda_width, 1 if dn else 0, out_dev,
                                        meta_dev, stream))


def render_thomson(em, l, b, nx, ny, nz):
    meta = np.zeros(10, np.float64)
    img = np.zeros(nx * ny, np.float32)
    _check(lib().fomo_render_thomson(em.id, l, b, nx, ny, nz, _ptr(img, ctypes.c_float), _ptr(meta, ctypes.c_double)))
    return img, meta


def autotune(ems, l, b, nx, ny, nz, reps=3):
    """Pick the fastest index configuration (cell scale x lookup subdivision x ray tile) for views
    like this one of the given emission cube(s) (a batch's lines); the dataset keeps it. Returns
    dict(cell_scale, subdiv, ms, ms_before, tile)."""
    ems = ems if isinstance(ems, (list, tuple)) else [ems]
    ids = (ctypes.c_int * len(ems))(*[e.id for e in ems])
    r = np.zeros(5, np.float64)
    _check(lib().fomo_autotune(len(ems), ids, l, b, nx, ny, nz, reps, _ptr(r, ctypes.c_double)))
    return {"cell_scale": float(r[0]), "subdiv": int(r[1]), "ms": float(r[2]), "ms_before": float(r[3]),
            "tile": int(r[4])}


def render_tet(em, l, b, nx, ny, nz, nl, lambda_width, dn=False, linear=False):
    """One view of the tetra path (nearest vertex of the containing Delaunay tet, or barycentric
    with linear=True); returns (image[ny*nx*max(nl,1)] float32, meta[10])."""
    meta = np.zeros(10, np.float64)
    img = np.zeros(nx * ny * max(nl, 1), np.float32)
    _check(lib().fomo_render_tet(em.id, l, b, nx, ny, nz, nl, lambda_width, 1 if dn else 0, 1 if linear else 0,
                                 _ptr(img, ctypes.c_float), _ptr(meta, ctypes.c_double)))
    return img, meta


def render_projection(em, l, b, nx, ny, nz, nl, lambda_width, dn=False):
    """One view of the projection method (each point's line profile added to the pixel of its
    rotated position); returns (image[ny*nx*max(nl,1)] float32, meta[10])."""
    meta = np.zeros(10, np.float64)
    img = np.zeros(nx * ny * max(nl, 1), np.float32)
    _check(lib().fomo_render_projection(em.id, l, b, nx, ny, nz, nl, lambda_width, 1 if dn else 0,
                                        _ptr(img, ctypes.c_float), _ptr(meta, ctypes.c_double)))
    return img, meta


def projection_pixels(ds, l, b, nx, ny, nz=2):
    """Pixel iy*nx + jx of every point (original order) for one projection view; returns (pixels, meta)."""
    out = np.zeros(ds.n, np.int32)
    meta = np.zeros(10, np.float64)
    _check(lib().fomo_projection_pixels(ds.id, l, b, nx, ny, nz, _ptr(out, ctypes.c_int32),
                                        _ptr(meta, ctypes.c_double)))
    return out, meta


def tet_vertices(ds, l, b, nx, ny, nz):
    out = np.zeros(nx * ny * nz, np.int32)
    meta = np.zeros(12, np.float64)
    _check(lib().fomo_tet_vertices(ds.id, l, b, nx, ny, nz, _ptr(out, ctypes.c_int32), _ptr(meta, ctypes.c_double)))
    return out, meta


def nn_indices(ds, l, b, nx, ny, nz, thomson=False):
    idx = np.zeros(nx * ny * nz, np.int32)
    meta = np.zeros(12, np.float64)
    _check(lib().fomo_nn_indices(ds.id, l, b, nx, ny, nz, 1 if thomson else 0, _ptr(idx, ctypes.c_int32),
                                 _ptr(meta, ctypes.c_double)))
    return idx, meta


class Batch:
    """All views x lines of one dataset in one captured HIP graph; results stay on the device."""

    def __init__(self, emissions, ls, bs, nx, ny, nz, nl, lambda_width, dn=False):
        ids = (ctypes.c_int * len(emissions))(*[e.id for e in emissions])
        la = np.ascontiguousarray(ls, np.float64)
        ba = np.ascontiguousarray(bs, np.float64)
        if la.size != ba.size:
            raise ValueError("l and b lists must have equal length (one entry per view)")
        out = ctypes.c_int()
        _check(lib().fomo_batch_create(len(emissions), ids, la.size, _ptr(la, ctypes.c_double),
                                       _ptr(ba, ctypes.c_double), nx, ny, nz, nl, lambda_width, 1 if dn else 0,
                                       ctypes.byref(out)))
        self.id = out.value
        self.shape = (la.size, len(emissions), ny, nx, max(nl, 1))

    def run(self):
        _check(lib().fomo_batch_run(self.id))

    def launch(self):
        _check(lib().fomo_batch_launch(self.id))

    def sync(self):
        _check(lib().fomo_batch_sync(self.id))

    def profile(self):
        """One eager pass of the batch's kernels with HIP events around each launch."""
        _check(lib().fomo_batch_profile(self.id))

    def read(self, out=None):
        """All images [view][line][ny][nx][nl] (into `out` when given) and the views' meta."""
        if out is None:
            out = np.zeros(self.shape, np.float32)
        elif out.shape != self.shape or out.dtype != np.float32 or not out.flags.c_contiguous:
            raise ValueError(f"out must be a contiguous float32 array of shape {self.shape}")
        meta = np.zeros((self.shape[0], 10), np.float64)
        _check(lib().fomo_batch_read(self.id, _ptr(out, ctypes.c_float), _ptr(meta, ctypes.c_double)))
        return out, meta

    def read_view(self, view, line):
        """One image (ny*nx*max(nl,1) floats) and its view's meta."""
        out = np.zeros(self.shape[2] * self.shape[3] * self.shape[4], np.float32)
        meta = np.zeros(10, np.float64)
        _check(lib().fomo_batch_read_view(self.id, view, line, _ptr(out, ctypes.c_float), _ptr(meta, ctypes.c_double)))
        return out, meta

    def close(self):
        if self.id is not None and _LIB is not None:
            lib().fomo_batch_destroy(self.id)
            self.id = None


METHOD_NN = 0
METHOD_THOMSON = 1
METHOD_TET = 2
METHOD_LINEAR = 3
METHOD_PROJECTION = 4


class Server:
    """Render server over resident emission cubes (continuous batching on several streams).

    submit() queues one view and returns a ticket; result() waits for it and returns
    (image, meta). Output arrays are allocated here and kept alive until the ticket finishes."""

    def __init__(self, device_bytes=4 << 30, host_bytes=1 << 30, max_inflight=8, nstreams=4):
        out = ctypes.c_int()
        _check(lib().fomo_server_create(int(device_bytes), int(host_bytes), max_inflight, nstreams, ctypes.byref(out)))
        self.id = out.value
        self._bufs = {}

    @staticmethod
    def footprint(em, method, nx, ny, nz, nl=1):
        d, h = ctypes.c_int64(), ctypes.c_int64()
        _check(lib().fomo_server_footprint(em.id, method, nx, ny, nz, nl, ctypes.byref(d), ctypes.byref(h)))
        return d.value, h.value

    def submit(self, em, method, l, b, nx, ny, nz, nl=1, lambda_width=0.0, dn=False, out=None):
        """out: optional float32 array of ny*nx*max(nl,1) values to receive the image."""
        nlo = 1 if method == METHOD_THOMSON else max(nl, 1)
        if out is None:
            img = np.empty(nx * ny * nlo, np.float32)
        else:
            img = out
            if img.dtype != np.float32 or not img.flags.c_contiguous or img.size != nx * ny * nlo:
                raise ValueError("out must be a contiguous float32 array of ny*nx*max(nl,1) values")
        meta = np.zeros(10, np.float64)
        t = ctypes.c_int64()
        _check(lib().fomo_server_submit(self.id, em.id, method, l, b, nx, ny, nz, nl, lambda_width, 1 if dn else 0,
                                        _ptr(img, ctypes.c_float), _ptr(meta, ctypes.c_double), ctypes.byref(t)))
        self._bufs[t.value] = (img, meta)
        return t.value

    def poll(self, ticket):
        st = ctypes.c_int()
        rc = lib().fomo_server_poll(self.id, ticket, ctypes.byref(st))
        if rc != 0 and st.value >= 0:
            _check(rc)
        return st.value

    def result(self, ticket, timeout=-1.0):
        """Wait for the ticket; returns (image, meta). Raises FomoError if it failed."""
        st = ctypes.c_int()
        _check(lib().fomo_server_wait(self.id, ticket, float(timeout), ctypes.byref(st)))
        return self._bufs.pop(ticket)

    def cancel(self, ticket):
        """Cancel a queued request. True when it was cancelled, False when it had already been
        admitted or had finished."""
        rc = lib().fomo_server_cancel(self.id, ticket)
        if rc == -5:
            return False
        _check(rc)
        self._bufs.pop(ticket, None)
        return True

    def wait_all(self, timeout=-1.0):
        _check(lib().fomo_server_wait_all(self.id, float(timeout)))

    def times(self, ticket):
        """(submitted, admitted, finished) in seconds since the server was created (-1: not yet)."""
        t = np.zeros(3, np.float64)
        _check(lib().fomo_server_times(self.id, ticket, _ptr(t, ctypes.c_double)))
        return t

    def stats(self):
        st = np.zeros(12, np.float64)
        _check(lib().fomo_server_stats(self.id, _ptr(st, ctypes.c_double)))
        keys = ("submitted", "completed", "failed", "queued", "running", "peak_running", "device_in_use",
                "device_peak", "device_budget", "host_in_use", "host_peak", "host_budget")
        return dict(zip(keys, (int(v) for v in st)))

    def close(self):
        if self.id is not None and _LIB is not None:
            lib().fomo_server_destroy(self.id)
            self.id = None
            self._bufs.clear()


def set_delaunay_threads(threads=0):
    """Worker threads of the host Delaunay builder (0: FOMO_DELAUNAY_THREADS, else OMP_NUM_THREADS,
    else the hardware threads). The tets and their order do not depend on it."""
    _check(lib().fomo_set_delaunay_threads(int(threads)))


def delaunay(x, y, z):
    """Delaunay tets of the points (host builder): int32 array (ntets, 4) of point indices, in a
    canonical order (by vertex, in the Morton order of the points)."""
    (xa, xp), (ya, yp), (za, zp) = _f32(x), _f32(y), _f32(z)
    n = xa.size
    cnt = ctypes.c_int64()
    cap = 8 * n + 64
    while True:
        out = np.zeros((cap, 4), np.int32)
        _check(lib().fomo_delaunay(n, xp, yp, zp, cap, ctypes.byref(cnt), _ptr(out, ctypes.c_int32)))
        if cnt.value <= cap:
            return out[: cnt.value].copy()
        cap = cnt.value


def gen_loop_cube(m, seed=1, t_in=1.0e6, t_out=2.0e6, noise_sigma=0.05, vamp=2.0e4, jitter=0.0, sausage=0.0):
    """Straight cylindrical loop in a 10 Mm box on an m^3 grid; returns dict of float32 arrays."""
    n = m ** 3
    arrs = {k: np.zeros(n, np.float32) for k in ("x", "y", "z", "rho", "T", "vx", "vy", "vz")}
    _check(lib().fomo_gen_loop_cube(m, seed, t_in, t_out, noise_sigma, vamp, jitter, sausage,
                                    *[_ptr(arrs[k], ctypes.c_float) for k in ("x", "y", "z", "rho", "T", "vx", "vy",
                                                                              "vz")]))
    return arrs


def gen_goft_table(path, ion="fe_9", lambda0=171.0, atweight=56.0, nt=101, nrho=41, peak_logt=5.9, sigma=0.15):
    _check(lib().fomo_gen_goft_table(path.encode(), ion.encode(), lambda0, atweight, nt, nrho, peak_logt, sigma))
    return path


def gen_corona(m, seed=3, sigma=0.1):
    cnt = ctypes.c_int64()
    _check(lib().fomo_gen_corona(m, seed, sigma, ctypes.byref(cnt), None, None, None, None))
    arrs = {k: np.zeros(cnt.value, np.float32) for k in ("x", "y", "z", "rho")}
    _check(lib().fomo_gen_corona(m, seed, sigma, ctypes.byref(cnt), *[_ptr(arrs[k], ctypes.c_float)
                                                                      for k in ("x", "y", "z", "rho")]))
    return arrs
