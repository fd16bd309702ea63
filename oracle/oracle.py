"""ctypes wrapper of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg, never by the product (fomo_amd/). Mode 0 = the reference's float pipeline,
mode 1 = FP64 throughout (the accuracy target).
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
_LIB = None

_f = ctypes.POINTER(ctypes.c_float)
_d = ctypes.POINTER(ctypes.c_double)
_i = ctypes.POINTER(ctypes.c_int32)
_ip = ctypes.POINTER(ctypes.c_int)
_u8 = ctypes.POINTER(ctypes.c_uint8)


def build():
    subprocess.check_call(["make", "-s", "-C", HERE])


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            build()
        _LIB = load(LIB_PATH)
    return _LIB


def load(path):
    """Open an oracle library at `path` with its entry points typed (tests also load variants
    built from edited copies of the source, to show their checks catch planted mistakes)."""
    L = ctypes.CDLL(path)
    L.orc_table_read.restype = ctypes.c_int
    L.orc_table_read.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int, _d, _d, ctypes.c_char_p,
                                 ctypes.c_int, _ip, _ip, _f, _f, _f]
    L.orc_emission.restype = ctypes.c_int64
    L.orc_emission.argtypes = [ctypes.c_int, ctypes.c_int64, _f, _f, ctypes.c_int, ctypes.c_int, _f, _f, _f,
                               ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_double, _d, _d]
    L.orc_thomson_factor.restype = None
    L.orc_thomson_factor.argtypes = [ctypes.c_int, ctypes.c_double, _d, _d]
    L.orc_thomson_emission.restype = None
    L.orc_thomson_emission.argtypes = [ctypes.c_int, ctypes.c_int64, _f, _f, _f, _f, _d, _d]
    L.orc_extents.restype = None
    L.orc_extents.argtypes = [ctypes.c_int64, _f, _f, _f, ctypes.c_double, ctypes.c_double, _d]
    for nm in ("orc_nn_bruteforce", "orc_nn_kdtree"):
        fn = getattr(L, nm)
        fn.restype = None
        fn.argtypes = [ctypes.c_int64, _f, _f, _f, ctypes.c_int64, _d, _d, _d, ctypes.c_double, _i]
    L.orc_render_nn.restype = ctypes.c_int
    L.orc_render_nn.argtypes = [ctypes.c_int, ctypes.c_int64, _f, _f, _f, _d, _d, _f, _f, _f, ctypes.c_double,
                                ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                ctypes.c_double, ctypes.c_double, ctypes.c_int, _d, _i, _d, _d, ctypes.c_int,
                                ctypes.c_int]
    L.orc_render_thomson.restype = ctypes.c_int
    L.orc_render_thomson.argtypes = [ctypes.c_int, ctypes.c_int64, _f, _f, _f, _d, _d, ctypes.c_double,
                                     ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_int, _d, _i, _u8, _d, _d,
                                     ctypes.c_int, ctypes.c_int]
    L.orc_render_nn_multi.restype = ctypes.c_int
    L.orc_render_nn_multi.argtypes = [ctypes.c_int, ctypes.c_int64, _f, _f, _f, ctypes.c_int, _d, _d, _f, _f, _f,
                                      ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                      ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_int, _d, _i, _u8, _d,
                                      _d, ctypes.c_int, ctypes.c_int]
    L.orc_render_projection.restype = ctypes.c_int
    L.orc_render_projection.argtypes = [ctypes.c_int, ctypes.c_int64, _f, _f, _f, _d, _d, _f, _f, _f,
                                        ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_int,
                                        ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                        ctypes.c_int, _d, _i, _d]
    L.orc_num_threads.restype = ctypes.c_int
    L.orc_rotate.restype = None
    L.orc_rotate.argtypes = [ctypes.c_int, ctypes.c_int64, _f, _f, _f, ctypes.c_double, ctypes.c_double, _d, _d,
                             _d, _d]
    L.orc_rotate_d.restype = None
    L.orc_rotate_d.argtypes = [ctypes.c_int64, _d, _d, _d, ctypes.c_double, ctypes.c_double, _d, _d, _d]
    L.orc_los_velocity.restype = None
    L.orc_los_velocity.argtypes = [ctypes.c_int64, _f, _f, _f, ctypes.c_double, ctypes.c_double, _d]
    L.orc_view_samples.restype = None
    L.orc_view_samples.argtypes = [ctypes.c_int, _d, ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                   ctypes.c_int, ctypes.c_int, _d, _d, _d]
    L.orc_maxdistance.restype = ctypes.c_double
    L.orc_maxdistance.argtypes = [ctypes.c_int, _d, ctypes.c_int, ctypes.c_int, ctypes.c_double]
    return L


def _pf(a):
    return np.ascontiguousarray(a, np.float32)


def _pd(a):
    return np.ascontiguousarray(a, np.float64)


def P(a, t):
    return a.ctypes.data_as(ctypes.POINTER(t))


def read_table(path):
    ion = ctypes.create_string_buffer(256)
    unit = ctypes.create_string_buffer(1024)
    l0, aw = ctypes.c_double(), ctypes.c_double()
    nt, nr = ctypes.c_int(), ctypes.c_int()
    rc = lib().orc_table_read(path.encode(), ion, 256, ctypes.byref(l0), ctypes.byref(aw), unit, 1024,
                              ctypes.byref(nt), ctypes.byref(nr), None, None, None)
    if rc != 0:
        raise IOError(f"oracle table read failed ({rc}) for {path}")
    logt = np.zeros(nt.value, np.float32)
    logn = np.zeros(nr.value, np.float32)
    g = np.zeros(nt.value * nr.value, np.float32)
    lib().orc_table_read(path.encode(), ion, 256, ctypes.byref(l0), ctypes.byref(aw), unit, 1024, ctypes.byref(nt),
                         ctypes.byref(nr), P(logt, ctypes.c_float), P(logn, ctypes.c_float), P(g, ctypes.c_float))
    return dict(ion=ion.value.decode(), lambda0=l0.value, atweight=aw.value, unit=unit.value.decode(), nt=nt.value,
                nrho=nr.value, logt=logt, logn=logn, g=g)


def emission(mode, rho, T, table, imaging=False, abundratio=1.0):
    rho, T = _pf(rho), _pf(T)
    n = rho.size
    peak = np.zeros(n)
    fwhm = np.zeros(n)
    lib().orc_emission(mode, n, P(rho, ctypes.c_float), P(T, ctypes.c_float), table["nt"], table["nrho"],
                       P(_pf(table["logt"]), ctypes.c_float), P(_pf(table["logn"]), ctypes.c_float),
                       P(_pf(table["g"]), ctypes.c_float), table["lambda0"], table["atweight"], 1 if imaging else 0,
                       abundratio, P(peak, ctypes.c_double), P(fwhm, ctypes.c_double))
    return peak, fwhm


def thomson_factor(mode, r):
    t, p = ctypes.c_double(), ctypes.c_double()
    lib().orc_thomson_factor(mode, r, ctypes.byref(t), ctypes.byref(p))
    return t.value, p.value


def thomson_emission(mode, x, y, z, dens):
    x, y, z, dens = _pf(x), _pf(y), _pf(z), _pf(dens)
    n = x.size
    tot, pol = np.zeros(n), np.zeros(n)
    lib().orc_thomson_emission(mode, n, P(x, ctypes.c_float), P(y, ctypes.c_float), P(z, ctypes.c_float),
                               P(dens, ctypes.c_float), P(tot, ctypes.c_double), P(pol, ctypes.c_double))
    return tot, pol


def extents(x, y, z, l, b):
    x, y, z = _pf(x), _pf(y), _pf(z)
    e = np.zeros(6)
    lib().orc_extents(x.size, P(x, ctypes.c_float), P(y, ctypes.c_float), P(z, ctypes.c_float), l, b,
                      P(e, ctypes.c_double))
    return e


def rotate(x, y, z, l, b, thomson=False):
    """Rotated coordinates of data points as the renders compute them. thomson=True: the blended
    rotation of fomo-thomson.cpp:136-191, returning sin^2 chi as a fourth array."""
    x, y, z = _pf(x), _pf(y), _pf(z)
    n = x.size
    xr, yr, zr, s2 = np.zeros(n), np.zeros(n), np.zeros(n), np.zeros(n)
    lib().orc_rotate(1 if thomson else 0, n, P(x, ctypes.c_float), P(y, ctypes.c_float), P(z, ctypes.c_float), l,
                     b, P(xr, ctypes.c_double), P(yr, ctypes.c_double), P(zr, ctypes.c_double),
                     P(s2, ctypes.c_double))
    return (xr, yr, zr, s2) if thomson else (xr, yr, zr)


def rotate_d(x, y, z, l, b):
    """The NearestNeighbour rotation applied to double coordinates (to rotate samples back)."""
    x, y, z = _pd(x), _pd(y), _pd(z)
    n = x.size
    xr, yr, zr = np.zeros(n), np.zeros(n), np.zeros(n)
    lib().orc_rotate_d(n, P(x, ctypes.c_double), P(y, ctypes.c_double), P(z, ctypes.c_double), l, b,
                       P(xr, ctypes.c_double), P(yr, ctypes.c_double), P(zr, ctypes.c_double))
    return xr, yr, zr


def los_velocity(vx, vy, vz, l, b):
    vx, vy, vz = _pf(vx), _pf(vy), _pf(vz)
    out = np.zeros(vx.size)
    lib().orc_los_velocity(vx.size, P(vx, ctypes.c_float), P(vy, ctypes.c_float), P(vz, ctypes.c_float), l, b,
                           P(out, ctypes.c_double))
    return out


def view_samples(ext6, l, b, nx, ny, nz, cgal=False):
    """Data-frame positions of every LOS sample, each of shape (ny, nx, nz)."""
    e = _pd(ext6)
    p = [np.zeros(nx * ny * nz) for _ in range(3)]
    lib().orc_view_samples(2 if cgal else 0, P(e, ctypes.c_double), l, b, nx, ny, nz,
                           *[P(a, ctypes.c_double) for a in p])
    return tuple(a.reshape(ny, nx, nz) for a in p)


def maxdistance(ext6, nx, ny, b=0.0, thomson=False):
    e = _pd(ext6)
    return lib().orc_maxdistance(1 if thomson else 0, P(e, ctypes.c_double), nx, ny, b)


def nn_query(x, y, z, px, py, pz, md, method="kdtree"):
    x, y, z = _pf(x), _pf(y), _pf(z)
    px, py, pz = _pd(px), _pd(py), _pd(pz)
    out = np.zeros(px.size, np.int32)
    fn = lib().orc_nn_kdtree if method == "kdtree" else lib().orc_nn_bruteforce
    fn(x.size, P(x, ctypes.c_float), P(y, ctypes.c_float), P(z, ctypes.c_float), px.size, P(px, ctypes.c_double),
       P(py, ctypes.c_double), P(pz, ctypes.c_double), md, P(out, ctypes.c_int32))
    return out


def render_nn(mode, x, y, z, peak, fwhm, vx, vy, vz, l, b, nx, ny, nz, nl, lambda_width, lambda0, dn=False,
              want_idx=False, rows=None):
    """Returns (image float64 [ny*nx*max(nl,1)], idx int32 [ny*nx*nz] or None, meta[10], timing[3])."""
    x, y, z, vx, vy, vz = map(_pf, (x, y, z, vx, vy, vz))
    peak, fwhm = _pd(peak), _pd(fwhm)
    out = np.zeros(nx * ny * max(nl, 1))
    idx = np.zeros(nx * ny * nz, np.int32) if want_idx else None
    meta = np.zeros(10)
    tim = np.zeros(3)
    rc = lib().orc_render_nn(mode, x.size, P(x, ctypes.c_float), P(y, ctypes.c_float), P(z, ctypes.c_float),
                             P(peak, ctypes.c_double), P(fwhm, ctypes.c_double), P(vx, ctypes.c_float),
                             P(vy, ctypes.c_float), P(vz, ctypes.c_float), l, b, nx, ny, nz, nl, lambda_width,
                             lambda0, 1 if dn else 0, P(out, ctypes.c_double),
                             P(idx, ctypes.c_int32) if want_idx else None, P(meta, ctypes.c_double),
                             P(tim, ctypes.c_double), rows[0] if rows else 0, rows[1] if rows else 0)
    if rc != 0:
        raise ValueError(f"oracle render_nn rejected the arguments ({rc})")
    return out, idx, meta, tim


def render_nn_multi(mode, x, y, z, peaks, fwhms, vx, vy, vz, l, b, nx, ny, nz, nl, lambda_width, lambda0,
                    dn=False, rows=None):
    """Several emission cubes (lists of per-point peak and fwhm arrays) of one dataset, one NN query
    per sample for all. Returns (images [nlines, ny*nx*max(nl,1)], idx [ny*nx*nz], tie flags
    [ny*nx*nz], meta[10], timing[3]); rows=(r0, r1) renders only those image rows."""
    x, y, z, vx, vy, vz = map(_pf, (x, y, z, vx, vy, vz))
    L = len(peaks)
    pk = _pd(np.concatenate([np.asarray(p, np.float64) for p in peaks]))
    fw = _pd(np.concatenate([np.asarray(f, np.float64) for f in fwhms]))
    out = np.zeros(L * nx * ny * max(nl, 1))
    idx = np.zeros(nx * ny * nz, np.int32)
    tie = np.zeros(nx * ny * nz, np.uint8)
    meta = np.zeros(10)
    tim = np.zeros(3)
    rc = lib().orc_render_nn_multi(mode, x.size, P(x, ctypes.c_float), P(y, ctypes.c_float), P(z, ctypes.c_float), L,
                                   P(pk, ctypes.c_double), P(fw, ctypes.c_double), P(vx, ctypes.c_float),
                                   P(vy, ctypes.c_float), P(vz, ctypes.c_float), l, b, nx, ny, nz, nl, lambda_width,
                                   lambda0, 1 if dn else 0, P(out, ctypes.c_double), P(idx, ctypes.c_int32),
                                   P(tie, ctypes.c_uint8), P(meta, ctypes.c_double), P(tim, ctypes.c_double),
                                   rows[0] if rows else 0, rows[1] if rows else 0)
    if rc != 0:
        raise ValueError(f"oracle render_nn_multi rejected the arguments ({rc})")
    return out.reshape(L, -1), idx, tie, meta, tim


def render_thomson(mode, x, y, z, totn, poln, l, b, nx, ny, nz, want_idx=False, rows=None, want_tie=False):
    """Returns (image [ny*nx], idx [ny*nx*nz] or None, meta[10], timing[3]); with want_tie also the
    per-sample tie flags. rows=(r0, r1) renders only those image rows (the rest stay 0)."""
    x, y, z = map(_pf, (x, y, z))
    totn, poln = _pd(totn), _pd(poln)
    out = np.zeros(nx * ny)
    idx = np.zeros(nx * ny * nz, np.int32) if want_idx else None
    tie = np.zeros(nx * ny * nz, np.uint8) if want_tie else None
    meta = np.zeros(10)
    tim = np.zeros(3)
    rc = lib().orc_render_thomson(mode, x.size, P(x, ctypes.c_float), P(y, ctypes.c_float), P(z, ctypes.c_float),
                                  P(totn, ctypes.c_double), P(poln, ctypes.c_double), l, b, nx, ny, nz,
                                  P(out, ctypes.c_double), P(idx, ctypes.c_int32) if want_idx else None,
                                  P(tie, ctypes.c_uint8) if want_tie else None, P(meta, ctypes.c_double),
                                  P(tim, ctypes.c_double), rows[0] if rows else 0, rows[1] if rows else 0)
    if rc != 0:
        raise ValueError(f"oracle render_thomson rejected the arguments ({rc})")
    if want_tie:
        return out, idx, meta, tim, tie
    return out, idx, meta, tim


def render_projection(mode, x, y, z, peak, fwhm, vx, vy, vz, l, b, nx, ny, nz, nl, lambda_width, lambda0,
                      dn=False, want_pix=False):
    """The reference's projection method (fomo-projection.cpp:17-214), sums in point order.
    Returns (image float64 [ny*nx*max(nl,1)], pixel per point int32 or None, meta[10])."""
    x, y, z, vx, vy, vz = map(_pf, (x, y, z, vx, vy, vz))
    peak, fwhm = _pd(peak), _pd(fwhm)
    out = np.zeros(nx * ny * max(nl, 1))
    pix = np.zeros(x.size, np.int32) if want_pix else None
    meta = np.zeros(10)
    rc = lib().orc_render_projection(mode, x.size, P(x, ctypes.c_float), P(y, ctypes.c_float),
                                     P(z, ctypes.c_float), P(peak, ctypes.c_double), P(fwhm, ctypes.c_double),
                                     P(vx, ctypes.c_float), P(vy, ctypes.c_float), P(vz, ctypes.c_float), l, b, nx,
                                     ny, nz, nl, lambda_width, lambda0, 1 if dn else 0, P(out, ctypes.c_double),
                                     P(pix, ctypes.c_int32) if want_pix else None, P(meta, ctypes.c_double))
    if rc != 0:
        raise ValueError(f"oracle render_projection rejected the arguments ({rc})")
    return out, pix, meta


def num_threads():
    return lib().orc_num_threads()


def rel_l2(a, ref):
    a = np.asarray(a, np.float64)
    ref = np.asarray(ref, np.float64)
    den = np.linalg.norm(ref)
    return float(np.linalg.norm(a - ref) / den) if den > 0 else float(np.linalg.norm(a - ref))


def render_tet(x, y, z, peak, fwhm, vx, vy, vz, l, b, nx, ny, nz, nl, lambda_width, lambda0, dn=False,
               linear=False, tri=None, rows=None):
    """The reference's CGAL render (fomo-CGAL.cpp:75-307) restated in numpy, FP64, on an independent
    triangulation: scipy.spatial.Delaunay (Qhull) and its find_simplex for the point location
    (CGAL's locate, :194-206). Each sample takes the values of the nearest vertex of its tet
    (nearest_vertex_in_cell, :211-222), none outside the convex hull (:225-228); linear=True takes
    the barycentric interpolation instead (the "Linear" extension). Extents come from the NN
    oracle's rotation (the same formula, :118-121). tri: a scipy Delaunay of these points to reuse
    (the reference triangulates once per render() call, :33-73). rows=(r0, r1): only those image
    rows are rendered (the others stay 0 in the image and -1 in the vertices). Returns (image
    [ny*nx*max(nl,1)], vertex per sample [(iy*nx + jx)*nz + k], original index or -1, meta[10])."""
    import math

    from scipy.spatial import Delaunay

    x, y, z, vx, vy, vz = map(_pf, (x, y, z, vx, vy, vz))
    peak, fwhm = _pd(peak), _pd(fwhm)
    pts = np.stack([x, y, z], 1).astype(np.float64)
    if tri is None:
        tri = Delaunay(pts)
    minx, maxx, miny, maxy, minz, maxz = extents(x, y, z, l, b)
    r0, r1 = rows if rows else (0, ny)
    shape = (r1 - r0, nx, nz)
    px, py, pz = (a[r0:r1] for a in view_samples([minx, maxx, miny, maxy, minz, maxz], l, b, nx, ny, nz,
                                                 cgal=True))  # :187-192
    q = np.stack([px.ravel(), py.ravel(), pz.ravel()], 1)
    s = tri.find_simplex(q)
    inside = s >= 0
    verts = tri.simplices[np.where(inside, s, 0)]
    vlos_pt = los_velocity(vx, vy, vz, l, b)                                       # :110
    if not linear:
        d = np.stack([((pts[verts[:, m], 0] - q[:, 0]) ** 2 + (pts[verts[:, m], 1] - q[:, 1]) ** 2)
                      + (pts[verts[:, m], 2] - q[:, 2]) ** 2 for m in range(4)], 1)
        nv = verts[np.arange(len(q)), np.argmin(d, 1)]
        vert = np.where(inside, nv, -1).astype(np.int32)
        pk = np.where(inside, peak[nv], 0.0)
        fw = np.where(inside, fwhm[nv], 1.0)
        vl = np.where(inside, vlos_pt[nv], 0.0)
    else:
        T = tri.transform[np.where(inside, s, 0)]
        bc = np.einsum("nij,nj->ni", T[:, :3], q - T[:, 3])
        w = np.concatenate([bc, 1 - bc.sum(1, keepdims=True)], 1)
        vert = np.where(inside, verts[:, 0], -1).astype(np.int32)
        pk = np.where(inside, np.sum(w * peak[verts], 1), 0.0)
        fw = np.where(inside, np.sum(w * fwhm[verts], 1), 1.0)
        vl = np.where(inside, np.sum(w * vlos_pt[verts], 1), 0.0)
    pk, fw, vl = pk.reshape(shape), fw.reshape(shape), vl.reshape(shape)
    C = 2.99792458e8
    nlo = max(nl, 1)
    img = np.zeros((ny, nx, nlo))
    if nl > 1:
        wla = lambda_width * lambda0 / C
        for il in range(nl):
            lam = il / (nl - 1) * wla - wla / 2.0                                  # :237
            with np.errstate(over="ignore", invalid="ignore", divide="ignore"):
                t = pk * np.exp(-((lam - vl / C * lambda0) ** 2) / fw ** 2 * 4.0 * math.log(2.0))
            img[r0:r1, :, il] = np.where(pk != 0, t, 0.0).sum(2)
    else:
        img[r0:r1, :, 0] = pk.sum(2)
    pathlength = (maxz - minz) / (nz - 1) if nz > 1 else 1.0                      # :263-271
    if maxz == minz:
        pathlength = 1.0
    apix = 1.0
    if dn:
        dxf = np.float32((maxx - minx) / (nx - 1))
        dyf = np.float32((maxy - miny) / (ny - 1))
        apix = (float(dxf) / 0.715) * (float(dyf) / 0.715) * (math.pi / 180.0 / 3600.0) ** 2
    scale = pathlength * 1e8 * apix
    meta = np.array([minx, maxx, miny, maxy, minz, maxz, 0.0, pathlength, scale, apix])
    vfull = np.full((ny, nx, nz), -1, np.int32)
    vfull[r0:r1] = vert.reshape(shape)
    return (img * scale).reshape(-1), vfull.reshape(-1), meta
