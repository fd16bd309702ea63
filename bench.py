#!/usr/bin/env python3
"""Benchmark: LOS samples x lambda per second, batched vs plain, on one MI355X.

Workload (one "step"): the batched multi-view multi-line nearest-neighbour render of a
256^3 structured loop datacube (16.8M points): 64 views (l = k pi/32, k=0..31, x b in {pi/2, pi/3})
x 4 spectral lines (synthetic G(T) tables peaked at log T = 5.8, 6.2, 6.4, 6.9, sigma 0.15) of
256x256 images with 128 LOS samples and 60 lambda pixels over 0.2 A, all in one captured HIP graph.
W per step = 64*4*256*256*128*60 = 1.29e11 LOS-sample x lambda evaluations. Inputs (datacube,
cell list, emission cubes) are resident in HBM before the timed region; results stay on the device.

Also measured: the plain path (one render per view and line, each re-computing the emissivity,
result copied to the host, as FoMoObject::render does per call; the spatial index stays resident),
the index build time, per-kernel durations from HIP events around each kernel of an eager pass of
the same work (roofline), and the CPU oracle (reference float pipeline) on a bounded sample.

Prints ONE JSON line (see the README of the harness contract).
"""
import argparse
import json
import os
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

C = 2.99792458e8
PEAK_HBM = 8.0e12  # B/s, MI355X HBM3E


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--m", type=int, default=256, help="datacube points per axis")
    ap.add_argument("--nviews", type=int, default=64)
    ap.add_argument("--nlines", type=int, default=4)
    ap.add_argument("--nx", type=int, default=256)
    ap.add_argument("--nz", type=int, default=128)
    ap.add_argument("--nl", type=int, default=60)
    ap.add_argument("--plain-renders", type=int, default=8)
    ap.add_argument("--cpu-rows", type=int, default=16, help="image rows in the CPU baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cell-scale", type=float, default=0.0)
    ap.add_argument("--workload", choices=["batched", "thomson"], default="batched")
    ap.add_argument("--corona-m", type=int, default=645, help="corona grid points per axis (thomson workload)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1 "
                                   "(multi-GPU sharding is described in DESIGN.md, not implemented)"}))
        sys.exit(2)

    import fomo_amd as F
    F.init(0)
    dev, hbm, cus = F.device_info()
    log(f"device {dev}, {hbm / 2**30:.0f} GiB HBM, {cus} CUs")
    if args.workload == "thomson":
        return thomson(F, args)

    m, nx, ny, nz, nl = args.m, args.nx, args.nx, args.nz, args.nl
    t0 = time.time()
    d = F.gen_loop_cube(m, seed=2, t_in=1.0e6, t_out=2.0e6)
    log(f"generated {m}^3 datacube in {time.time() - t0:.1f}s")
    tmp = tempfile.mkdtemp()
    peaks = [5.8, 6.2, 6.4, 6.9][: args.nlines]
    paths = [F.gen_goft_table(os.path.join(tmp, f"line{i}.dat"), ion="fe_9", lambda0=171.0, atweight=56.0,
                              peak_logt=p) for i, p in enumerate(peaks)]
    tabs = [F.Table.from_file(p) for p in paths]
    lambda0, atw = 171.0, 56.0
    width = 0.2 * C / lambda0

    t0 = time.time()
    ds = F.Dataset(d["x"], d["y"], d["z"], [d["rho"], d["T"], d["vx"], d["vy"], d["vz"]], cell_scale=args.cell_scale)
    n, dims, maxc = ds.info()
    log(f"dataset upload + index: {time.time() - t0:.2f}s, grid {dims}, max cell occupancy {maxc}")
    ems = [F.Emission.chianti(ds, t, lambda0, atw) for t in tabs]

    nv = args.nviews
    ls = [np.pi * (k % (nv // 2)) / (nv // 2) for k in range(nv)]
    bs = [np.pi / 2 if k < nv // 2 else np.pi / 3 for k in range(nv)]
    W_step = float(nv * len(ems)) * nx * ny * nz * nl
    S_view = nx * ny * nz
    R = nx * ny

    t0 = time.time()
    bt = F.Batch(ems, ls, bs, nx, ny, nz, nl, width)
    log(f"batch graph capture: {time.time() - t0:.2f}s")

    # ---- batched: warmup + timed steps --------------------------------------------------------
    for _ in range(args.warmup):
        bt.run()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        bt.launch()
    bt.sync()
    dt = time.perf_counter() - t0
    value = W_step * args.steps / dt
    ms_step = dt / args.steps * 1e3
    log(f"batched: {ms_step:.2f} ms/step, {value:.3e} W/s")

    # ---- device copy probe (plain float4 copy of 1 GiB: 1 GiB read + 1 GiB written) --------------
    copy_ms = F.copy_probe(1 << 30, 10)
    copy_gbps = 2 * (1 << 30) / (copy_ms * 1e-3) / 1e9
    log(f"copy probe: {copy_ms:.3f} ms per 1 GiB copy, {copy_gbps:.0f} GB/s read+write")

    # ---- per-kernel timing: HIP events around each launch of an eager pass of the same work ----
    F.profile_reset()
    bt.profile()
    bt.profile()
    stats = {k: F.profile_stats(k) for k in ("k_view_rotate_nn", "k_nn", "k_nn_rays", "k_los_spectral")}
    log("per-kernel stats (count, total ms): " + json.dumps(stats))
    bytes_per = {
        # algorithmic bytes per launch
        "k_nn": S_view * 4 + n * 16 + (dims[0] * dims[1] * dims[2] + 1) * 4,
        "k_los_spectral": S_view * 16 + R * nl * 4,
        "k_view_rotate_nn": n * 36,
    }
    dom = max(stats, key=lambda k: stats[k][1])
    cnt, tot = stats[dom]
    roof = None
    if cnt > 0:
        avg_s = tot / cnt / 1e3
        ach = bytes_per[dom] / avg_s
        roof = {"kernel": dom, "bound": "hbm", "achieved": ach / 1e9, "peak": PEAK_HBM / 1e9, "unit": "GB/s",
                "frac": ach / PEAK_HBM, "traffic": None, "avg_launch_ms": avg_s * 1e3,
                "algorithmic_bytes_per_launch": bytes_per[dom],
                "share_of_step": {k: stats[k][1] / max(1e-9, sum(v[1] for v in stats.values())) for k in stats}}

    # ---- plain: one render() per (view, line), as FoMoObject::render does per call: emissivity
    #      recomputed, one view rendered, result copied to the host; the index stays resident
    nplain = min(args.plain_renders, nv * len(ems))
    t0 = time.perf_counter()
    for r in range(nplain):
        v, g = r // len(ems), r % len(ems)
        e = F.Emission.chianti(ds, tabs[g], lambda0, atw)
        img, _ = F.render_nn(e, ls[v], bs[v], nx, ny, nz, nl, width)
        e.close()
    dtp = time.perf_counter() - t0
    W_render = float(nx) * ny * nz * nl
    plain_value = W_render * nplain / dtp
    log(f"plain: {dtp / nplain * 1e3:.2f} ms/render, {plain_value:.3e} W/s")
    # index (cell list + candidate lookup grid) build time, once per dataset
    t0 = time.perf_counter()
    ds.reindex(args.cell_scale)
    t_index = time.perf_counter() - t0
    linfo = ds.lookup_info()
    log(f"index rebuild: {t_index * 1e3:.1f} ms, lookup {linfo}")

    # ---- CPU baseline: reference float pipeline (oracle), 1 view x 1 line, row sample ---------
    cpu = None
    if not args.no_cpu:
        from oracle import oracle as O
        tab = O.read_table(paths[0])
        t0 = time.perf_counter()
        pk, fw = O.emission(0, d["rho"], d["T"], tab)
        t_em = time.perf_counter() - t0
        rows = (ny // 2 - args.cpu_rows // 2, ny // 2 + args.cpu_rows // 2)
        _, _, _, tim = O.render_nn(0, d["x"], d["y"], d["z"], pk, fw, d["vx"], d["vy"], d["vz"], ls[1], bs[1], nx, ny,
                                   nz, nl, width, lambda0, rows=rows)
        t_view = tim[0] + tim[1] + tim[2] * ny / (rows[1] - rows[0])
        cpu_value = W_render / (t_view + t_em)
        cpu = {"value": cpu_value, "unit": "LOS samples*lambda/s", "cores": O.num_threads(), "kind": "port",
               "sample": f"1 view (l=pi/32, b=pi/2) x 1 line of the same workload; rotate {tim[0]:.2f}s + kd-tree "
                         f"{tim[1]:.2f}s + ray loop over {rows[1] - rows[0]} of {ny} image rows {tim[2]:.2f}s "
                         f"(scaled to all rows) + emissivity {t_em:.2f}s; reference float pipeline (oracle mode 0)"}
        log(f"cpu: {cpu_value:.3e} W/s on {O.num_threads()} threads")

    out = {
        "metric": "LOS samples x lambda per second, batched render (64 views x 4 lines in one HIP graph)",
        "value": value,
        "unit": "LOS samples*lambda/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"batched NN render, {m}^3 loop cube, {nv} views x {len(ems)} lines, "
                               f"{nx}x{ny} image, nz {nz}, nl {nl}",
                   "points": n, "views": nv, "lines": len(ems), "image": [nx, ny], "z_pixels": nz,
                   "lambda_pixels": nl, "W_per_step": W_step},
        "plain": {"value": plain_value, "ms_per_render": dtp / nplain * 1e3, "renders_timed": nplain,
                  "what": "one render per view and line: emissivity + render + copy to host, index resident"},
        "index_build_ms": t_index * 1e3,
        "copy_probe_GBps": copy_gbps,
        "lookup": linfo,
        "batched_vs_plain": value / plain_value,
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    bt.close()
    for e in ems:
        e.close()
    ds.close()
    F.shutdown()


def thomson(F, args):
    """Thomson-scattering white-light render (pB): corona on 1 < r <= 3 Rsun, 16 views at b = pi/2
    (l = 2 pi k / 16), 1024^2 images with 512 LOS samples. W = LOS samples (nl = 1)."""
    t0 = time.time()
    d = F.gen_corona(args.corona_m, seed=3, sigma=0.1)
    n = d["x"].size
    log(f"corona: {n} points (grid {args.corona_m}^3 inside 1 < r <= 3 Rsun) in {time.time() - t0:.1f}s")
    t0 = time.time()
    ds = F.Dataset(d["x"], d["y"], d["z"], [d["rho"]], cell_scale=args.cell_scale)
    em = F.Emission.thomson(ds)
    log(f"upload + index + emission: {time.time() - t0:.2f}s, lookup {ds.lookup_info()}")
    nx = ny = args.nx if args.nx != 256 else 1024
    nz = args.nz if args.nz != 128 else 512
    nv = 16
    ls = [2 * np.pi * k / nv for k in range(nv)]
    b = np.pi / 2
    for _ in range(max(1, args.warmup)):
        t1 = time.perf_counter()
        F.render_thomson(em, ls[0], b, nx, ny, nz)
        log(f"warmup view: {(time.perf_counter() - t1) * 1e3:.1f} ms")
    t0 = time.perf_counter()
    for k in range(nv):
        t1 = time.perf_counter()
        img, meta = F.render_thomson(em, ls[k], b, nx, ny, nz)
        log(f"view {k}: {(time.perf_counter() - t1) * 1e3:.1f} ms")
    dt = time.perf_counter() - t0
    W = float(nv) * nx * ny * nz
    F.profile_reset()
    F.profile_enable(True)
    F.render_thomson(em, ls[1], b, nx, ny, nz)
    F.profile_enable(False)
    st = {k: F.profile_stats(k) for k in ("k_view_rotate_thomson", "k_nn", "k_nn_rays", "k_los_thomson")}
    log("per-kernel (count, ms): " + json.dumps(st))
    out = {"metric": "LOS samples per second, Thomson pB render (16 views)", "value": W / dt,
           "unit": "LOS samples/s", "n_gpus": 1, "steps": nv, "warmup": args.warmup, "ms_per_step": dt / nv * 1e3,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
           "config": {"workload": f"Thomson pB, corona {n} points, 16 views b=pi/2, {nx}x{ny} image, nz {nz}"},
           "kernels_ms": {k: (v[1] / v[0] if v[0] else None) for k, v in st.items()}}
    print(json.dumps(out), flush=True)
    em.close()
    ds.close()
    F.shutdown()


if __name__ == "__main__":
    main()
