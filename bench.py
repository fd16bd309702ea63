#!/usr/bin/env python3
"""Benchmark: LOS samples x lambda per second, batched vs plain, on one MI355X.

Workload (one "step"): the batched multi-view multi-line nearest-neighbour render of a
256^3 structured loop datacube (16.8M points): 64 views (l = k pi/32, k=0..31, x b in {pi/2, pi/3})
x 4 spectral lines (synthetic G(T) tables peaked at log T = 5.8, 6.2, 6.4, 6.9, sigma 0.15) of
256x256 images with 128 LOS samples and 60 lambda pixels over 0.2 A, all in one captured HIP graph.
W per step = 64*4*256*256*128*60 = 1.29e11 LOS-sample x lambda evaluations. Inputs (datacube,
cell list, emission cubes) are resident in HBM before the timed region; results stay on the device.

Also measured, beside the headline:
- a self-check: one oblique view of the batch (all 4 lines) on a band of rows against the FP64
  oracle (NN choices bit for bit with exact-distance ties counted apart, images within 1e-5);
- the batched step plus the copy of all its images to the host;
- the plain path (one render per view and line, each re-computing the emissivity, result copied
  to the host, as FoMoObject::render does per call), with the spatial index resident and with it
  rebuilt before every render (the reference rebuilds its R-tree per view);
- the index build time and per-kernel durations from HIP events around each kernel of an eager
  pass of the same work (roofline);
- serving: the mixed request stream of the serving configuration over resident cubes (this one
  among them), saturated and at 0.5, 0.8 and 0.95 of that rate, p50/p99 latency;
- the CPU oracle (reference float pipeline) on a bounded sample, median of 3 runs.

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
PEAK_FP64 = 78.6e12  # FLOP/s, MI355X dense FP64 (vector and matrix peaks coincide)
# Algorithmic FP64 work of the reference's integration per LOS sample x lambda pixel (W): the
# Gaussian term and its accumulation, about 8 FLOP besides the exp (fomo-nearestneighbour.cpp:200)
FLOP_PER_W = 8.0
# HBM bytes per view (read + write) of the default workload's fused pass, from two rocprofv3 --pmc
# runs (FETCH_SIZE x 2.000, WRITE_SIZE x 1.000, the factors a 1 GiB copy kernel gives) divided by
# the views of a launch, per index kind of the cube: profiles/batched_render_256cube_counters.md
MEASURED_TRAFFIC = {"scattered": {"k_nn_los": 1.177e9}, "rectilinear": {"k_nn_los": 5.667e8}}
# FP64 work the fused pass executes per view (FLOP: 2 per FMA, 1 per MUL/ADD/transcendental,
# from SQ_INSTS_VALU_{FMA,MUL,ADD,TRANS}_F64 x 64 lanes) on the rectilinear default cube, beside the
# reference formula's count: profiles/batched_render_256cube_counters.md
EXECUTED_FP64 = {"rectilinear": {"k_nn_los": 1.509e9}}


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
    ap.add_argument("--autotune", action="store_true",
                    help="let the autotuner pick the index configuration first (for C2 the default one is as fast "
                         "inside the batch, where the NN pass shares the GPU with the integration)")
    ap.add_argument("--workload", choices=["batched", "thomson", "serve", "tet", "projection"], default="batched")
    ap.add_argument("--requests", type=int, default=10000, help="requests per arrival rate (serve workload)")
    ap.add_argument("--serve-requests", type=int, default=10000,
                    help="requests per arrival rate in the serving phase of the default run (0: no serving phase)")
    ap.add_argument("--corona-m", type=int, default=645, help="corona grid points per axis (thomson workload)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1 "
                                   "(sharding over several GPUs is not implemented)"}))
        sys.exit(2)

    import fomo_amd as F
    F.init(0)
    dev, hbm, cus = F.device_info()
    log(f"device {dev}, {hbm / 2**30:.0f} GiB HBM, {cus} CUs")
    if args.workload == "thomson":
        return thomson(F, args)
    if args.workload == "serve":
        return serve(F, args)
    if args.workload == "tet":
        return tet(F, args)
    if args.workload == "projection":
        return projection(F, args)

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
    rect = ds.rect_axes()
    index_kind = "rectilinear" if rect else "scattered"
    log(f"dataset upload + index: {time.time() - t0:.2f}s, {index_kind} index "
        f"({'axes ' + str(rect) if rect else f'grid {dims}, max cell occupancy {maxc}'})")
    ems = [F.Emission.chianti(ds, t, lambda0, atw) for t in tabs]

    nv = args.nviews
    ls = [np.pi * (k % (nv // 2)) / (nv // 2) for k in range(nv)]
    bs = [np.pi / 2 if k < nv // 2 else np.pi / 3 for k in range(nv)]
    # index configuration picked by the autotuner on one view of the workload (setup, not timed)
    tune = None
    if args.autotune and args.cell_scale <= 0:
        t0 = time.time()
        tune = F.autotune(ems, ls[min(1, nv - 1)], bs[min(1, nv - 1)], nx, ny, nz, reps=5)
        log(f"autotune: cell scale {tune['cell_scale']}, lookup subdivision {tune['subdiv']}: NN pass "
            f"{tune['ms_before']:.3f} -> {tune['ms']:.3f} ms ({time.time() - t0:.1f}s)")
        n, dims, maxc = ds.info()
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

    # ---- HBM probes: a streaming copy of 1 GiB (1 GiB read + 1 GiB written) and 16 M random
    #      128-byte lines of a 4 GiB buffer (the NN passes' access pattern) ------------------------
    copy_ms = F.copy_probe(1 << 30, 10)
    copy_gbps = 2 * (1 << 30) / (copy_ms * 1e-3) / 1e9
    gather_ms = F.gather_probe(4 << 30, 16 << 20, 10)
    gather_gbps = (16 << 20) * 128 / (gather_ms * 1e-3) / 1e9
    log(f"HBM probes: copy {copy_gbps:.0f} GB/s read+write, random 128-byte lines {gather_gbps:.0f} GB/s")

    # ---- per-kernel timing: HIP events around each launch of an eager pass of the same work ----
    F.profile_reset()
    bt.profile()
    bt.profile()
    stats = {k: F.profile_stats(k) for k in ("k_view_extents_multi", "k_nn_los", "k_nn", "k_nn_far",
                                             "k_los_spectral")}
    stats = {k: v for k, v in stats.items() if v[0] > 0}
    log("per-kernel stats (count, total ms): " + json.dumps(stats))
    # views per launch: the batch runs every view's fused pass in one launch (two profile passes)
    vpl = {k: 1 for k in stats}
    if "k_nn_los" in stats:
        vpl["k_nn_los"] = max(1, round(2 * nv / stats["k_nn_los"][0]))
    bytes_per = {
        # algorithmic bytes per view
        "k_nn": S_view * 4 + n * 16 + (dims[0] * dims[1] * dims[2] + 1) * 4,
        # fused NN + integration: what the unfused pipeline moves for the same work, the NN pass
        # (S * 4 written, N * 12 read) plus each line's integration (S * 16 gathered, R * nl * 4
        # written); the fused pass never writes or gathers the per-sample values
        "k_nn_los": S_view * 4 + n * 12 + len(ems) * (S_view * 16 + R * nl * 4),
        "k_los_spectral": S_view * 16 + R * nl * 4,
        "k_view_extents_multi": n * 16,
    }
    dom = max(stats, key=lambda k: stats[k][1])
    cnt, tot = stats[dom]
    roof = None
    if cnt > 0:
        avg_s = tot / cnt / 1e3
        views = vpl.get(dom, 1)
        # HBM bytes per launch measured with rocprofv3 FETCH_SIZE / WRITE_SIZE (calibrated on a plain
        # copy kernel) for this exact workload: profiles/batched_render_256cube_counters.md
        default_cfg = (m, nx, nz, nl, len(ems)) == (256, 256, 128, 60, 4)
        traffic = MEASURED_TRAFFIC[index_kind].get(dom) if default_cfg else None
        traffic = traffic * views if traffic else None
        share = {k: stats[k][1] / max(1e-9, sum(v[1] for v in stats.values())) for k in stats}
        if dom == "k_nn_los":
            # the fused pass is bound by the 128-byte lines its NN queries fetch at random (lookup
            # cell records, winners' records); its FP64 integration is reported beside, with the
            # algorithmic FLOPs of the reference's formula
            alg = bytes_per[dom] * views
            flops = FLOP_PER_W * S_view * nl * len(ems) * views
            exe = EXECUTED_FP64.get(index_kind, {}).get(dom) if default_cfg else None
            exe = exe * views if exe else None
            roof = {"kernel": dom, "bound": "hbm", "achieved": alg / avg_s / 1e9, "peak": PEAK_HBM / 1e9,
                    "unit": "GB/s", "frac": alg / avg_s / PEAK_HBM, "traffic": traffic,
                    "avg_launch_ms": avg_s * 1e3, "views_per_launch": views, "ms_per_view": avg_s * 1e3 / views,
                    "algorithmic_bytes_per_launch": alg, "algorithmic_bytes_per_view": bytes_per[dom],
                    "measured": {"GBps": traffic / avg_s / 1e9 if traffic else None,
                                 "frac_of_peak": traffic / avg_s / PEAK_HBM if traffic else None,
                                 "frac_of_copy_probe": traffic / avg_s / 1e9 / copy_gbps if traffic else None,
                                 "frac_of_random_line_probe": traffic / avg_s / 1e9 / gather_gbps if traffic
                                 else None},
                    "fp64": {"reference_formula_flops_per_launch": flops, "flop_per_W": FLOP_PER_W,
                             "reference_formula_TFLOPs": flops / avg_s / 1e12,
                             "executed_flops_per_launch": exe,
                             "executed_TFLOPs": exe / avg_s / 1e12 if exe else None,
                             "peak_TFLOPs": PEAK_FP64 / 1e12,
                             "frac": exe / avg_s / PEAK_FP64 if exe else None},
                    "share_of_step": share}
        else:
            ach = bytes_per[dom] * views / avg_s
            roof = {"kernel": dom, "bound": "hbm", "achieved": ach / 1e9, "peak": PEAK_HBM / 1e9, "unit": "GB/s",
                    "frac": ach / PEAK_HBM, "traffic": traffic, "avg_launch_ms": avg_s * 1e3,
                    "views_per_launch": views, "algorithmic_bytes_per_launch": bytes_per[dom] * views,
                    "share_of_step": share}

    # ---- self-check: one oblique view of the batch, all lines, on a band of rows, against the FP64
    #      oracle (NN choices of the same view bit-exact, ties counted; images within 1e-5)
    check = None
    if not args.no_cpu:
        check = self_check(F, d, paths, ds, bt, nv // 2 + 3, ls, bs, nx, ny, nz, nl, width, lambda0, args.cpu_rows)

    # ---- batched with the images copied to the host: the same step plus one read of every image
    # (the first read also faults in the fresh host buffer; the second reuses it, as a consumer
    # that keeps its output buffer does)
    cube = np.empty(bt.shape, np.float32)
    t0 = time.perf_counter()
    bt.read(cube)
    t_first = time.perf_counter() - t0
    t0 = time.perf_counter()
    bt.read(cube)
    t_copy = time.perf_counter() - t0
    del cube
    nbytes = float(nv * len(ems)) * R * nl * 4
    with_copy = {"ms_per_step": ms_step + t_copy * 1e3, "value": W_step / (dt / args.steps + t_copy),
                 "copy_ms": t_copy * 1e3, "copy_GBps": nbytes / t_copy / 1e9,
                 "first_copy_ms": t_first * 1e3, "bytes": nbytes,
                 "what": "batched step + device-to-host copy of all its images into a reused host buffer"}
    log(f"batched + copy of all {with_copy['bytes'] / 2**30:.1f} GiB of images to the host: "
        f"{with_copy['ms_per_step']:.1f} ms/step, {with_copy['value']:.3e} W/s")

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
    # the batch graph holds pointers into the index, so it goes before any re-index
    bt.close()
    # ---- plain with the spatial index rebuilt for every render, as the reference rebuilds its
    #      R-tree per view (fomo-nearestneighbour.cpp:93-100)
    nrb = 2
    t0 = time.perf_counter()
    for r in range(nrb):
        v, g = r // len(ems), r % len(ems)
        ds.reindex(tune["cell_scale"] if tune else args.cell_scale)
        e = F.Emission.chianti(ds, tabs[g], lambda0, atw)
        img, _ = F.render_nn(e, ls[v], bs[v], nx, ny, nz, nl, width)
        e.close()
    dtr = time.perf_counter() - t0
    plain_rebuild = {"value": W_render * nrb / dtr, "ms_per_render": dtr / nrb * 1e3, "renders_timed": nrb,
                     "what": "one render per view and line with the spatial index rebuilt first"}
    log(f"plain with index rebuild: {dtr / nrb * 1e3:.1f} ms/render")
    # index (cell list + candidate lookup grid) build time, once per dataset
    t0 = time.perf_counter()
    ds.reindex(tune["cell_scale"] if tune else args.cell_scale)
    t_index = time.perf_counter() - t0
    linfo = ds.lookup_info()
    log(f"index rebuild: {t_index * 1e3:.1f} ms, lookup {linfo}")

    # ---- the same batched step on the scattered index (cell list + candidate lookup grid), the one
    #      every dataset that is not a product grid uses; the headline runs on the index the cube
    #      gets by default
    scattered = None
    if rect:
        t0 = time.perf_counter()
        ds.set_rectilinear(False)
        t_scat = time.perf_counter() - t0
        bt2 = F.Batch(ems, ls, bs, nx, ny, nz, nl, width)
        for _ in range(2):
            bt2.run()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            bt2.launch()
        bt2.sync()
        dt2 = time.perf_counter() - t0
        F.profile_reset()
        bt2.profile()
        c2, tot2 = F.profile_stats("k_nn_los")
        bt2.close()
        scattered = {"ms_per_step": dt2 / args.steps * 1e3, "value": W_step * args.steps / dt2,
                     "index_build_ms": t_scat * 1e3, "k_nn_los_ms_per_view": tot2 / nv if c2 else None,
                     "lookup": ds.lookup_info(),
                     "what": "the batched step with the cube indexed as scattered points (its product-grid axes "
                             "not used)"}
        log(f"scattered index: {scattered['ms_per_step']:.2f} ms/step, index {t_scat * 1e3:.0f} ms, k_nn_los "
            f"{scattered['k_nn_los_ms_per_view']} ms per view")
        ds.set_rectilinear(True)

    # ---- serving: the request mix of the serving configuration (NN, CGAL/Linear tetra and Thomson
    #      requests over resident cubes, this 256^3 cube among them), Poisson arrivals at 0.5, 0.8
    #      and 0.95 x the saturated request rate; latency = submit -> image in the caller's buffer
    serving = None
    if args.serve_requests > 0:
        emi = F.Emission.chianti(ds, tabs[1 % len(tabs)], lambda0, atw, imaging=True)
        targets, objs = c4_targets(F, paths[1 % len(paths)], lambda0, atw,
                                   resident_256=(ems[1 % len(ems)], emi, f"{m}^3 loop (NN)"))
        rng = np.random.default_rng(11)
        serving = serve_measure(F, targets, width, rng, args.serve_requests, [0.5, 0.8, 0.95])
        for o in objs:
            o.close()
        emi.close()

    # ---- CPU baseline: reference float pipeline (oracle), 1 view x 1 line, row sample, 3 runs
    cpu = None
    if not args.no_cpu:
        from oracle import oracle as O
        tab = O.read_table(paths[0])
        rows = (ny // 2 - args.cpu_rows // 2, ny // 2 + args.cpu_rows // 2)
        runs = []
        for _ in range(3):
            t0 = time.perf_counter()
            pk, fw = O.emission(0, d["rho"], d["T"], tab)
            t_em = time.perf_counter() - t0
            _, _, _, tim = O.render_nn(0, d["x"], d["y"], d["z"], pk, fw, d["vx"], d["vy"], d["vz"], ls[1], bs[1], nx,
                                       ny, nz, nl, width, lambda0, rows=rows)
            runs.append((tim[0] + tim[1] + tim[2] * ny / (rows[1] - rows[0]) + t_em, tim, t_em))
        runs.sort(key=lambda r: r[0])
        t_view, tim, t_em = runs[1]
        cpu_value = W_render / t_view
        cpu = {"value": cpu_value, "unit": "LOS samples*lambda/s", "cores": O.num_threads(), "kind": "port",
               "cpu": cpu_model(),
               "sample": f"1 view (l=pi/32, b=pi/2) x 1 line of the same workload, median of 3 runs "
                         f"({', '.join(f'{r[0]:.2f}s' for r in runs)} per view); the median run: rotate {tim[0]:.2f}s "
                         f"+ kd-tree {tim[1]:.2f}s + ray loop over {rows[1] - rows[0]} of {ny} image rows "
                         f"{tim[2]:.2f}s (scaled to all rows) + emissivity {t_em:.2f}s; reference float pipeline "
                         f"(oracle mode 0)"}
        log(f"cpu: {cpu_value:.3e} W/s on {O.num_threads()} threads ({cpu['cpu']})")

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
        "batched_with_copy": with_copy,
        "plain": {"value": plain_value, "ms_per_render": dtp / nplain * 1e3, "renders_timed": nplain,
                  "what": "one render per view and line: emissivity + render + copy to host, index resident"},
        "plain_rebuild_index": plain_rebuild,
        "batched_vs_plain": value / plain_value,
        "batched_with_copy_vs_plain": with_copy["value"] / plain_value,
        "index_build_ms": t_index * 1e3,
        "copy_probe_GBps": copy_gbps,
        "random_line_probe_GBps": gather_gbps,
        "index": {"kind": index_kind, "axes": list(rect) if rect else None},
        "lookup": linfo,
        "scattered_index": scattered,
        "autotune": tune,
        "self_check": check,
        "roofline": roof,
        "cpu_baseline": cpu,
        "serving": serving,
    }
    print(json.dumps(out), flush=True)
    for e in ems:
        e.close()
    ds.close()
    F.shutdown()


def cpu_model():
    """CPU model name and the cores this process may use."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return f"{model}; {len(os.sched_getaffinity(0))} cores available"


def self_check(F, d, paths, ds, bt, v, ls, bs, nx, ny, nz, nl, width, lambda0, nrows):
    """One view of the batch (all its lines) on nrows image rows against the FP64 oracle: the
    relative L2 error of every line's band, and the NN choices of the same view (F.nn_indices)
    bit for bit, with the samples whose nearest distance is an exact tie counted apart."""
    from oracle import oracle as O
    t0 = time.perf_counter()
    rows = (ny // 2 - nrows // 2, ny // 2 + nrows // 2)
    nlines = bt.shape[1]
    em_o = [O.emission(1, d["rho"], d["T"], O.read_table(p)) for p in paths[:nlines]]
    imgs, ridx, tie, rmeta, _ = O.render_nn_multi(1, d["x"], d["y"], d["z"], [e[0] for e in em_o],
                                                  [e[1] for e in em_o], d["vx"], d["vy"], d["vz"], ls[v], bs[v], nx,
                                                  ny, nz, nl, width, lambda0, rows=rows)
    idx, _ = F.nn_indices(ds, ls[v], bs[v], nx, ny, nz)
    gi, oi = idx.reshape(ny, -1)[rows[0]:rows[1]], ridx.reshape(ny, -1)[rows[0]:rows[1]]
    ti = tie.reshape(ny, -1)[rows[0]:rows[1]].astype(bool)
    errs = []
    meta_ok = True
    for g in range(nlines):
        img, meta = bt.read_view(v, g)
        meta_ok = meta_ok and bool(np.array_equal(meta[:7], rmeta[:7]))
        errs.append(O.rel_l2(img.reshape(ny, -1)[rows[0]:rows[1]], imgs[g].reshape(ny, -1)[rows[0]:rows[1]]))
    res = {"view": v, "l": ls[v], "b": bs[v], "rows": list(rows), "lines": nlines, "samples": int(gi.size),
           "nn_mismatch": int((gi != oi).sum()), "nn_mismatch_non_tie": int(((gi != oi) & ~ti).sum()),
           "tie_samples": int(ti.sum()), "empty_box_samples": int((oi < 0).sum()),
           "rel_l2_per_line": errs, "rel_l2_max": max(errs), "tolerance": 1e-5, "meta_equal": meta_ok,
           "ok": bool((gi == oi).all() and max(errs) <= 1e-5 and meta_ok), "seconds": time.perf_counter() - t0}
    log(f"self-check: view {v}, rows {rows}: NN mismatches {res['nn_mismatch']} ({res['tie_samples']} tie samples), "
        f"max rel L2 {res['rel_l2_max']:.2e}, ok={res['ok']}")
    return res


# Request mix of the serving workload: image sizes x LOS samples x lambda pixels drawn uniformly.
C4_RES = (64, 128, 256, 512)
C4_NZ = (64, 128, 256)
C4_NL = (1, 30, 100)


def draw_requests(F, targets, width, rng, count):
    """targets: dicts with "spec" (spectroscopic emission), "imag" (imaging emission, nl = 1) or
    "thomson" (Thomson emission), a "weight" and a "name"; "tet": True sends the target's requests
    through the tetra path. Views l ~ U(0, 2 pi), b ~ U(0.1, pi/2)."""
    w = np.array([t["weight"] for t in targets], float)
    w /= w.sum()
    reqs = []
    for _ in range(count):
        t = targets[rng.choice(len(targets), p=w)]
        nx = int(rng.choice(C4_RES))
        nz = int(rng.choice(C4_NZ))
        nl = int(rng.choice(C4_NL))
        l, b = float(rng.uniform(0, 2 * np.pi)), float(rng.uniform(0.1, np.pi / 2))
        method = F.METHOD_NN
        if "tet" in t:  # tetra path: the CGAL method or its linear variant, half each
            method = F.METHOD_TET if rng.random() < 0.5 else F.METHOD_LINEAR
        if "thomson" in t:
            reqs.append((t["thomson"], F.METHOD_THOMSON, l, b, nx, nx, nz, 1, 0.0))
        elif nl == 1:
            reqs.append((t["imag"], method, l, b, nx, nx, nz, 1, 0.0))
        else:
            reqs.append((t["spec"], method, l, b, nx, nx, nz, nl, width))
    return reqs


def req_work(r):
    return float(r[4]) * r[5] * r[6] * max(r[7], 1)


class _Buffers:
    """Reusable float32 output buffers by size, so that a long request stream holds only the images
    of the requests that are still outstanding."""

    def __init__(self):
        self.free = {}

    def get(self, n):
        lst = self.free.get(n)
        return lst.pop() if lst else np.empty(n, np.float32)

    def put(self, a):
        self.free.setdefault(a.size, []).append(a)


def _out_size(r):
    return r[4] * r[5] * max(r[7], 1)


def serve_measure(F, targets, width, rng, count, loads, device_bytes=24 << 30, host_bytes=2 << 30,
                  max_inflight=8, nstreams=4, window=256):
    """Saturated throughput (requests kept `window` deep in the queue), the same requests one
    synchronous render at a time, then open-loop Poisson arrivals at each load x the saturated rate:
    p50/p99 latency from submit to the image in the caller's buffer."""
    import collections
    sv = F.Server(device_bytes=device_bytes, host_bytes=host_bytes, max_inflight=max_inflight, nstreams=nstreams)
    bufs = _Buffers()

    def submit(r):
        out = bufs.get(_out_size(r))
        return sv.submit(*r[:8], lambda_width=r[8], out=out)

    def finish(t, timeout=-1.0):
        img, _ = sv.result(t, timeout)
        bufs.put(img)

    try:
        warm = draw_requests(F, targets, width, rng, 8)
        for t in [submit(r) for r in warm]:
            finish(t, 300)
        reqs = draw_requests(F, targets, width, rng, count)
        W = sum(req_work(r) for r in reqs)
        q = collections.deque()
        t0 = time.perf_counter()
        for r in reqs:
            if len(q) >= window:
                finish(q.popleft(), 600)
            q.append(submit(r))
        while q:
            finish(q.popleft(), 600)
        dt_sat = time.perf_counter() - t0
        sat_rps = count / dt_sat
        log(f"serving saturated: {count} requests in {dt_sat:.2f}s, {sat_rps:.1f} req/s, {W / dt_sat:.3e} W/s")
        t0 = time.perf_counter()
        for r in reqs:
            if r[1] == F.METHOD_THOMSON:
                F.render_thomson(r[0], *r[2:7])
            elif r[1] in (F.METHOD_TET, F.METHOD_LINEAR):
                F.render_tet(r[0], *r[2:8], r[8], linear=r[1] == F.METHOD_LINEAR)
            else:
                F.render_nn(r[0], *r[2:8], r[8])
        dt_seq = time.perf_counter() - t0
        log(f"one synchronous render at a time: {dt_seq:.2f}s, {W / dt_seq:.3e} W/s")
        runs = []
        for load in loads:
            rate = load * sat_rps
            reqs = draw_requests(F, targets, width, rng, count)
            arr = np.cumsum(rng.exponential(1.0 / rate, count))
            tks, q = [], collections.deque()
            t0 = time.perf_counter()
            for a, r in zip(arr, reqs):
                # hand back the buffers of finished requests, oldest first
                while q and sv.poll(q[0]) == 2:
                    finish(q.popleft(), 0)
                d = a - (time.perf_counter() - t0)
                if d > 0:
                    time.sleep(d)
                t = submit(r)
                tks.append(t)
                q.append(t)
            while q:
                finish(q.popleft(), 900)
            span = time.perf_counter() - t0
            lat = np.array([sv.times(t)[2] - sv.times(t)[0] for t in tks]) * 1e3
            Wl = sum(req_work(r) for r in reqs)
            run = {"load": load, "arrival_rate_req_s": rate, "requests": count, "p50_ms": float(np.percentile(lat, 50)),
                   "p99_ms": float(np.percentile(lat, 99)), "mean_ms": float(lat.mean()),
                   "max_ms": float(lat.max()), "throughput_W_s": Wl / span}
            log(f"serving at {load:.2f} x saturation ({rate:.1f} req/s, {count} requests): p50 {run['p50_ms']:.2f} ms, "
                f"p99 {run['p99_ms']:.2f} ms, {Wl / span:.3e} W/s")
            runs.append(run)
        st = sv.stats()
    finally:
        sv.close()
    return {"mix": " + ".join(t["name"] for t in targets) + "; image {64,128,256,512}^2 x nz {64,128,256} x "
                   "nl {1,30,100} uniform, l ~ U(0,2pi), b ~ U(0.1,pi/2)",
            "requests_per_rate": count, "saturation_window": window,
            "saturated_req_s": sat_rps, "saturated_W_s": W / dt_sat, "sequential_W_s": W / dt_seq,
            "served_vs_sequential": dt_seq / dt_sat, "max_inflight": max_inflight, "streams": nstreams,
            "device_budget_GiB": device_bytes / 2**30, "device_peak_GiB": st["device_peak"] / 2**30,
            "runs": runs,
            "p50_ms": runs[0]["p50_ms"] if len(runs) == 1 else None,
            "p99_ms": runs[0]["p99_ms"] if len(runs) == 1 else None}


def c4_targets(F, table_path, lambda0, atw, resident_256=None):
    """Resident objects of the serving configuration: 128^3 and 256^3 loop cubes (a spectroscopic
    and an imaging emission cube each) for nearest-neighbour requests, a triangulated jittered 128^3
    cube for tetra (CGAL and Linear) requests and a corona for Thomson requests. resident_256:
    (spectroscopic, imaging, name) of a 256^3 cube already resident, used instead of a new one.
    Returns (targets, objects to close)."""
    tab = F.Table.from_file(table_path)
    objs, targets = [tab], []
    for m, seed in ((128, 4), (256, 2)):
        if m == 256 and resident_256 is not None:
            es, ei, name = resident_256
            targets.append({"spec": es, "imag": ei, "weight": 0.35, "name": name})
            continue
        d = F.gen_loop_cube(m, seed=seed)
        ds = F.Dataset(d["x"], d["y"], d["z"], [d["rho"], d["T"], d["vx"], d["vy"], d["vz"]])
        es = F.Emission.chianti(ds, tab, lambda0, atw)
        ei = F.Emission.chianti(ds, tab, lambda0, atw, imaging=True)
        objs += [es, ei, ds]
        targets.append({"spec": es, "imag": ei, "weight": 0.35, "name": f"{m}^3 loop (NN)"})
    # linear-interpolation requests: the tetra path on a jittered 128^3 cube (triangulated once)
    d = F.gen_loop_cube(128, seed=5, jitter=0.3)
    dst = F.Dataset(d["x"], d["y"], d["z"], [d["rho"], d["T"], d["vx"], d["vy"], d["vz"]])
    ntets, tsec = dst.triangulate()
    log(f"triangulated the jittered 128^3 cube: {ntets} tets in {tsec:.1f}s")
    es = F.Emission.chianti(dst, tab, lambda0, atw)
    ei = F.Emission.chianti(dst, tab, lambda0, atw, imaging=True)
    objs += [es, ei, dst]
    targets.append({"spec": es, "imag": ei, "tet": True, "weight": 0.2,
                    "name": "jittered 128^3 loop (CGAL / Linear tetra path)"})
    c = F.gen_corona(128, seed=3, sigma=0.1)
    dsc = F.Dataset(c["x"], c["y"], c["z"], [c["rho"]])
    et = F.Emission.thomson(dsc)
    objs += [et, dsc]
    targets.append({"thomson": et, "weight": 0.10, "name": f"corona {c['x'].size} pts (Thomson)"})
    log("resident: " + ", ".join(t["name"] for t in targets))
    # close in reverse order of creation (emission cubes before their datasets)
    return targets, objs[::-1]


def serve(F, args):
    """Serving workload: the resident objects of c4_targets; requests drawn uniformly from the size
    mix, Poisson arrivals at 0.5, 0.8 and 0.95 x the saturated request rate."""
    tmp = tempfile.mkdtemp()
    path = F.gen_goft_table(os.path.join(tmp, "fe9.dat"), ion="fe_9", lambda0=171.0, atweight=56.0, peak_logt=6.2)
    lambda0, atw = 171.0, 56.0
    width = 0.2 * C / lambda0
    targets, objs = c4_targets(F, path, lambda0, atw)
    rng = np.random.default_rng(11)
    res = serve_measure(F, targets, width, rng, args.requests, [0.5, 0.8, 0.95])
    mid = [r for r in res["runs"] if r["load"] == 0.8][0]
    out = {"metric": "render latency p50/p99 under Poisson load (continuous batching server)",
           "value": res["saturated_W_s"], "unit": "LOS samples*lambda/s", "n_gpus": 1, "steps": args.requests,
           "warmup": 8, "ms_per_step": 1e3 / res["saturated_req_s"], "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f64", "data": "synthetic",
           "config": {"workload": "serving mix: " + res["mix"]},
           "p50_ms_at_0.8": mid["p50_ms"], "p99_ms_at_0.8": mid["p99_ms"], "serving": res}
    print(json.dumps(out), flush=True)
    for o in objs:
        o.close()
    F.shutdown()


def tet(F, args):
    """Tetra-path render at scale, linear interpolation in the containing Delaunay tet (the Linear
    method; the reference's CGAL method takes the tet's nearest vertex, timed beside it): a 256^3
    loop cube jittered by +-0.3 cell (16.8M points, unstructured Delaunay), 2.5 MK corona, 10 %
    sausage mode; one view (l=0, b=pi/2) of 512x512 pixels, 256 LOS samples, 100 lambda pixels over
    0.2 A. The Delaunay triangulation (host, once per dataset) is timed separately; a step is one
    render."""
    m = args.m
    nx = ny = args.nx if args.nx != 256 else 512
    nz = args.nz if args.nz != 128 else 256
    nl = args.nl if args.nl != 60 else 100
    t0 = time.time()
    d = F.gen_loop_cube(m, seed=1, t_in=1.0e6, t_out=2.5e6, jitter=0.3, sausage=0.1)
    log(f"generated {m}^3 jittered loop cube in {time.time() - t0:.1f}s")
    tmp = tempfile.mkdtemp()
    path = F.gen_goft_table(os.path.join(tmp, "fe9.dat"), ion="fe_9", lambda0=171.0, atweight=56.0, peak_logt=5.9)
    tab = F.Table.from_file(path)
    lambda0, atw = 171.0, 56.0
    width = 0.2 * C / lambda0
    ds = F.Dataset(d["x"], d["y"], d["z"], [d["rho"], d["T"], d["vx"], d["vy"], d["vz"]])
    t0 = time.time()
    os.environ["FOMO_PROGRESS"] = "1"  # the builder reports every 1M inserted points
    ntets, tbuild = ds.triangulate()
    log(f"Delaunay: {ntets} tets in {tbuild:.1f}s (host build + upload), {time.time() - t0:.1f}s wall")
    em = F.Emission.chianti(ds, tab, lambda0, atw)
    l, b = 0.0, np.pi / 2
    for _ in range(max(1, args.warmup)):
        F.render_tet(em, l, b, nx, ny, nz, nl, width, linear=True)
    steps = max(1, args.steps)
    t0 = time.perf_counter()
    for _ in range(steps):
        img, meta = F.render_tet(em, l, b, nx, ny, nz, nl, width, linear=True)
    dt = (time.perf_counter() - t0) / steps
    W = float(nx) * ny * nz * nl
    F.profile_reset()
    F.profile_enable(True)
    F.render_tet(em, l, b, nx, ny, nz, nl, width, linear=True)
    F.profile_enable(False)
    st = {k: F.profile_stats(k) for k in ("k_view_rotate_nn", "k_tet_walk_linear", "k_los_spectral")}
    log(f"linear tet render: {dt * 1e3:.1f} ms/view, {W / dt:.3e} W/s; kernels " + json.dumps(st))
    # the CGAL method (nearest vertex of the tet) on the same view
    t0 = time.perf_counter()
    for _ in range(steps):
        F.render_tet(em, l, b, nx, ny, nz, nl, width)
    dtc = (time.perf_counter() - t0) / steps
    log(f"nearest-vertex tet render: {dtc * 1e3:.1f} ms/view")
    # same view, nearest-neighbour method, for comparison
    t0 = time.perf_counter()
    F.render_nn(em, l, b, nx, ny, nz, nl, width)
    dtn = time.perf_counter() - t0
    # roofline of the walk (the profile scope holds the walk and the exact re-walk of the samples it
    # cannot place): per sample the containing tet's record (80 B), the 4 vertex records it
    # interpolates (4 x 32 B) and the per-sample outputs (12 B)
    cnt, tot = st["k_tet_walk_linear"]
    roof = None
    if cnt:
        avg_s = tot / cnt / 1e3
        byt = float(nx) * ny * nz * (80 + 4 * 32 + 12)
        roof = {"kernel": "k_tet_walk_linear", "bound": "hbm", "achieved": byt / avg_s / 1e9, "peak": PEAK_HBM / 1e9,
                "unit": "GB/s", "frac": byt / avg_s / PEAK_HBM, "traffic": None, "avg_launch_ms": avg_s * 1e3,
                "algorithmic_bytes_per_launch": byt}
    cpu = None
    if not args.no_cpu:
        # oracle (numpy FP64 on scipy's Qhull triangulation, find_simplex per sample, linear
        # interpolation) on a 64^3 cube of the same model, 128^2 x 64 samples x the same lambda grid;
        # the triangulation is built first and not timed (as on the GPU)
        from oracle import oracle as O
        from scipy.spatial import Delaunay
        ds_ = F.gen_loop_cube(64, seed=1, t_in=1.0e6, t_out=2.5e6, jitter=0.3, sausage=0.1)
        otab = O.read_table(path)
        pk, fw = O.emission(1, ds_["rho"], ds_["T"], otab)
        tri = Delaunay(np.stack([ds_["x"], ds_["y"], ds_["z"]], 1).astype(np.float64))
        cn, cz = 128, 64
        t0 = time.perf_counter()
        O.render_tet(ds_["x"], ds_["y"], ds_["z"], pk, fw, ds_["vx"], ds_["vy"], ds_["vz"], l, b, cn, cn, cz, nl,
                     width, lambda0, linear=True, tri=tri)
        tc = time.perf_counter() - t0
        Wc = float(cn) * cn * cz * nl
        cpu = {"value": Wc / tc, "unit": "LOS samples*lambda/s", "cores": 1, "kind": "port",
               "sample": f"64^3 jittered loop cube (262144 points), {cn}^2 x {cz} samples x {nl} lambda pixels, "
                         f"{tc:.2f}s: the oracle's linear tetra render (numpy FP64, scipy Qhull find_simplex), "
                         f"triangulation not timed"}
        log(f"cpu: {Wc / tc:.3e} W/s on 1 thread")
    out = {"metric": "LOS samples x lambda per second, tetra linear-interpolation render", "value": W / dt,
           "unit": "LOS samples*lambda/s", "n_gpus": 1, "steps": steps, "warmup": args.warmup,
           "ms_per_step": dt * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
           "data": "synthetic",
           "config": {"workload": f"tetra linear-interpolation render, {m}^3 jittered loop cube ({d['x'].size} "
                                  f"points), 1 view, {nx}x{ny} image, nz {nz}, nl {nl}"},
           "nearest_vertex_ms": dtc * 1e3,
           "delaunay": {"tets": ntets, "build_s": tbuild},
           "kernels_ms": {k: (v[1] / v[0] if v[0] else None) for k, v in st.items()},
           "nn_same_view_ms": dtn * 1e3, "roofline": roof, "cpu_baseline": cpu}
    print(json.dumps(out), flush=True)
    em.close()
    ds.close()
    tab.close()
    F.shutdown()


def projection(F, args):
    """Projection method at scale (fomo-projection.cpp:17-214): the 256^3 loop cube jittered by
    +-0.3 cell (16.8M points) of the tetra workload, one view (l=0.3, b=1.2) of 512x512 pixels,
    100 lambda pixels over 0.2 A (nz 256 only sets the path length). A step is one render, image
    copied to the host. W = points x lambda pixels. The per-(pixel, lambda) sums are FP64-bound:
    one exp per point and lambda pixel, in point order."""
    m = args.m
    nx = ny = args.nx if args.nx != 256 else 512
    nz = args.nz if args.nz != 128 else 256
    nl = args.nl if args.nl != 60 else 100
    d = F.gen_loop_cube(m, seed=1, t_in=1.0e6, t_out=2.5e6, jitter=0.3, sausage=0.1)
    n = d["x"].size
    tmp = tempfile.mkdtemp()
    path = F.gen_goft_table(os.path.join(tmp, "fe9.dat"), ion="fe_9", lambda0=171.0, atweight=56.0, peak_logt=5.9)
    tab = F.Table.from_file(path)
    lambda0, atw = 171.0, 56.0
    width = 0.2 * C / lambda0
    ds = F.Dataset(d["x"], d["y"], d["z"], [d["rho"], d["T"], d["vx"], d["vy"], d["vz"]])
    em = F.Emission.chianti(ds, tab, lambda0, atw)
    l, b = 0.3, 1.2
    for _ in range(max(1, args.warmup)):
        F.render_projection(em, l, b, nx, ny, nz, nl, width)
    steps = max(1, args.steps)
    t0 = time.perf_counter()
    for _ in range(steps):
        img, meta = F.render_projection(em, l, b, nx, ny, nz, nl, width)
    dt = (time.perf_counter() - t0) / steps
    W = float(n) * nl
    F.profile_reset()
    F.profile_enable(True)
    F.render_projection(em, l, b, nx, ny, nz, nl, width)
    F.profile_enable(False)
    st = {k: F.profile_stats(k) for k in ("k_view_rotate_nn", "k_proj_keys", "proj_sort", "k_proj_spectral")}
    log(f"projection render: {dt * 1e3:.1f} ms/view, {W / dt:.3e} point*lambda/s; kernels " + json.dumps(st))
    cnt, tot = st["k_proj_spectral"]
    roof = None
    if cnt:
        avg_s = tot / cnt / 1e3
        flops = FLOP_PER_W * W
        roof = {"kernel": "k_proj_spectral", "bound": "mfma", "what": "FP64 (vector and matrix peaks coincide)",
                "achieved": flops / avg_s / 1e12, "peak": PEAK_FP64 / 1e12, "unit": "TFLOP/s",
                "frac": flops / avg_s / PEAK_FP64, "traffic": None, "avg_launch_ms": avg_s * 1e3,
                "algorithmic_flops_per_launch": flops, "flop_per_W": FLOP_PER_W}
    cpu = None
    if not args.no_cpu:
        # oracle (reference float pipeline, sums in point order, one thread) on every 64th point
        from oracle import oracle as O
        otab = O.read_table(path)
        sel = slice(0, n, 64)
        pk, fw = O.emission(0, d["rho"][sel], d["T"][sel], otab)
        t0 = time.perf_counter()
        O.render_projection(0, d["x"][sel], d["y"][sel], d["z"][sel], pk, fw, d["vx"][sel], d["vy"][sel],
                            d["vz"][sel], l, b, nx, ny, nz, nl, width, lambda0)
        tc = time.perf_counter() - t0
        ns = d["x"][sel].size
        cpu = {"value": ns * nl / tc, "unit": "point*lambda/s", "cores": 1, "kind": "port",
               "sample": f"every 64th point ({ns} points) of the same view and spectral grid, {tc:.2f}s; the "
                         f"oracle's projection (reference float pipeline, point-order sums, one thread)"}
        log(f"cpu: {ns * nl / tc:.3e} point*lambda/s on 1 thread")
    out = {"metric": "points x lambda pixels per second, projection render", "value": W / dt,
           "unit": "point*lambda/s", "n_gpus": 1, "steps": steps, "warmup": args.warmup, "ms_per_step": dt * 1e3,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
           "config": {"workload": f"projection render, {m}^3 jittered loop cube ({n} points), 1 view, "
                                  f"{nx}x{ny} image, nz {nz}, nl {nl}"},
           "kernels_ms": {k: (v[1] / v[0] if v[0] else None) for k, v in st.items()},
           "roofline": roof, "cpu_baseline": cpu}
    print(json.dumps(out), flush=True)
    em.close()
    ds.close()
    tab.close()
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
    st = {k: F.profile_stats(k) for k in ("k_view_rotate_thomson", "k_nn", "k_nn_far", "k_nn_deep", "k_los_thomson")}
    log("per-kernel (count, ms): " + json.dumps(st))
    S = float(nx) * ny * nz
    roof = None
    if st["k_nn"][0] and st["k_nn_far"][0]:
        # the NN pass (lookup + exact octree query, and its range walk, for the samples the lookup
        # cannot resolve): algorithmic bytes S*4 written + N*12 read per view; the integration:
        # 16 B per sample
        t_nn = sum(st[k][1] / st[k][0] for k in ("k_nn", "k_nn_far", "k_nn_deep") if st[k][0]) / 1e3
        bnn = S * 4 + n * 12.0
        roof = {"kernel": "k_nn + k_nn_far + k_nn_deep (NN pass)", "bound": "hbm", "achieved": bnn / t_nn / 1e9,
                "peak": PEAK_HBM / 1e9, "unit": "GB/s", "frac": bnn / t_nn / PEAK_HBM, "traffic": None,
                "avg_launch_ms": t_nn * 1e3, "algorithmic_bytes_per_launch": bnn,
                "note": "the exact octree query of far-field samples is a chain of dependent node loads"}
        if st["k_los_thomson"][0]:
            t_int = st["k_los_thomson"][1] / st["k_los_thomson"][0] / 1e3
            roof["integration"] = {"kernel": "k_los_thomson", "achieved": S * 16 / t_int / 1e9,
                                   "frac": S * 16 / t_int / PEAK_HBM, "avg_launch_ms": t_int * 1e3}
    cpu = None
    if not args.no_cpu:
        # oracle (C++, OpenMP, kd-tree) on a 160^3-grid corona, one view of 256^2 x 128 samples
        from oracle import oracle as O
        dc = F.gen_corona(160, seed=3, sigma=0.1)
        tn, pn = O.thomson_emission(0, dc["x"], dc["y"], dc["z"], dc["rho"])
        _, _, _, tim = O.render_thomson(0, dc["x"], dc["y"], dc["z"], tn, pn, ls[1], b, 256, 256, 128)
        tc = float(tim[0] + tim[1] + tim[2])
        cpu = {"value": 256 * 256 * 128 / tc, "unit": "LOS samples/s", "cores": O.num_threads(), "kind": "port",
               "sample": f"corona on a 160^3 grid ({dc['x'].size} points), one view of 256^2 x 128 samples, "
                         f"{tc:.2f}s (rotation {tim[0]:.2f}s, kd-tree {tim[1]:.2f}s, ray loop {tim[2]:.2f}s): the "
                         f"oracle's Thomson render (reference float pipeline)"}
        log(f"cpu: {256 * 256 * 128 / tc:.3e} samples/s on {O.num_threads()} threads")
    out = {"metric": "LOS samples per second, Thomson pB render (16 views)", "value": W / dt,
           "unit": "LOS samples/s", "n_gpus": 1, "steps": nv, "warmup": args.warmup, "ms_per_step": dt / nv * 1e3,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
           "config": {"workload": f"Thomson pB, corona {n} points, 16 views b=pi/2, {nx}x{ny} image, nz {nz}"},
           "kernels_ms": {k: (v[1] / v[0] if v[0] else None) for k, v in st.items()},
           "roofline": roof, "cpu_baseline": cpu}
    print(json.dumps(out), flush=True)
    em.close()
    ds.close()
    F.shutdown()


if __name__ == "__main__":
    main()
