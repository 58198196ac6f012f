#!/usr/bin/env python3
"""Throughput of the staggered-grid elastic solver on one MI355X.

Workload (default): the 512x512x384 four-layer + 5 % random medium with a
20-cell PML, dt 0.8 ms, FP32, the configuration the stated target is quoted on.
A step is one full time step (source, velocity sweep, stress sweep, station
gather) over the whole grid; inputs are resident in HBM before timing.

Prints one JSON line: Gpt-updates/s over the full grid (PML and halo
included), the roofline position of the dominant sweep (HIP-event timed over
the timed region, algorithmic bytes from the lean split-PML word count), and
the CPU oracle timed on a bounded sample on this host.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "staggered-grid_amd"))

HBM_PEAK_GBPS = 8000.0
METRIC = "Gpt-updates/s + % HBM roofline; FP16-store vs FP32 at fixed seismogram error"


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_baseline(wl, planes=60, steps=10):
    """CPU leg on a bounded sample of the same workload: the first `planes`
    z-planes of the model (same x/y extent and PML), `steps` time steps, one
    thread.  Times the reference's own CPU solver
    (3DCPUwithPML_3DArray_GridIncludePMLandOrder, built into oracle/_ref by
    `make -C oracle ref`) when that build is present, else the C restatement
    in oracle/."""
    import numpy as np

    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O

    nz = min(planes, wl.nz)
    sl = slice(0, nz)
    b, lam, mu = (np.ascontiguousarray(a[sl]) for a in (wl.b, wl.lam, wl.mu))
    src = (wl.src[0], wl.src[1], min(wl.src[2], nz // 2))
    pts = wl.nx * wl.ny * nz
    exe = os.path.join(ROOT, "oracle", "_ref", "ref_workload")
    if os.path.exists(exe):
        P = wl.pml
        el = O.ref_time("workload", (wl.nx - 2 * P - 10, wl.ny - 2 * P - 10, nz - 2 * P - 10), P, steps, src,
                        wl.dx, wl.dy, wl.dz, wl.dt, wl.f0 or 15.0, wl.damp_v, b, lam, mu)
        return {"value": pts * steps / el / 1e9, "unit": "Gpt-updates/s", "cores": 1, "kind": "reference",
                "sample": f"reference CPU solver (ElasticWaveEquation3D.cpp, g++ -O2, 1 thread) on z-planes "
                          f"0..{nz - 1} of the same model ({wl.nx}x{wl.ny}x{nz}, same PML), {steps} steps, "
                          f"{el:.1f} s timed", "host_cpus": os.cpu_count()}
    p = O.OracleParams(wl.nx, wl.ny, nz, wl.pml, wl.dt, wl.dx, wl.dy, wl.dz, steps, src,
                       coef=tuple(float(c) for c in wl.coef), damp_mode=O.DAMP_QUADRATIC, damp_v=wl.damp_v,
                       nthreads=1)
    w = np.asarray(wl.wavelet[:steps], np.float32)
    t = time.perf_counter()
    O.run(p, b, lam, mu, w, None, fields=False)
    el = time.perf_counter() - t
    return {"value": pts * steps / el / 1e9, "unit": "Gpt-updates/s", "cores": 1, "kind": "port",
            "sample": f"oracle/sg_oracle.c (gcc -O2, 1 thread) on z-planes 0..{nz - 1} of the same model "
                      f"({wl.nx}x{wl.ny}x{nz}, same PML), {steps} steps, {el:.1f} s",
            "host_cpus": os.cpu_count()}


def cpu_all_cores(wl, planes=60, steps=10):
    """The C restatement of the reference CPU solver with its loop nest
    split over up to 16 host threads, on the same bounded sample as
    cpu_baseline (informative beside the one-thread reference)."""
    import numpy as np

    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O

    nz = min(planes, wl.nz)
    sl = slice(0, nz)
    b, lam, mu = (np.ascontiguousarray(a[sl]) for a in (wl.b, wl.lam, wl.mu))
    src = (wl.src[0], wl.src[1], min(wl.src[2], nz // 2))
    threads = max(1, min(16, os.cpu_count() or 1))
    p = O.OracleParams(wl.nx, wl.ny, nz, wl.pml, wl.dt, wl.dx, wl.dy, wl.dz, steps, src,
                       coef=tuple(float(c) for c in wl.coef), damp_mode=O.DAMP_QUADRATIC, damp_v=wl.damp_v,
                       nthreads=threads)
    w = np.asarray(wl.wavelet[:steps], np.float32)
    t = time.perf_counter()
    O.run(p, b, lam, mu, w, None, fields=False)
    el = time.perf_counter() - t
    return {"value": wl.nx * wl.ny * nz * steps / el / 1e9, "unit": "Gpt-updates/s", "cores": threads,
            "kind": "port", "sample": f"oracle/sg_oracle.c (gcc -O2, {threads} threads) on z-planes 0..{nz - 1} "
                                      f"of the same model, {steps} steps, {el:.2f} s"}


def read_traffic(workload, precision):
    """HBM bytes per stress launch measured by rocprofv3 --pmc (FETCH_SIZE +
    WRITE_SIZE, copy-kernel calibrated), if a summary is committed."""
    path = os.path.join(ROOT, "profiles", f"{workload}_{precision}_traffic.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_stress_launch:"):
            return float(line.split(":", 1)[1])
    return None


def bench_batched(args):
    """32 uneven shots advanced together (one launch per work-list and
    half-step, hipGraph replay); value = sum over shots of grid points x
    steps / time."""
    import sgfd
    from sgfd import configs

    log("bench: building 32 shots")
    t = time.perf_counter()
    shots = configs.batched_shots(nt=args.warmup + args.steps)
    log(f"bench: shots ready in {time.perf_counter() - t:.1f} s")
    solvers = [wl.solver() for wl in shots]
    b = None
    tuned = None
    try:
        b = sgfd.Batch(solvers)
        if args.tile:
            b.set_launch(*args.tile)
        elif not args.no_tune:
            # z-chunk, register bound and grid of the batched launch, timed
            # outside the measured steps (resets every shot)
            log("bench: batch autotune")
            tile, grid, tms = b.autotune(3)
            tuned = {"tile": list(tile), "grid": grid, "ms_per_step": tms}
            log(f"bench: autotune chose tile {tile} grid {grid} ({tms:.2f} ms per step)")
        b.step(args.warmup)
        b.synchronize()
        ms = b.step_timed(args.steps)
    finally:
        if b is not None:
            b.close()
        for s in solvers:
            s.close()
    pts = sum(wl.points for wl in shots)
    K = args.steps
    value = pts * K / (ms * 1e-3) / 1e9
    seq_value = None
    if args.compare_sequential:
        log("bench: the same shots one solver at a time")
        total = 0.0
        for wl in shots:
            s = wl.solver()
            s.step(args.warmup)
            s.synchronize()
            _, _, t_ms = s.step_timed(K)
            total += t_ms
            s.close()
        seq_value = pts * K / (total * 1e-3) / 1e9
    vb = sb = 0.0
    for wl in shots:
        s = wl.solver()
        v, st = s.sweep_bytes()
        s.close()
        vb += v
        sb += st
    step_frac = (vb + sb) * K / (ms * 1e-3) / 1e9 / HBM_PEAK_GBPS
    out = {"metric": METRIC, "value": value, "unit": "Gpt-updates/s", "n_gpus": 1, "steps": K, "warmup": args.warmup,
           "ms_per_step": ms / K, "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
           "data": "synthetic",
           "config": {"workload": "32 shots, nx,ny U{128..320}, nz U{96..256}, P12, 800 steps (batched)",
                      "points": pts, "parallelism": "single-gpu"},
           "roofline": {"bound": "hbm", "kernel": "full step (both sweeps, all shots)",
                        "achieved": (vb + sb) / (ms / K * 1e-3) / 1e9, "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                        "frac": step_frac, "traffic": None},
           "sequential_value": seq_value,
           "launch": tuned,
           "cpu_baseline": None}
    print(json.dumps(out))


def host_buffers(args):
    """The whole job with the host copies the boundary makes: the model
    uploaded from host arrays (sg_set_medium, host Courant check included),
    the timed steps, the station traces and all nine final fields copied
    back; each phase timed with a synchronisation.  Reported beside the
    HBM-resident rate; never the bench value."""
    import sgfd
    from sgfd import configs

    wl = configs.BUILDERS[args.workload]()
    K = min(args.steps, wl.nt)
    s = wl.solver()  # the first upload and a few steps warm the path
    try:
        s.step(min(10, K))
        s.synchronize()
        s.reset()
        s.synchronize()
        t0 = time.perf_counter()
        s.set_medium(wl.b, wl.lam, wl.mu)
        s.synchronize()
        t1 = time.perf_counter()
        s.step(K)
        s.synchronize()
        t2 = time.perf_counter()
        tr = s.stations()
        t3 = time.perf_counter()
        F = s.fields()
        t4 = time.perf_counter()
    finally:
        s.close()
    pts = wl.points * K
    print(json.dumps({"metric": "Gpt-updates/s with host copies", "workload": wl.description, "steps": K,
                      "upload_s": t1 - t0, "steps_s": t2 - t1, "stations_s": t3 - t2, "fields_s": t4 - t3,
                      "model_bytes": 3 * wl.b.nbytes, "station_bytes": int(tr.nbytes), "field_bytes": int(F.nbytes),
                      "gpt_resident": pts / (t2 - t1) / 1e9, "gpt_with_model_and_stations": pts / (t3 - t0) / 1e9,
                      "gpt_with_final_fields": pts / (t4 - t0) / 1e9}))


def bench_medium(args):
    """The seeded random-medium builder (the reference's spectral method,
    randomMedium/RandomMedia_3_d.m, with its own spectrum and parameters) on
    a 512x512x384 grid: value = grid points / device time of one build (the
    FFT plan and the copy of the result to the host are outside the timing)."""
    import sgfd

    nx, ny, nz = 512, 512, 384
    n = nx * ny * nz
    K = min(args.steps, 20)
    for w in range(max(args.warmup, 1)):
        sgfd.random_medium_timed(nx, ny, nz, "reference", seed=w)
    ms = []
    for k in range(K):
        _, t = sgfd.random_medium_timed(nx, ny, nz, "reference", seed=100 + k)
        ms.append(t)
    avg = sum(ms) / K
    # spectrum write 8 B, transform read + write 16 B (one-pass lower bound),
    # two moment passes 16 B, scaling read 8 B + write 4 B
    alg = 52.0 * n
    out = {"metric": "random-medium build rate", "value": n / (avg * 1e-3) / 1e6, "unit": "Mpt/s", "n_gpus": 1,
           "steps": K, "warmup": max(args.warmup, 1), "ms_per_step": avg, "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f32 (FP64 moments)", "data": "synthetic",
           "config": {"workload": "random medium 512x512x384, the reference script's spectrum (a=10, p=10, l=50), "
                                  "v0 2000, eps 0.1", "grid": [nx, ny, nz], "parallelism": "single-gpu"},
           "roofline": {"bound": "hbm", "kernel": "whole build (spectrum, hipFFT C2C, moments, scaling)",
                        "achieved": alg / (avg * 1e-3) / 1e9, "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                        "frac": alg / (avg * 1e-3) / 1e9 / HBM_PEAK_GBPS, "traffic": None,
                        "alg_bytes_per_launch": alg, "avg_launch_ms": avg}}
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import medium as M

        log("bench: CPU baseline (numpy oracle, 128x128x96)")
        cx, cy, cz = 128, 128, 96
        t = time.perf_counter()
        reps = 0
        while reps < 3 or time.perf_counter() - t < 10.0:
            M.random_medium(cx, cy, cz, "reference", seed=reps)
            reps += 1
        sec = (time.perf_counter() - t) / reps
        out["cpu_baseline"] = {"value": cx * cy * cz / sec / 1e6, "unit": "Mpt/s", "cores": 1, "kind": "port",
                               "sample": f"numpy restatement, {cx}x{cy}x{cz}, {reps} builds"}
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out))


def precision_study(args):
    """The metric's second half: each storage precision runs the full
    workload (the same source, medium and stations); its Gpt-updates/s
    (HIP events over the timed steps) and its seismogram rel-L2 against the
    FP64 build's traces.  16-bit storage uses scales from an FP32 pilot run.
    Reports the FP16/FP32 speed ratio and whether FP16 meets rel-L2 <= 1e-3.
    Two diagnostic rows split the FP16 error: FP32 storage with only the
    model (b, lambda, mu) rounded to FP16 under its power-of-two scales (the
    model's share), and FP16 storage with every field scale multiplied by
    1/8 and by 4 (if the error does not move, the rounding of the 11-bit
    significand sets it, not the scale, so finer scales cannot help)."""
    import numpy as np
    from sgfd import configs

    wl = configs.BUILDERS[args.workload]()
    K = wl.nt
    res = {}
    ref = None
    # (name, storage, rounding, compensated fields, variant)
    modes = [("fp64", "fp64", None, None, None), ("fp32", "fp32", None, None, None),
             ("fp16", "fp16", None, None, None), ("bf16", "bf16", None, None, None),
             ("fp16-stochastic", "fp16", "stochastic", None, None), ("fp16-compensated", "fp16", None, "all", None),
             ("fp16-compensated-velocity", "fp16", None, "velocity", None),
             ("fp16-compensated-stress", "fp16", None, "stress", None),
             ("bf16-stochastic", "bf16", "stochastic", None, None), ("bf16-compensated", "bf16", None, "all", None),
             ("fp32-model-fp16", "fp32", None, None, "model16"), ("fp16-scales-x1/8", "fp16", None, None, 0.125),
             ("fp16-scales-x4", "fp16", None, None, 4.0)]
    if args.quick_study:
        modes = [m for m in modes if m[0] in ("fp64", "fp32", "fp16", "fp32-model-fp16", "fp16-scales-x1/8",
                                              "fp16-scales-x4")]
    for name, prec, rounding, comp, variant in modes:
        log(f"bench: {name}, {K} steps")
        s = wl.solver(precision=prec)
        try:
            if variant == "model16":
                s.set_medium(*[model_rounded_fp16(a) for a in (wl.b, wl.lam, wl.mu)])
            elif variant is not None:
                s.set_field_scales(wl._scales * np.float32(variant))
            if rounding:
                s.set_rounding(rounding)
            if comp:
                s.set_compensation(comp)
            _, _, ms = s.step_timed(K)
            tr = s.stations().astype(np.float64)
        finally:
            s.close()
        if ref is None:
            ref = tr
        err = float(np.linalg.norm(tr - ref) / np.linalg.norm(ref))
        res[name] = {"gpt_per_s": wl.points * K / (ms / 1e3) / 1e9, "ms_per_step": ms / K,
                     "seismogram_rel_l2_vs_fp64": err}
    ratio = res["fp16"]["gpt_per_s"] / res["fp32"]["gpt_per_s"]
    out = {"metric": "FP16-store vs FP32 at fixed seismogram error", "value": ratio, "unit": "x (FP16/FP32 Gpt/s)",
           "n_gpus": 1, "steps": K, "higher_is_better": True, "data": "synthetic",
           "config": {"workload": wl.description, "grid": [wl.nx, wl.ny, wl.nz], "pml": wl.pml,
                      "stations": int(len(wl.stations)), "components": list(wl.comps)},
           "target": {"ratio": 1.7, "seismogram_rel_l2": 1e-3},
           "fp16_meets_error_target": res["fp16"]["seismogram_rel_l2_vs_fp64"] <= 1e-3,
           "precisions": res}
    print(json.dumps(out))


def model_rounded_fp16(a):
    """The model array as FP16 storage holds it: scaled by the power of two
    that maps its largest |value| into [2^13, 2^14), rounded to half, scaled
    back (float32)."""
    import numpy as np

    a = np.asarray(a, np.float32)
    m = float(np.abs(a).max())
    sc = np.float32(2.0 ** np.floor(np.log2(16384.0 / m)))
    return ((a * sc).astype(np.float16).astype(np.float32) / sc).astype(np.float32)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="layered512")
    ap.add_argument("--precision", default="fp32", choices=["fp32", "fp64", "fp16", "bf16"])
    ap.add_argument("--kernel", default=None, choices=["tiled", "simple"],
                    help="default: tiled (FP64 runs the per-point kernels)")
    ap.add_argument("--tile", type=int, nargs=5, default=None, metavar=("TY", "NY", "PTY", "PNY", "ZCHUNK"),
                    help="interior box rows and rows per thread, PML box rows and rows per thread, "
                         "z planes per box (0 = auto)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-fp16", action="store_true", help="skip the FP16-storage leg of an FP32 run")
    ap.add_argument("--no-tune", action="store_true", help="batched32: keep the default launch (no autotuner)")
    ap.add_argument("--tune", action="store_true",
                    help="single-shot workloads: run the solver autotuner (sg_autotune) before timing")
    ap.add_argument("--streams", type=int, default=None, help="streams per sweep (1, 2, 3)")
    ap.add_argument("--copy-calib", action="store_true", help="only measure the copy-kernel bandwidth")
    ap.add_argument("--compare-sequential", action="store_true",
                    help="batched32: also time the shots one solver at a time")
    ap.add_argument("--quick-study", action="store_true",
                    help="precision study: only FP64, FP32, FP16 and the FP16 error-split rows")
    ap.add_argument("--host-buffers", action="store_true",
                    help="the whole job with the model upload and the station and field downloads, FP32")
    ap.add_argument("--precision-study", action="store_true",
                    help="every storage precision and 16-bit mode over the full workload: speed and seismogram "
                         "error against FP64")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py: this solver runs on one GPU; --gpus {args.gpus} is not supported "
              "(multi-GPU sharding is not implemented)", file=sys.stderr)
        sys.exit(2)

    import sgfd
    from sgfd import configs

    if args.copy_calib:
        gbps = sgfd.copy_bandwidth(1 << 30, 20)
        print(json.dumps({"copy_gbps": gbps, "bytes": 1 << 30}))
        return
    if args.workload == "batched32":
        bench_batched(args)
        return
    if args.workload == "medium512":
        bench_medium(args)
        return
    if args.host_buffers:
        host_buffers(args)
        return
    if args.precision_study:
        precision_study(args)
        return

    log(f"bench: building workload {args.workload}")
    t = time.perf_counter()
    wl = configs.BUILDERS[args.workload]()
    log(f"bench: workload ready in {time.perf_counter() - t:.1f} s, courant {wl.courant():.3f}")
    K = args.steps
    r = timed_run(sgfd, wl, args, args.precision)
    sec = r["ms_total"] / 1e3
    value = wl.points * K / sec / 1e9
    # updated points (5-cell halo excluded) and the undamped interior among them
    upd = (wl.nx - 10) * (wl.ny - 10) * (wl.nz - 10)
    inner = max(wl.nx - 10 - 2 * wl.pml, 0) * max(wl.ny - 10 - 2 * wl.pml, 0) * max(wl.nz - 10 - 2 * wl.pml, 0)
    # the roofline also as a fraction of this device's measured copy bandwidth
    copy_gbps = sgfd.copy_bandwidth(1 << 30, 20)
    roof = r["roofline"]
    roof["copy_gbps"] = copy_gbps
    roof["copy_frac"] = roof["achieved"] / copy_gbps
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "Gpt-updates/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": args.warmup,
        "ms_per_step": r["ms_total"] / K,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": DTYPES[args.precision],
        "data": "synthetic",
        "config": {"workload": wl.description, "grid": [wl.nx, wl.ny, wl.nz], "pml": wl.pml,
                   "points": {"total": wl.points, "updated": upd, "interior": inner},
                   "precision": args.precision, "kernel": kernel_of(args), "parallelism": "single-gpu"},
        "roofline": roof,
        "sweeps_ms": r["sweeps_ms"],
        "launch": r["launch"],
        "step_hbm_frac": r["step_hbm_frac"],
        "energy_after": r["energy"],
    }
    if args.precision == "fp32" and not args.no_fp16:
        # the metric's second half: the same workload with FP16 storage
        # (scales from an FP32 pilot run), timed the same way
        r16 = timed_run(sgfd, wl, args, "fp16")
        v16 = wl.points * K / (r16["ms_total"] / 1e3) / 1e9
        out["fp16"] = {"value": v16, "unit": "Gpt-updates/s", "ratio_to_fp32": v16 / value,
                       "ms_per_step": r16["ms_total"] / K, "dtype": DTYPES["fp16"], "roofline": r16["roofline"],
                       "sweeps_ms": r16["sweeps_ms"], "target_ratio": 1.7,
                       "seismogram_rel_l2_vs_fp64": read_precision_error(args.workload, "fp16")}
    if not args.no_cpu:
        log("bench: CPU baseline (reference CPU solver, bounded sample)")
        out["cpu_baseline"] = cpu_baseline(wl)
        log("bench: CPU restatement on all host cores (same sample)")
        out["cpu_all_cores"] = cpu_all_cores(wl)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out))


def kernel_of(args, precision=None):
    """The sweep kernels a run uses: --kernel, else the tiled sweeps (the
    per-point kernels for FP64, the only ones the library runs in FP64)."""
    p = precision or args.precision
    return args.kernel or ("simple" if p == "fp64" else "tiled")


DTYPES = {"fp32": "f32", "fp64": "f64", "fp16": "f16-storage/f32", "bf16": "bf16-storage/f32"}


def timed_run(sgfd, wl, args, precision):
    """args.warmup untimed steps, then args.steps steps with HIP events
    around each sweep (sg_step_timed); the dominant sweep's roofline."""
    s = wl.solver(precision=precision, kernel=args.kernel, time_length=args.warmup + args.steps)
    try:
        if args.tile:
            s.set_tile(*args.tile)
        if args.streams:
            s.set_streams(args.streams)
        launch = None
        if args.tune and kernel_of(args, precision) == "tiled":
            log(f"bench: {precision} autotune")
            tile, tms = s.autotune(4)
            launch = {"tile": list(tile), "streams": s.streams(), "autotune_ms_per_step": tms}
            log(f"bench: autotune chose tile {tile}, {launch['streams']} stream(s) ({tms:.3f} ms per step)")
        log(f"bench: {precision} warmup {args.warmup} steps")
        s.step(args.warmup)
        s.synchronize()
        log(f"bench: {precision} timing {args.steps} steps")
        ms_v, ms_s, ms_tot = s.step_timed(args.steps)
        s.synchronize()
        vbytes, sbytes = s.sweep_bytes()
        energy = s.energy()
    finally:
        s.close()
    K = args.steps
    avg_v, avg_s = ms_v / K, ms_s / K
    dom = ("stress", avg_s, sbytes) if avg_s >= avg_v else ("velocity", avg_v, vbytes)
    achieved = dom[2] / (dom[1] * 1e-3) / 1e9
    roof = {"bound": "hbm", "kernel": dom[0], "achieved": achieved, "peak": HBM_PEAK_GBPS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBPS, "traffic": read_traffic(args.workload, precision),
            "alg_bytes_per_launch": dom[2], "avg_launch_ms": dom[1]}
    if launch is not None:
        roof["traffic"] = None  # the committed traffic summary covers the default launch only
    return {"ms_total": ms_tot, "roofline": roof, "launch": launch,
            "sweeps_ms": {"velocity_avg": avg_v, "stress_avg": avg_s},
            "step_hbm_frac": (vbytes + sbytes) * K / (ms_tot / 1e3) / 1e9 / HBM_PEAK_GBPS, "energy": list(energy)}


def read_precision_error(workload, precision):
    """Seismogram rel-L2 against the FP64 build over the full run, as
    measured by the full-size GPU test and recorded in the committed
    precision summary (profiles/<workload>_precision_study.md)."""
    path = os.path.join(ROOT, "profiles", f"{workload}_precision_study.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith(f"seismogram_rel_l2_{precision}:"):
            return {"value": float(line.split(":", 1)[1]), "source": os.path.relpath(path, ROOT)}
    return None


if __name__ == "__main__":
    main()
