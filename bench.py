"""Benchmark: one AL-DIC frame on one MI355X.

Default workload (one "step" = one frame): 4096x4096 12-bit speckle stored as uint16
(600k Gaussian dots, sigma U[1.5, 3] px), 0.5 % tension plus a Gaussian shear band (width
200 px, peak shear 3 %), noise sigma 2; subsets 32 px on an 8 px grid (247,009 subsets),
integer search radius 24: normalisation + gradients, integer search, 6-DOF IC-GN with the
bad-point rule and natural-neighbour refill, step-1 global step, and 3 ADMM iterations
(2-DOF IC-GN + gridfit refit, Q4 FEM global step, nodal strain, dual update).  The raw
pair is resident in HBM before the timed region.  --workload 2 runs the 2048^2 uint8
ws32/step16 frame (14,884 subsets), --workload 4 the quadtree plate with a hole.

value = converged subsets (6-DOF IC-GN status 1..100) per second of whole frame.
accuracy_vs_fp64 and cpu_baseline come from the FP64 oracle on a seeded 2 % sample of
the subsets (search + 6-DOF IC-GN + three 2-DOF IC-GN passes per subset, OpenMP; the
full-frame local-step time is that sample scaled by S / n_sample; the global steps are
not in the CPU figure); workload 2 runs the oracle's whole frame instead.

  python bench.py [--steps K] [--warmup W] [--workload 2|3|4] [--no-cpu] [--no-stages]
"""
from __future__ import annotations

import argparse
import importlib
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
INT8_PEAK_TOPS = 5000.0    # MI355X dense int8 MFMA
FP32_PEAK_TFLOPS = 157.3   # MI355X FP32
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "hbm_traffic.json")
VALU_CLOCK_HZ = 2.4e9  # MI355X peak engine clock


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def copy_calibration(nbytes):
    """One plain device-to-device copy of nbytes (hipMemcpy's copy kernel): the reference
    point for scaling the FETCH_SIZE / WRITE_SIZE counters to HBM bytes."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    a, b = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(a), C.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(C.byref(b), C.c_size_t(nbytes)) == 0
    hip.hipMemset(a, 1, C.c_size_t(nbytes))
    assert hip.hipMemcpy(b, a, C.c_size_t(nbytes), 3) == 0  # device to device
    hip.hipDeviceSynchronize()
    hip.hipFree(a)
    hip.hipFree(b)
    log(f"[bench] copy calibration: {nbytes} bytes device-to-device")


def measure_copy_bw(nbytes=1 << 30, reps=5):
    """Device-to-device copy bandwidth (read + write bytes / time) over `reps` copies of
    nbytes, timed with HIP events: the measured HBM ceiling beside the nominal 8 TB/s."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    a, b = C.c_void_p(), C.c_void_p()
    if hip.hipMalloc(C.byref(a), C.c_size_t(nbytes)) or hip.hipMalloc(C.byref(b), C.c_size_t(nbytes)):
        return None
    e0, e1 = C.c_void_p(), C.c_void_p()
    hip.hipEventCreate(C.byref(e0))
    hip.hipEventCreate(C.byref(e1))
    hip.hipMemset(a, 1, C.c_size_t(nbytes))
    hip.hipMemcpy(b, a, C.c_size_t(nbytes), 3)
    hip.hipDeviceSynchronize()
    hip.hipEventRecord(e0, None)
    for _ in range(reps):
        hip.hipMemcpyAsync(b, a, C.c_size_t(nbytes), 3, None)
    hip.hipEventRecord(e1, None)
    hip.hipEventSynchronize(e1)
    ms = C.c_float(0.0)
    hip.hipEventElapsedTime(C.byref(ms), e0, e1)
    hip.hipEventDestroy(e0)
    hip.hipEventDestroy(e1)
    hip.hipFree(a)
    hip.hipFree(b)
    return round(2.0 * nbytes * reps / (ms.value / 1e3) / 1e9, 1) if ms.value > 0 else None


def cpu_model():
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-stages", action="store_true",
                    help="skip the untimed per-stage breakdown frames")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--workload", type=int, default=3, choices=(2, 3, 4),
                    help="3: 4096^2 12-bit ws32/step8 s24, 247,009 subsets (default); 2: "
                         "2048^2 uint8 ws32/step16 s8; 4: 2048^2 plate with a hole on a "
                         "quadtree mesh, windows 16-64 px")
    ap.add_argument("--sample", type=float, default=0.02,
                    help="fraction of the subsets the FP64 oracle checks and times (workload 3)")
    ap.add_argument("--no-smooth", action="store_true",
                    help="skip the smoothing-on frames (funSmoothDisp / funSmoothStrain)")
    ap.add_argument("--frames", type=int, default=1,
                    help="workload 4: above 1, the incremental sequence of that many frames "
                         "(far-field strain 0.05 %% per frame, each frame seeded with the "
                         "previous one's U) on the ~29k-node mesh")
    ap.add_argument("--no-learned", action="store_true",
                    help="skip LocalICGN from the learned start against the search start")
    ap.add_argument("--no-mg", action="store_true",
                    help="skip the multigrid integer search line (funIntegerSearchMg)")
    ap.add_argument("--copy-calib", action="store_true",
                    help="also run one 256 MiB device-to-device copy (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the multi-GPU "
              "sharding is a design note, not a measured path.", file=sys.stderr)
        sys.exit(2)

    pkg = importlib.import_module("2d_aldic_amd")
    sp = pkg.speckle
    if args.workload == 4:
        return run_quadtree_sequence(args, pkg) if args.frames > 1 else run_quadtree(args, pkg)
    t = time.perf_counter()
    if args.workload == 3:
        pair = sp.config3()
        ws, step, srad, side = 32, 8, 24, 4096
        wname = "config3: 4096x4096 12-bit speckle in uint16, 0.5% tension + shear band, ws32/step8"
    else:
        pair = sp.config2()
        ws, step, srad, side = 32, 16, 8, 2048
        wname = "config2: 2048x2048 uint8 speckle, sinusoidal field, ws32/step16"
    log(f"[bench] workload generated in {time.perf_counter() - t:.1f} s")
    grid = [1, side, 1, side]
    mu = 1e-3
    beta = 1e-2 * step * step * mu
    n_admm = 3

    ctx = pkg.Context(0)
    if args.copy_calib:
        copy_calibration(256 << 20)
    ctx.frame_stage(pair.f, pair.g)
    # timed frames carry events only around the search, the 6-DOF IC-GN, their two main
    # kernels and the whole frame; the per-stage breakdown comes from a second, untimed
    # set of frames with an event at every stage boundary
    S = ctx.frame_setup(ws, step, srad, grid, tol=1e-2, mu=mu, beta=beta, n_admm=n_admm,
                        cg_maxit=2000, cg_rtol=1e-10, stage_times=False)
    log(f"[bench] frame set up: {S} subsets")
    for _ in range(args.warmup):
        ctx.frame_run()
    log("[bench] warmup done")
    acc_t = {}
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ctx.frame_run()
        for k, v in ctx.frame_times().items():
            acc_t[k] = acc_t.get(k, 0.0) + v
    wall = time.perf_counter() - t0
    kt = {k: v / args.steps for k, v in acc_t.items()}
    ms_per_step = wall * 1e3 / args.steps
    out = ctx.frame_get()
    counters = ctx.frame_counters()
    log(f"[bench] {ms_per_step:.3f} ms/frame (search kernel {kt['search_kernel']:.3f} ms, "
        f"6-DOF kernel {kt['icgn6_kernel']:.3f} ms)")
    stage = {}
    if not args.no_stages:
        ctx.frame_setup(ws, step, srad, grid, tol=1e-2, mu=mu, beta=beta, n_admm=n_admm,
                        cg_maxit=2000, cg_rtol=1e-10, stage_times=True)
        nst = max(3, min(args.steps, 10))
        for _ in range(nst):
            ctx.frame_run()
            for k, v in ctx.frame_times().items():
                stage[k] = stage.get(k, 0.0) + v
        stage = {k: v / nst for k, v in stage.items()}
        log(f"[bench] stages {stage}")
    smooth_ms = None
    if not args.no_smooth:
        # main_ALDIC.m's smoothing passes on (funSmoothStrain of F1 and every F2,
        # funSmoothDisp of U1 and U2 in the ADMM steps: main_ALDIC.m:233,320,381,400,407)
        ctx.frame_setup(ws, step, srad, grid, tol=1e-2, mu=mu, beta=beta, n_admm=n_admm,
                        cg_maxit=2000, cg_rtol=1e-10, stage_times=False, smooth=True)
        ctx.frame_run()
        nsm = max(3, min(args.steps, 10))
        t0 = time.perf_counter()
        for _ in range(nsm):
            ctx.frame_run()
        smooth_ms = (time.perf_counter() - t0) * 1e3 / nsm
        log(f"[bench] smoothing on: {smooth_ms:.3f} ms/frame")
    learned = None
    frame_learned = None
    if args.workload == 3 and not args.no_learned:
        learned = learned_start(ctx, pkg, pair, grid, ws, step, srad, args.steps)
        log(f"[bench] learned start: {learned}")
        frame_learned = learned_frame(ctx, pair, out, ms_per_step, ws, step, srad, grid, mu,
                                      beta, n_admm, args.steps, args.warmup)
        log(f"[bench] frame with the learned start: {frame_learned}")
    mgs = None
    if not args.no_mg:
        mgs = multigrid_search(ctx, pair, grid, ws, step, args.steps)
        log(f"[bench] multigrid search: {mgs}")
    it6 = out["iters6"]
    conv = int(np.sum((it6 >= 1) & (it6 <= 100)))
    value = conv / (ms_per_step / 1e3)

    npx = side * side
    traffic = {}
    try:
        with open(TRAFFIC_FILE) as fh:
            traffic = json.load(fh).get(f"workload{args.workload}", {})
    except (OSError, ValueError):
        pass

    def measured(kernel, key="hbm_bytes_per_launch"):
        return traffic.get(kernel, {}).get(key)

    # The integer search: int8 matrix cores, algorithmic ops = the NCC numerator's
    # multiply-adds x 2, 2 (2s+3)^2 (w+1)^2 per subset; bytes = raw pair once + outputs
    # (+ the window box-sum planes of the multi-tile kernel).  Timed live: the kernel's
    # own events on the library's stream, averaged over the timed frames.
    ops = 2.0 * (2 * srad + 3) ** 2 * (ws + 1) ** 2 * S
    t_s = kt["search_kernel"] / 1e3
    small = pair.f.dtype == np.uint8 and srad <= 14
    bpp = pair.f.dtype.itemsize
    b12 = bpp == 2 and int(pair.f.max()) < 4096 and int(pair.g.max()) < 4096
    skern = "k_search_i8" if small else ("k_search_k3" if b12 else "k_search_mt")
    engine = {"k_search_i8": "int8 MFMA (32x32x32), one 8-bit plane",
              "k_search_k3": "int8 MFMA (32x32x32), base-64 digit planes, 3 GEMMs (Karatsuba)",
              "k_search_mt": "int8 MFMA (32x32x32)" + ("" if bpp == 1 else
                                                        ", two digit planes (4 GEMMs)")}[skern]
    search_bytes = npx * 2 * bpp + 32 * S + (0 if small else npx * 36)
    vs = measured(skern, "valu_instr_per_launch")
    search_roof = dict(kernel=skern, bound="mfma", engine=engine,
                       achieved=round(ops / t_s / 1e12, 3), peak=INT8_PEAK_TOPS,
                       unit="TOP/s", frac=round(ops / t_s / 1e12 / INT8_PEAK_TOPS, 5),
                       traffic=measured(skern), algorithmic_ops=ops,
                       algorithmic_bytes=search_bytes, avg_ms=round(kt["search_kernel"], 5),
                       valu_frac=(round(vs * 4 / (1024 * VALU_CLOCK_HZ * t_s), 4) if vs else None),
                       mfma_per_launch=measured(skern, "mfma_instr_per_launch"))
    # The 6-DOF IC-GN (one k_icgn launch per frame): bandwidth class, compulsory bytes
    # 16 N_px (f, fx, fy, g as FP32) + 124 S (2 words in; 6 + 21 + 2 words out) per launch
    icgn_bytes = 16 * npx + 124 * S
    t_i = kt["icgn6_kernel"] / 1e3
    vi = measured("k_icgn_6dof", "valu_instr_per_launch")
    icgn_roof = dict(kernel="k_icgn (6-DOF IC-GN, ws 32)", bound="hbm",
                     achieved=round(icgn_bytes / t_i / 1e9, 2), peak=HBM_PEAK_GBS, unit="GB/s",
                     frac=round(icgn_bytes / t_i / 1e9 / HBM_PEAK_GBS, 5),
                     traffic=measured("k_icgn_6dof"), algorithmic_bytes=icgn_bytes,
                     avg_ms=round(kt["icgn6_kernel"], 5),
                     valu_frac=(round(vi * 4 / (1024 * VALU_CLOCK_HZ * t_i), 4) if vi else None),
                     note="VALU-issue-bound (FP32 Keys gathers and sums per pixel pass)")
    # the dominant kernel of the frame is the roofline line
    if kt["search_kernel"] >= kt["icgn6_kernel"]:
        roof, extra_roof, extra_key = search_roof, icgn_roof, "icgn_roofline"
    else:
        roof, extra_roof, extra_key = icgn_roof, search_roof, "search_roofline"
    roof["copy_gbs_measured"] = measure_copy_bw()

    cpu = None
    acc = {}
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        threads = args.cpu_threads or int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or \
            len(os.sched_getaffinity(0))
        os.environ["OMP_NUM_THREADS"] = str(threads)
        import oracle
        oracle.build()
        if args.workload == 3:
            cpu, acc = cpu_sample(oracle, pair, out, S, ws, step, srad, grid, mu, args.sample,
                                  threads)
        else:
            cpu, acc = cpu_full_frame(oracle, pair, out, S, ws, step, srad, grid, mu, beta,
                                      n_admm, threads)
        log(f"[bench] CPU baseline {cpu}")
        log(f"[bench] accuracy vs FP64: {acc}")

    line = {
        "metric": "converged subsets/s at <=0.01 px RMS vs FP64; ADMM frame time (ms)",
        "value": round(value, 2),
        "unit": "converged subsets/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"{wname} ({S} subsets), search {srad}, 3 ADMM iterations, "
                               "FEM Q4",
                   "subsets": S, "converged": conv, "frame_ms": round(ms_per_step, 4),
                   "frame_ms_smoothing_on": None if smooth_ms is None else round(smooth_ms, 4),
                   "kernel_ms": {k: round(kt[k], 4) for k in ("search_kernel", "icgn6_kernel")},
                   "stage_ms": {k: round(v, 4) for k, v in stage.items()},
                   "stage_ms_from": "separate untimed frames with an event at every stage boundary",
                   "gridfit": counters,
                   "icgn6_subsets_per_s": round(conv / t_i, 1),
                   "learned_start": learned, "frame_learned_start": frame_learned,
                   "multigrid_search": mgs},
        "roofline": roof,
        extra_key: extra_roof,
        "cpu_baseline": cpu,
        "accuracy_vs_fp64": acc or None,
    }
    print(json.dumps(line), flush=True)


def learned_start(ctx, pkg, pair, grid, ws, step, srad, reps):
    """LocalICGN's start + 6-DOF IC-GN over the whole grid (aldic_local_icgn), from the
    integer search and from the learned warp (search fallback on the subsets that miss
    20 iterations or ZNCC 0.9), on the same pair: device time from HIP events around the
    stage (results' copies to the host excluded), converged subsets, iterations, fallback
    count, and both forms' U against each other and against the analytic field."""
    cnn = importlib.import_module("2d_aldic_amd.cnn")
    ctx.set_images(pair.f, pair.g)
    cnn.load_weights(ctx)
    res, ms = {}, {}
    for start in ("search", "learned"):
        ctx.local_icgn(grid, ws, step, [srad, srad], start=start)          # warmup
        acc = np.zeros(5)
        n = max(3, min(reps, 10))
        for _ in range(n):
            r = ctx.local_icgn(grid, ws, step, [srad, srad], start=start)
            acc += r["times_ms"]
        res[start], ms[start] = r, acc / n
    out = {}
    xy = np.column_stack([res["search"]["x"], res["search"]["y"]])
    ua, va = pair.ufun(xy[:, 0], xy[:, 1])
    for start, r in res.items():
        ok = (r["iters"] >= 1) & (r["iters"] <= 100)
        U = r["U"].reshape(-1, 2)
        err = np.hypot(U[ok, 0] - ua[ok], U[ok, 1] - va[ok])
        out[start] = dict(stage_ms=round(float(ms[start][4]), 4),
                          start_ms=round(float(ms[start][0]), 4),
                          icgn_ms=round(float(ms[start][1]), 4),
                          fallback_ms=round(float(ms[start][2] + ms[start][3]), 4),
                          converged=int(ok.sum()),
                          converged_subsets_per_s=round(float(ok.sum()) / (ms[start][4] / 1e3), 1),
                          mean_iters=round(float(r["iters"][ok].mean()), 4),
                          fallback=int(r["n_fallback"]),
                          rms_vs_field_px=round(float(np.sqrt(np.mean(err ** 2))), 5))
    both = ((res["search"]["iters"] >= 1) & (res["search"]["iters"] <= 100) &
            (res["learned"]["iters"] >= 1) & (res["learned"]["iters"] <= 100))
    d = (res["learned"]["U"] - res["search"]["U"]).reshape(-1, 2)[both]
    out["learned_vs_search_rms_px"] = float(np.sqrt(np.mean(np.sum(d ** 2, 1))))
    out["learned_vs_search_max_px"] = float(np.abs(d).max())
    out["speedup"] = round(out["learned"]["converged_subsets_per_s"] /
                           out["search"]["converged_subsets_per_s"], 3)
    out["note"] = ("LocalICGN stage only (start + 6-DOF IC-GN over 247,009 subsets); the "
                   "search start is the exact 12-bit integer search")
    return out


def learned_frame(ctx, pair, out_s, ms_s, ws, step, srad, grid, mu, beta, n_admm, steps, warmup):
    """The same device frame with LocalICGN started from the learned warp (frame option
    learned_start; the search runs only on the fallback list): wall time per frame as the
    headline's, converged subsets, fallback count, and its final U against the searched
    frame's and against the analytic field."""
    ctx.frame_stage(pair.f, pair.g)
    ctx.frame_setup(ws, step, srad, grid, tol=1e-2, mu=mu, beta=beta, n_admm=n_admm,
                    cg_maxit=2000, cg_rtol=1e-10, stage_times=False, learned_start=True)
    for _ in range(max(1, warmup)):
        ctx.frame_run()
    t0 = time.perf_counter()
    for _ in range(steps):
        ctx.frame_run()
    ms = (time.perf_counter() - t0) * 1e3 / steps
    r = ctx.frame_get()
    fb = ctx.frame_counters()["learned_fallback"]
    it6 = r["iters6"]
    conv = int(np.sum((it6 >= 1) & (it6 <= 100)))
    S = it6.size
    x1, y1, Nx, Ny = ctx.search_grid(grid, ws, step, [srad, srad])
    X, Y = np.meshgrid(x1 + np.arange(Nx) * step + 0.5 * ws, y1 + np.arange(Ny) * step + 0.5 * ws)
    xy = np.column_stack([X.ravel(), Y.ravel()])
    d = (r["U"] - out_s["U"]).reshape(-1, 2)
    res = dict(frame_ms=round(ms, 4), converged=conv, converged_subsets_per_s=round(conv / (ms / 1e3), 1),
               fallback=fb, vs_searched_frame_rms_px=float(np.sqrt(np.mean(np.sum(d ** 2, 1)))),
               vs_searched_frame_max_px=float(np.abs(d).max()),
               speedup_vs_searched_frame=round(ms_s / ms, 3))
    ua, va = pair.ufun(xy[:, 0], xy[:, 1])
    for name, UU in (("rms_vs_field_px", r["U"]), ("searched_frame_rms_vs_field_px", out_s["U"])):
        U = UU.reshape(-1, 2)
        res[name] = float(np.sqrt(np.mean((U[:, 0] - ua) ** 2 + (U[:, 1] - va) ** 2)))
    res["note"] = (f"{S} subsets; the search runs only for the fallback list; outlier stage off "
                   "(no search field to test)")
    return res


def multigrid_search(ctx, pair, grid, ws, step, reps):
    """funIntegerSearchMg (IntegerSearch.m's method 0) over the workload's ROI on the
    staged pair: wall time of the whole call (level-1 FFT correlation, every later level,
    the host's region bookkeeping between levels and the final gridfit), levels, final
    regions, and the fitted grid against the analytic field."""
    ctx.set_images(pair.f, pair.g)
    g4 = (grid[0], grid[1], grid[2], grid[3])
    r = ctx.integer_search_mg(g4, ws, step)                 # warmup
    n = max(3, min(reps, 10))
    t0 = time.perf_counter()
    for _ in range(n):
        r = ctx.integer_search_mg(g4, ws, step)
    ms = (time.perf_counter() - t0) * 1e3 / n
    ua, va = pair.ufun(r["x"], r["y"])
    err = np.hypot(r["u"] - ua, r["v"] - va)
    ok = ~np.isnan(r["tempu"])
    ta, tv = pair.ufun(r["tempx"][ok], r["tempy"][ok])
    terr = np.hypot(r["tempu"][ok] - ta, r["tempv"][ok] - tv)
    return dict(ms=round(ms, 3), levels=r["levels"], regions=int(r["tempx"].size),
                regions_nan=r["n_nan"], grid=[r["Nx"], r["Ny"]], fit_cg_iters=r["fit_iters"],
                region_rms_vs_field_px=round(float(np.sqrt(np.mean(terr ** 2))), 4),
                grid_rms_vs_field_px=round(float(np.sqrt(np.mean(err ** 2))), 4),
                note="wall time per call incl. host bookkeeping and per-level waits")


def cpu_full_frame(oracle, pair, out, S, ws, step, srad, grid, mu, beta, n_admm, threads):
    """The oracle's whole frame (workload 2): CPU baseline and accuracy of the device frame."""
    log(f"[bench] CPU baseline: FP64 oracle frame on {threads} threads")
    o = oracle.frame(pair.f, pair.g, ws, step, srad, grid, tol=1e-2, mu=mu, beta=beta,
                     n_admm=n_admm)
    it6 = out["iters6"]
    itc = o["iters6"]
    convc = int(np.sum((itc >= 1) & (itc <= 100)))
    cpu = dict(value=round(convc / o["times"]["frame"], 2), unit="converged subsets/s",
               cores=threads, kind="port", cpu_model=cpu_model(),
               sample=f"one full frame of the same workload ({S} subsets), FP64 oracle, "
                      f"{o['times']['frame']:.2f} s",
               stage_s={k: round(v, 4) for k, v in o["times"].items()})
    same = (it6 == itc) & (it6 >= 1) & (it6 <= 100)
    d6 = (out["U6"] - o["U6"]).reshape(-1, 2)[same]
    dF = (out["F6"] - o["F6"]).reshape(-1, 4)[same]
    acc = dict(icgn6_status_agree=round(float(np.mean(it6 == itc)), 5),
               icgn6_rms_px=float(np.sqrt(np.mean(d6 ** 2))),
               icgn6_max_px=float(np.abs(d6).max()),
               icgn6_max_F=float(np.abs(dF).max()),
               frame_rms_px=float(np.sqrt(np.nanmean((out["U"] - o["U"]) ** 2))),
               search_max_px=float(np.nanmax(np.abs(out["U0"] - o["U0"]))))
    acc["gate_0p01px"] = bool(acc["icgn6_rms_px"] <= 0.01)
    return cpu, acc


def cpu_sample(oracle, pair, out, S, ws, step, srad, grid, mu, frac, threads):
    """Workload 3: the FP64 oracle on a seeded sample of the subsets.  Accuracy: the
    device's search peaks and 6-DOF IC-GN results on the sample against the oracle's.
    CPU time: normalisation + gradients of the whole pair once, then search + 6-DOF
    IC-GN + three 2-DOF IC-GN passes on the sample (median of 3), scaled by S / n."""
    n = max(1, int(round(frac * S)))
    q = np.sort(np.random.default_rng(2026).choice(S, n, replace=False)).astype(np.int32)
    log(f"[bench] FP64 oracle on a seeded {n}-subset sample ({100 * n / S:.1f} %), "
        f"{threads} threads")
    t = time.perf_counter()
    f, _ = oracle.normalize(pair.f.astype(np.float64))
    g, _ = oracle.normalize(pair.g.astype(np.float64))
    gx, gy = oracle.gradient(f)
    t_prep = time.perf_counter() - t
    s2 = [srad, srad]
    times = []
    for rep in range(3):
        t = time.perf_counter()
        r = oracle.search_sample(f, g, grid, ws, step, s2, q)
        t1 = time.perf_counter()
        xy = np.stack([r["x"], r["y"]], 1)
        U0 = np.stack([r["u"], r["v"]], 1).ravel()
        U6, F6, H, it6o, _ = oracle.icgn6(f, g, gx, gy, xy, U0, ws)
        t2 = time.perf_counter()
        z4, z2 = np.zeros(4 * n), np.zeros(2 * n)
        for _ in range(3):
            oracle.icgn2(f, g, gx, gy, xy, ws, H, mu, z4, z2, U6, F6)
        t3 = time.perf_counter()
        times.append((t3 - t, t1 - t, t2 - t1, t3 - t2))
    times.sort()
    tot, ts, t6, t2 = times[1]
    conv_o = int(np.sum((it6o >= 1) & (it6o <= 100)))
    frame_s = t_prep + tot * S / n
    cpu = dict(value=round(conv_o * S / n / frame_s, 2), unit="converged subsets/s",
               cores=threads, kind="port", cpu_model=cpu_model(),
               sample=f"seeded {n}-subset sample ({100 * n / S:.1f} %) of the same frame: "
                      f"search + 6-DOF + 3 x 2-DOF IC-GN in {tot:.2f} s (median of 3), scaled "
                      f"by S/n to {frame_s:.1f} s per frame with the whole-image "
                      f"normalisation ({t_prep:.2f} s); global steps not included",
               stage_s=dict(prep=round(t_prep, 3), search=round(ts * S / n, 2),
                            icgn6=round(t6 * S / n, 2), icgn2_x3=round(t2 * S / n, 2),
                            frame_local=round(frame_s, 2)))
    Ud0 = out["U0"].reshape(-1, 2)[q]
    Ud6 = out["U6"].reshape(-1, 2)[q]
    Fd6 = out["F6"].reshape(-1, 4)[q]
    itd = out["iters6"][q]
    fin = np.isfinite(U0.reshape(-1, 2)).all(1)
    same = (itd == it6o) & (itd >= 1) & (itd <= 100)
    d6 = Ud6[same] - U6.reshape(-1, 2)[same]
    dF = Fd6[same] - F6.reshape(-1, 4)[same]
    acc = dict(sample=int(n),
               search_peaks_identical=bool(np.array_equal(Ud0[fin], U0.reshape(-1, 2)[fin])),
               search_max_px=float(np.abs(Ud0[fin] - U0.reshape(-1, 2)[fin]).max(initial=0.0)),
               icgn6_status_agree=round(float(np.mean(itd == it6o)), 5),
               icgn6_rms_px=float(np.sqrt(np.mean(d6 ** 2))),
               icgn6_max_px=float(np.abs(d6).max(initial=0.0)),
               icgn6_max_F=float(np.abs(dF).max(initial=0.0)))
    acc["gate_0p01px"] = bool(acc["icgn6_rms_px"] <= 0.01)
    return cpu, acc


def run_quadtree_sequence(args, pkg):
    """Workload 4 as an incremental sequence (main_ALDIC_Quadtree.m's frame loop with
    cumulative displacements): frames at far-field strain 0.05 %, 0.1 %, ... of the plate
    with a hole against one reference image, on the ~29k-node quadtree mesh; frame 1
    starts from the analytic field rounded to whole pixels, frame k from frame k-1's U
    (main_ALDIC.m:134-137).  Every frame is timed (wall clock around qframe_run)."""
    qt = importlib.import_module("2d_aldic_amd.quadtree")
    t = time.perf_counter()
    f, gs, ufuns, P = pkg.speckle.config4_sequence(
        frames=args.frames, progress=lambda k: log(f"[bench] frame {k + 1} rendered "
                                                   f"({time.perf_counter() - t:.0f} s)"))
    xy, el, dr, ws = P["coords"], P["elems"], P["dirichlet"], P["ws"]
    S = len(xy)
    log(f"[bench] workload 4 sequence generated in {time.perf_counter() - t:.1f} s: "
        f"{len(gs)} frames, {S} nodes, {len(el)} elements")
    ctx = pkg.Context(0)
    ctx.set_images(f, gs[0], mask=P["mask"])
    ctx.set_quadtree(xy, el, dr)
    u, v = ufuns[0](xy[:, 0], xy[:, 1])
    U0 = np.column_stack([np.rint(u), np.rint(v)]).ravel()
    _, _, gxy = ctx.nodal_strain_quadtree(U0)
    tri, w = qt.linear_weights(gxy, xy)
    mu = 1e-3
    beta = 1e-3 * 16.0 ** 2 * mu
    ctx.qframe_setup(ws, tri, w, mu=mu, beta=beta, n_admm=3, cg_maxit=5000, cg_rtol=1e-10)
    far = np.hypot(xy[:, 0] - P["hole"][0], xy[:, 1] - P["hole"][1]) > P["hole"][2] + 16
    for _ in range(args.warmup):          # warm-up on frame 1 (not part of the sequence)
        ctx.qframe_run(U0)
    ms, conv, rms = [], [], []
    Useed = U0
    for k, g in enumerate(gs):
        ctx.set_images(f, g, mask=P["mask"])
        t0 = time.perf_counter()
        ctx.qframe_run(Useed)
        r = ctx.qframe_get()
        ms.append((time.perf_counter() - t0) * 1e3)
        it6 = r["iters6"]
        conv.append(int(np.sum((it6 >= 1) & (it6 <= 100))))
        ut, vt = ufuns[k](xy[far, 0], xy[far, 1])
        rms.append(float(np.sqrt(np.mean((r["U"][0::2][far] - ut) ** 2 + (r["U"][1::2][far] - vt) ** 2))))
        log(f"[bench] frame {k + 1}: {ms[-1]:.3f} ms, converged {conv[-1]}, rms {rms[-1]:.4f} px")
        Useed = r["U"]
    total = sum(ms) / 1e3
    line = {
        "metric": "converged subsets/s at <=0.01 px RMS vs FP64; ADMM frame time (ms)",
        "value": round(sum(conv) / total, 2),
        "unit": "converged subsets/s",
        "n_gpus": 1, "steps": len(gs), "warmup": args.warmup,
        "ms_per_step": round(float(np.mean(ms)), 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"config4 sequence: {len(gs)} incremental frames of a 2048x2048 uint8 "
                               "plate with a hole (r 300, Kirsch, far-field strain 0.05 % per frame), "
                               f"quadtree mesh ({S} nodes, {len(el)} elements, cells 8-32 px), windows "
                               "16-64 px, each frame LocalICGN + 3 ADMM steps seeded with the previous U",
                   "subsets": S, "frames": len(gs),
                   "windows": {int(kk): int(cc) for kk, cc in zip(*np.unique(ws, return_counts=True))},
                   "frame_ms": [round(m, 4) for m in ms], "converged": conv,
                   "rms_err_vs_analytic_px": [round(e, 5) for e in rms],
                   "timing": "wall clock around each frame's qframe_run (images staged before)"},
        "roofline": None, "cpu_baseline": None, "accuracy_vs_fp64": None,
    }
    print(json.dumps(line), flush=True)


def run_quadtree(args, pkg):
    """Workload 4: the quadtree frame (main_ALDIC_Quadtree.m's LocalICGN + 3 ADMM steps)
    on a 2048^2 plate with a hole of radius 300 (Kirsch field, far-field strain 1 %),
    cells 8-32 px refined towards the hole edge, windows 2 x cell size (16-64 px).  The
    initial guess is the analytic field rounded to whole pixels (what the integer search
    or the previous incremental frame hands over); the search is not part of the frame."""
    qt = importlib.import_module("2d_aldic_amd.quadtree")
    t = time.perf_counter()
    pair = pkg.speckle.config4()
    P = pair.params
    xy, el, dr, ws = P["coords"], P["elems"], P["dirichlet"], P["ws"]
    S = len(xy)
    log(f"[bench] workload 4 generated in {time.perf_counter() - t:.1f} s: {S} nodes, "
        f"{len(el)} elements")
    if not args.no_cpu:
        log("[bench] workload 4: no CPU baseline (the FP64 restatement of the quadtree frame "
            "is the test oracle, pure Python)")
    ctx = pkg.Context(0)
    ctx.set_images(pair.f, pair.g, mask=P["mask"])
    ctx.set_quadtree(xy, el, dr)
    u, v = pair.ufun(xy[:, 0], xy[:, 1])
    U0 = np.column_stack([np.rint(u), np.rint(v)]).ravel()
    _, _, gxy = ctx.nodal_strain_quadtree(U0)
    tri, w = qt.linear_weights(gxy, xy)
    mu = 1e-3
    beta = 1e-3 * 16.0 ** 2 * mu
    ctx.qframe_setup(ws, tri, w, mu=mu, beta=beta, n_admm=3, cg_maxit=5000, cg_rtol=1e-10)
    for _ in range(args.warmup):
        ctx.qframe_run(U0)
    loc = 0.0
    pcg = 0
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ctx.qframe_run(U0)
        r = ctx.qframe_get()
        loc += r["local_ms"]
        pcg += r["pcg_iters"]
    wall = time.perf_counter() - t0
    ms = wall * 1e3 / args.steps
    it6 = r["iters6"]
    conv = int(np.sum((it6 >= 1) & (it6 <= 100)))
    far = np.hypot(xy[:, 0] - P["hole"][0], xy[:, 1] - P["hole"][1]) > P["hole"][2] + 16
    ut, vt = pair.ufun(xy[far, 0], xy[far, 1])
    err = np.hypot(r["U"][0::2][far] - ut, r["U"][1::2][far] - vt)
    npx = pair.f.size
    loc /= args.steps
    icgn_bytes = 16 * npx + 124 * S
    log(f"[bench] {ms:.3f} ms/frame (LocalICGN {loc:.3f} ms, {pcg / args.steps:.0f} PCG "
        f"iterations, budget {r['pcg_budget']}, re-runs {r['pcg_reruns']})")
    smooth = None
    if not args.no_smooth:
        # the reference's default flow: funSmoothStrainQuadtree at main_ALDIC_Quadtree.m:172,
        # 221, 283 (StrainSmoothness 1e-4, elementMinSize = the 8 px cells)
        ctx.qframe_setup(ws, tri, w, mu=mu, beta=beta, n_admm=3, cg_maxit=5000, cg_rtol=1e-10,
                         smooth_h=8.0, smoothness=1e-4)
        ctx.qframe_run(U0)
        ns = max(2, min(args.steps, 4))
        t1 = time.perf_counter()
        for _ in range(ns):
            ctx.qframe_run(U0)
        rs = ctx.qframe_get()
        es = np.hypot(rs["U"][0::2][far] - ut, rs["U"][1::2][far] - vt)
        smooth = dict(frame_ms=round((time.perf_counter() - t1) * 1e3 / ns, 3),
                      rms_err_vs_analytic_px=float(np.sqrt(np.mean(es ** 2))),
                      note="strain smoothing on (4 funSmoothStrainQuadtree calls per frame)")
        log(f"[bench] strain smoothing on: {smooth}")
    line = {
        "metric": "converged subsets/s at <=0.01 px RMS vs FP64; ADMM frame time (ms)",
        "value": round(conv / (ms / 1e3), 2),
        "unit": "converged subsets/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(ms, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "config4: 2048x2048 uint8 plate with a hole (r 300, Kirsch, 1 %), "
                               f"quadtree mesh ({S} nodes, {len(el)} elements, cells 8-32 px), "
                               "windows 16-64 px, LocalICGN + 3 ADMM steps",
                   "subsets": S, "converged": conv,
                   "windows": {int(k): int(c) for k, c in zip(*np.unique(ws, return_counts=True))},
                   "frame_ms": round(ms, 4), "local_icgn_ms": round(loc, 4),
                   "pcg_iters_per_frame": pcg // args.steps,
                   "pcg_budget": r["pcg_budget"], "pcg_reruns": r["pcg_reruns"],
                   "rms_err_vs_analytic_px": float(np.sqrt(np.mean(err ** 2))),
                   "strain_smoothing_on": smooth},
        "roofline": dict(kernel="k_icgn (6-DOF IC-GN, all window classes) + bad-point step",
                         bound="hbm", achieved=round(icgn_bytes / (loc / 1e3) / 1e9, 2),
                         peak=HBM_PEAK_GBS, unit="GB/s",
                         frac=round(icgn_bytes / (loc / 1e3) / 1e9 / HBM_PEAK_GBS, 5),
                         traffic=None, algorithmic_bytes=icgn_bytes, avg_ms=round(loc, 5)),
        "cpu_baseline": None,
        "accuracy_vs_fp64": None,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
