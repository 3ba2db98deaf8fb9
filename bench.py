"""Benchmark: one AL-DIC frame on one MI355X.

Workload (one "step" = one frame): 2048x2048 uint8 Gaussian-dot speckle (150k dots,
sigma 2.5 px), sinusoidal field u = 2 sin(2 pi X/256), v = 0.5 cos(2 pi Y/512), noise
sigma 2, subsets 32 px on a 16 px grid (14,884 subsets), integer search radius 8:
normalisation + gradients, integer search, 6-DOF IC-GN with the bad-point rule,
step-1 global step, and 3 ADMM iterations (2-DOF IC-GN, Q4 FEM global step, nodal
strain, dual update).  The raw pair is resident in HBM before the timed region.

value = converged subsets (6-DOF IC-GN status 1..100) per second of whole frame.
The CPU baseline is the FP64 restatement of the same frame on the host cores.

  python bench.py [--steps K] [--warmup W] [--no-cpu] [--no-stages]
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
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-stages", action="store_true",
                    help="skip the untimed per-stage breakdown frames")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--workload", type=int, default=2, choices=(2, 3, 4),
                    help="2: 2048^2 uint8 ws32/step16 s8 (default); 3: 4096^2 12-bit "
                         "ws32/step8 s24, ~247k subsets (generation takes ~1.5 min); 4: "
                         "2048^2 plate with a hole on a quadtree mesh, windows 16-64 px")
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
        return run_quadtree(args, pkg)
    t = time.perf_counter()
    if args.workload == 3:
        if not args.no_cpu:
            log("[bench] workload 3: the FP64 CPU frame would take hours; CPU baseline skipped")
            args.no_cpu = True
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
    # timed frames carry events only around the search and the whole frame; the
    # per-stage breakdown comes from a second, untimed set of frames with an event at
    # every stage boundary
    S = ctx.frame_setup(ws, step, srad, grid, tol=1e-2, mu=mu, beta=beta, n_admm=n_admm,
                        cg_maxit=2000, cg_rtol=1e-10, stage_times=False)
    log(f"[bench] frame set up: {S} subsets")
    for _ in range(args.warmup):
        ctx.frame_run()
    log("[bench] warmup done")
    search_ms = icgn6_ms = 0.0
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ctx.frame_run()
        t = ctx.frame_times()
        search_ms += t["search"]
        icgn6_ms += t["icgn6"]
    wall = time.perf_counter() - t0
    search_ms /= args.steps
    icgn6_ms /= args.steps
    ms_per_step = wall * 1e3 / args.steps
    out = ctx.frame_get()
    counters = ctx.frame_counters()
    stage = {}
    if not args.no_stages:
        ctx.frame_setup(ws, step, srad, grid, tol=1e-2, mu=mu, beta=beta, n_admm=n_admm,
                        cg_maxit=2000, cg_rtol=1e-10, stage_times=True)
        for _ in range(args.warmup + args.steps):
            ctx.frame_run()
            for k, v in ctx.frame_times().items():
                stage[k] = stage.get(k, 0.0) + v
        stage = {k: v / (args.warmup + args.steps) for k, v in stage.items()}
    it6 = out["iters6"]
    conv = int(np.sum((it6 >= 1) & (it6 <= 100)))
    value = conv / (ms_per_step / 1e3)
    log(f"[bench] {ms_per_step:.3f} ms/frame, stages {stage}")

    # Roofline of the dominant kernel, the 6-DOF IC-GN (one k_icgn launch per frame plus
    # the device-listed relaunch of enlarged windows, event-timed on the library's
    # stream over the timed frames).  Bandwidth class: achieved = compulsory HBM bytes
    # per launch / time, with 16 N_px (f, fx, fy, g as FP32) + 124 S (2 words in; 6 + 21
    # + 2 words out, 4 B each) per launch.  "traffic" is the measured HBM bytes per
    # launch (rocprofv3 FETCH_SIZE x 2.0 + WRITE_SIZE x 1.0, the factors a 256 MiB
    # device copy reads on the same counters), kept in profiles/hbm_traffic.json.
    npx = side * side
    traffic = {}
    try:
        with open(TRAFFIC_FILE) as fh:
            traffic = json.load(fh)
    except (OSError, ValueError):
        pass

    def measured(key):
        return traffic.get(key, {}).get("hbm_bytes_per_launch")

    icgn_bytes = 16 * npx + 124 * S
    t_i = icgn6_ms / 1e3
    if args.workload != 2:
        traffic = {}  # the measured table is for the default workload
    roof = dict(kernel="k_icgn (6-DOF IC-GN, ws 32)", bound="hbm",
                achieved=round(icgn_bytes / t_i / 1e9, 2), peak=HBM_PEAK_GBS, unit="GB/s",
                frac=round(icgn_bytes / t_i / 1e9 / HBM_PEAK_GBS, 5),
                traffic=measured("k_icgn_6dof"), algorithmic_bytes=icgn_bytes,
                avg_ms=round(icgn6_ms, 5),
                note="VALU- and latency-bound (FP32 Keys gathers and sums, 66 VALU per pixel pass)",
                copy_gbs_measured=measure_copy_bw())
    # VALU issue share: SQ_INSTS_VALU per launch (rocprofv3 counter run) x 4 cycles per
    # wave64 VALU instruction / (1024 SIMDs x 2.4 GHz x the live average duration)
    vi = traffic.get("k_icgn_6dof", {}).get("valu_instr_per_launch")
    roof["valu_frac"] = round(vi * 4 / (1024 * VALU_CLOCK_HZ * t_i), 4) if vi else None
    # The integer search beside it: int8 matrix cores, algorithmic ops = the NCC
    # numerator's multiply-adds, 2 (2s+3)^2 (w+1)^2 per subset
    ops = 2.0 * (2 * srad + 3) ** 2 * (ws + 1) ** 2 * S
    t_s = search_ms / 1e3
    if pair.f.dtype in (np.uint8, np.uint16):
        small = pair.f.dtype == np.uint8 and srad <= 14
        bpp = pair.f.dtype.itemsize
        # raw pair once + outputs (+ the window box-sum planes for the multi-tile kernel)
        search_bytes = npx * 2 * bpp + 32 * S + (0 if small else npx * 36)
        search_roof = dict(kernel="k_search_i8" if small else "k_search_mt", bound="mfma",
                           engine="int8 MFMA (32x32x32)" + ("" if bpp == 1 else
                                                            ", two digit planes (4 GEMMs)"),
                           achieved=round(ops / t_s / 1e12, 3), peak=INT8_PEAK_TOPS,
                           unit="TOP/s", frac=round(ops / t_s / 1e12 / INT8_PEAK_TOPS, 5),
                           traffic=measured("k_search_i8") if small else None,
                           algorithmic_ops=ops, algorithmic_bytes=search_bytes,
                           avg_ms=round(search_ms, 5))
        # the kernel is bound by VALU issue, not the matrix cores: its VALU share as for
        # IC-GN (counter run of the default workload only)
        svi = traffic.get("k_search_i8", {}).get("valu_instr_per_launch") if small else None
        search_roof["valu_frac"] = (round(svi * 4 / (1024 * VALU_CLOCK_HZ * t_s), 4)
                                    if svi else None)
    else:
        search_bytes = npx * 8 + 32 * S  # the FP32 pair once + outputs
        search_roof = dict(kernel="k_search", bound="mfma",
                           engine="VALU FP32 FMA (packed FP32 off)",
                           achieved=round(ops / t_s / 1e12, 3), peak=FP32_PEAK_TFLOPS,
                           unit="TFLOP/s", frac=round(ops / t_s / 1e12 / FP32_PEAK_TFLOPS, 5),
                           traffic=None, algorithmic_ops=ops, algorithmic_bytes=search_bytes,
                           avg_ms=round(search_ms, 5))

    cpu = None
    acc = {}
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        os.environ["OMP_NUM_THREADS"] = str(threads)
        import oracle
        oracle.build()
        log(f"[bench] CPU baseline: FP64 oracle frame on {threads} threads")
        o = oracle.frame(pair.f, pair.g, ws, step, srad, grid, tol=1e-2, mu=mu, beta=beta,
                         n_admm=n_admm)
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
                   "stage_ms": {k: round(v, 4) for k, v in stage.items()},
                   "stage_ms_from": "separate untimed frames with an event at every stage boundary",
                   "gridfit": counters,
                   "icgn6_subsets_per_s": round(conv / (icgn6_ms / 1e3), 1)},
        "roofline": roof,
        "search_roofline": search_roof,
        "cpu_baseline": cpu,
        "accuracy_vs_fp64": acc or None,
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
                   "rms_err_vs_analytic_px": float(np.sqrt(np.mean(err ** 2)))},
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
