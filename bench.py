"""Benchmark: CG GLS map-making on one MI355X.

Workload (default): 16 detectors at angles k*pi/8, NSIDE 512, 2^26 samples per
detector in segments of 2^20 (1,073,741,824 TOD samples), spin/precession
scan, white + 1/f noise (fknee 0.1 Hz, index 1).  Synthetic inputs are
generated in host RAM first (untimed) and copied into pinned host buffers,
the form a caller hands to the library.

Three complete solves are timed, each from inputs in pinned host memory to
a true relative residual |b - A x| / |b| <= 1e-6 (H2D, pointing, ang2pix,
sort, b, preconditioner, CG; file I/O excluded):
  baseline   FP64 operator, plain 3x3 block-Jacobi PCG
  headline   FP32 operator with FP64 CG scalars and FP64 residual refresh,
             flexible PCG with the learned U-Net preconditioner (trained
             offline by sanepic_amd.train_precond on disjoint seeds) and the
             monopole deflation; block-Jacobi mixed PCG if no trained
             weights exist for the workload
  second     the headline's pipeline with block-Jacobi instead of the
             network ("block_jacobi_same_precision"), or with the network
             when the headline is not the learned solver ("learned")
Reported:
  value            TOD samples per second per CG iteration of the headline
                   solve (W warm-up iterations, then exactly K timed)
  time_to_1e-6_s   the headline solve's time to solution; the baseline's and
                   the ratio are in "baseline"
  roofline         the headline operator's dominant kernel: algorithmic bytes
                   per launch over its average launch duration (HIP events on
                   the library stream)
  cpu_baseline     the CPU oracle (an FP64 restatement of the reference's
                   path) on a bounded sample of the same workload
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12
CLASS_NAMES = ["gather+column FFT", "row FFT + noise filter", "inverse column FFT",
               "pixel-sorted P^T reduction"]


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def class_bytes_per_sample(precision):
    """Algorithmic HBM bytes per TOD sample of each operator kernel class
    (reads + writes of its streams; the map gathers hit the caches).
    FP64: pixel int32, (cos, sin) 2psi as two doubles, complex FP64 pair
    buffer (8 bytes per sample), FP64 z.  FP32: pixel int32, 2psi as a 32-bit
    fraction of a turn, complex FP32 pair buffer (4 bytes per sample), FP32 z."""
    fp64 = precision == "fp64"
    ang = 16 if fp64 else 4
    w = 8 if fp64 else 4
    z = 8 if fp64 else 4
    return [4 + ang + w,   # gather + column FFT: pixel, angle in; pair buffer out
            2 * w,         # row FFT + filter: pair buffer in and out
            w + z,         # inverse column FFT: pair buffer in, z out
            4 + z + ang]   # pixel-sorted P^T: permutation, z (gathered), sorted angle


def calibration_copy(nbytes=1 << 30):
    """A plain device-to-device copy of known size (HIP's blit kernel), run
    before the workload so hardware byte counters (FETCH_SIZE / WRITE_SIZE)
    can be scaled by what they read for exactly nbytes in and nbytes out."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(a), ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(ctypes.byref(b), ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMemset(a, 1, ctypes.c_size_t(nbytes)) == 0
    assert hip.hipDeviceSynchronize() == 0
    assert hip.hipMemcpy(b, a, ctypes.c_size_t(nbytes), 3) == 0  # device to device
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)
    log(f"calibration copy of {nbytes} bytes done")


def cpu_baseline(wl, budget_s=20.0):
    """The oracle on a bounded sample: one detector, a few segments of the
    same length, same NSIDE; timed over full A-applications."""
    import oracle
    from sanepic_amd import synth
    ncores = os.cpu_count() or 1
    ns = wl.get("ns", 2 ** 20)
    nseg = 4 if ns >= 2 ** 20 else 8
    cfg = synth.Config(nside=wl["nside"], ndet=1, nseg=nseg, ns=ns)
    ds = synth.make_dataset(cfg, workers=min(8, ncores))
    P = oracle.Problem(cfg.nside, 1, nseg, cfg.ns, nthreads=ncores)
    P.set_detector(0, 0.0, 0.0, 0.0, ds["ispf"][0])
    P.pixelize(ds["ra"], ds["dec"], ds["psi"])
    P.load_tod(ds["tod"])
    x = np.random.default_rng(0).standard_normal(3 * P.npix)
    P.apply_A(x)  # warm
    n, t0 = 0, time.perf_counter()
    while True:
        P.apply_A(x)
        n += 1
        if time.perf_counter() - t0 > budget_s or n >= 20:
            break
    dt = (time.perf_counter() - t0) / n
    nsamp = nseg * cfg.ns
    return {"value": nsamp / dt, "unit": "samples/s per CG iteration (one A-application)",
            "cores": ncores, "kind": "port",
            "sample": f"1 detector x {nseg} segments of {cfg.ns} samples, NSIDE {cfg.nside}, "
                      f"{n} A-applications timed ({dt * 1e3:.1f} ms each)"}


def run_solve(cfg, pin, precision, precond, deflate, weights, steps, warmup, maxiter, profile, net_nside=None,
              refresh_fp32=False, refresh=10):
    """One timed pipeline: create -> uploads -> pixelize -> setup -> solve to a
    true residual <= 1e-6.  Returns timings, per-iteration time and kernel
    times of `steps` iterations measured after `warmup`."""
    import sanepic_amd
    prec = sanepic_amd.FP64 if precision == "fp64" else sanepic_amd.MIXED
    pc = {"block": sanepic_amd.PRECOND_BLOCK3, "diag": sanepic_amd.PRECOND_DIAG,
          "learned": sanepic_amd.PRECOND_LEARNED}[precond]
    un = None
    if precond == "learned":
        un = sanepic_amd.UNet(net_nside or cfg.nside)
        un.set_params(weights)  # loading the trained network is not part of the solve
    t0 = time.perf_counter()
    if hasattr(cfg, "seglen"):  # per-detector segment lengths
        ctx = sanepic_amd.Context.uneven(cfg.nside, cfg.seglen, cfg.nsamp, precision=prec)
        ctx.load_uneven(cfg, pin, queued=True)
    else:
        ctx = sanepic_amd.Context(cfg.nside, cfg.ndet, cfg.nseg, cfg.ns, precision=prec)
        ctx.load_async(cfg, pin)
    t_up = time.perf_counter()
    npix, ndef = ctx.pixelize()
    t_pix = time.perf_counter()
    ctx.setup()
    t_setup = time.perf_counter()
    if un is not None:
        ctx.learn_attach(un)
    ctx.solve_begin(pc, refresh, None, deflate, refresh_fp32=refresh_fp32)
    t_begin = time.perf_counter()
    # time to solution: iterate until crit <= 1e-12, then confirm the true
    # residual with an explicit A x
    hist = []
    converged, true_res = False, None
    while len(hist) < maxiter:
        hist.append(ctx.step())
        if hist[-1]["crit"] <= 1e-12:
            true_res = ctx.true_residual()
            if true_res <= 1e-6:
                converged = True
                break
    t_end = time.perf_counter()
    if true_res is None:
        true_res = ctx.true_residual()
    lstats = ctx.learn_stats() if un is not None else None
    # per-iteration time: a fresh start, `warmup` iterations, then exactly
    # `steps` iterations with the kernel timers on
    kt, ms_per_step = None, None
    if profile:
        ctx.solve_begin(pc, refresh, None, deflate, refresh_fp32=refresh_fp32)
        for _ in range(warmup):
            ctx.step()
        ctx.set_profiling(True)
        ts = time.perf_counter()
        for _ in range(steps):
            ctx.step()
        ms_per_step = (time.perf_counter() - ts) * 1e3 / steps
        kt = [ctx.kernel_time(k) for k in range(5)]
        ctx.set_profiling(False)
    its = len(hist)  # the first iteration whose true residual confirmed 1e-6
    out = dict(precision=precision, precond=precond, deflate=deflate, npix=npix, deferred=ndef,
               upload_s=t_up - t0, pixelize_s=t_pix - t_up, setup_s=t_setup - t_pix,
               begin_s=t_begin - t_setup, solve_s=t_end - t_begin, iterations=its, converged=converged,
               final_true_residual=true_res,
               time_to_solution_s=(t_end - t0) if converged else None,
               device_bytes=ctx.device_bytes(), ms_per_step=ms_per_step, kernel_times=kt)
    if un is not None:
        out["learned_stats"] = lstats
    log(f"[{precision} {precond}{' +defl' if deflate else ''}] upload {out['upload_s']:.2f} s, "
        f"pixelize {out['pixelize_s']:.2f} s, setup {out['setup_s']:.2f} s, begin {out['begin_s']:.2f} s; "
        f"{its} iterations, true residual {true_res:.3e}, time to solution "
        f"{out['time_to_solution_s'] if converged else float('nan'):.3f} s"
        + (f", {ms_per_step:.2f} ms/iteration" if ms_per_step else "")
        + (f", {out['learned_stats']}" if un is not None else ""))
    ctx.close()
    if un is not None:
        un.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="c1")
    ap.add_argument("--precision", default="mixed", choices=["fp64", "mixed"])
    ap.add_argument("--precond", default="learned", choices=["learned", "block", "diag"])
    ap.add_argument("--no-second", action="store_true",
                    help="skip the second pipeline (mixed block-Jacobi beside a learned headline, the "
                         "learned solver beside any other)")
    ap.add_argument("--no-deflate", action="store_true")
    ap.add_argument("--mixed-refresh", type=int, default=20,
                    help="FP64 explicit residual every this many iterations in the mixed-precision pipelines "
                         "(the FP64 baseline keeps the reference's 10)")
    ap.add_argument("--no-baseline", action="store_true", help="skip the FP64 block-Jacobi solve")
    ap.add_argument("--maxiter", type=int, default=600)
    ap.add_argument("--settle", type=float, default=5.0,
                    help="idle seconds before each pipeline after the first, outside the timed regions: a "
                         "context created right after another released ~100 GB of device memory has stalled "
                         "for up to 3 s inside its set-up, so every pipeline starts from an idle device")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--weights", default=None, help="trained network (default: the workload's committed one)")
    ap.add_argument("--calib-copy", action="store_true",
                    help="run a 1 GiB device copy first (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: there is no multi-GPU path (sharding the (detector, "
              "segment-pair) blocks over GPUs is a design, not built)", file=sys.stderr)
        sys.exit(2)

    import sanepic_amd
    from sanepic_amd import synth, train_precond

    if args.calib_copy:
        calibration_copy()
    uneven = args.workload in synth.UNEVEN_WORKLOADS
    wl = synth.UNEVEN_WORKLOADS[args.workload] if uneven else synth.WORKLOADS[args.workload]
    cfg = synth.workload_config(args.workload)
    tg = time.perf_counter()
    workers = min(16, os.cpu_count() or 1)
    ds = synth.make_uneven_dataset(cfg, workers) if uneven else synth.make_dataset(cfg, workers)
    pin = sanepic_amd.PinnedInputs(cfg, ds)
    del ds
    ntot = sum(cfg.used(d) for d in range(cfg.ndet)) if uneven else cfg.ndet * cfg.nseg * cfg.ns
    log(f"inputs generated in {time.perf_counter() - tg:.1f} s "
        f"({cfg.ndet} det, {ntot} samples, NSIDE {cfg.nside}), in pinned host memory")

    precond, weights, record = args.precond, None, None
    path = args.weights or train_precond.weights_path(args.workload)
    if os.path.exists(path):
        weights, record = train_precond.load_weights(path)
    if precond == "learned" and weights is None:
        log(f"no trained network at {path}: block-Jacobi instead")
        precond = "block"
    deflate = not args.no_deflate

    base = None
    if not args.no_baseline:
        base = run_solve(cfg, pin, "fp64", "block", False, None, args.steps, args.warmup,
                         args.maxiter, profile=False)
    rf = args.mixed_refresh if args.precision == "mixed" else 10
    if base is not None:
        time.sleep(args.settle)
    head = run_solve(cfg, pin, args.precision, precond, deflate, weights, args.steps, args.warmup,
                     args.maxiter, profile=True, net_nside=(record or {}).get("net_nside"), refresh=rf)
    learned = second_block = None
    if not args.no_second:
        time.sleep(args.settle)
        if precond != "learned" and weights is not None:
            # the same pipeline with the trained network (reported beside the headline)
            learned = run_solve(cfg, pin, args.precision, "learned", deflate, weights, 2, 1, args.maxiter,
                                profile=False, net_nside=record.get("net_nside"), refresh=rf)
        elif precond == "learned":
            # the same pipeline with block-Jacobi: what the network buys
            second_block = run_solve(cfg, pin, args.precision, "block", deflate, None, 2, 1, args.maxiter,
                                     profile=False, refresh=rf)
    pin.close()

    kt = head["kernel_times"]
    ms_per_step = head["ms_per_step"]
    per = [(kt[k][0], kt[k][1], k) for k in range(4) if kt[k][1] > 0]
    tot_ms, cnt, dom = max(per)
    # per operator application: with per-detector lengths one application
    # launches each kernel class once per detector group
    n_ops = max(kt[4][1], 1)
    avg_ms = tot_ms / n_ops
    cbs = class_bytes_per_sample(args.precision)
    alg = ntot * cbs[dom]  # algorithmic bytes of the dominant kernel class per operator application
    achieved = alg / (avg_ms * 1e-3)
    for k in range(4):
        if kt[k][1]:
            log(f"  {CLASS_NAMES[k]}: {kt[k][0] / n_ops:.3f} ms per operator ({kt[k][1]} launches)")
    apply_ms = kt[4][0] / max(kt[4][1], 1)
    s = 8 if args.precision == "fp64" else 4
    lens = cfg.seglen if uneven else [cfg.ns] * cfg.ndet
    b_alg = ntot * sum(cbs) + sum(L // 2 + 1 for L in lens) * s + 12 * 3 * head["npix"] * 8
    tts = head["time_to_solution_s"]
    out = {
        "metric": "time-to-1e-6 residual (s) + TOD samples/s per CG iter, NSIDE512 16det, 1 GPU",
        "value": ntot / (ms_per_step * 1e-3),
        "unit": "TOD samples/s per CG iteration",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32 operator, f64 CG scalars",
        "data": "synthetic (spin/precession scan, Gaussian I/Q/U sky, white + 1/f noise)",
        "config": {"workload": (f"{cfg.ndet} det, {ntot} samples, NSIDE {cfg.nside}, per-detector "
                                f"segment lengths m 2^l (m in 1,3,5), {precond} PCG" if uneven else
                                f"{cfg.ndet} det x {cfg.nseg * cfg.ns} samples, NSIDE {cfg.nside}, "
                                f"segments of {cfg.ns}, {precond} PCG")
                               + (" + monopole deflation" if deflate else ""),
                   "nside": cfg.nside, "ndet": cfg.ndet, "samples": ntot,
                   "segment": sorted(set(lens)) if uneven else cfg.ns, "npix": head["npix"], "precond": precond,
                   "deflate": deflate, "precision": args.precision, "fp64_refresh_every": rf},
        "time_to_1e-6_s": tts,
        "iterations_to_1e-6": head["iterations"] if head["converged"] else None,
        "final_true_residual": head["final_true_residual"],
        "setup_s": head["upload_s"] + head["pixelize_s"] + head["setup_s"] + head["begin_s"],
        "phases_s": {k: head[k] for k in ("upload_s", "pixelize_s", "setup_s", "begin_s")},
        "idle_before_each_pipeline_s": args.settle,
        "operator_ms": apply_ms,
        "operator_frac_hbm": b_alg / (apply_ms * 1e-3) / HBM_PEAK,
        "device_bytes": head["device_bytes"],
        "roofline": {"bound": "hbm", "kernel": CLASS_NAMES[dom], "achieved": achieved / 1e9,
                     "peak": HBM_PEAK / 1e9, "unit": "GB/s", "frac": achieved / HBM_PEAK,
                     "traffic": None, "bytes_per_operator": alg, "ms_per_operator": avg_ms,
                     "launches_per_operator": cnt // n_ops},
    }
    lrun = head if "learned_stats" in head else learned
    if lrun is not None:
        out["learned"] = {"solver": f"{args.precision} operator, flexible PCG with the trained network"
                                    + (" + monopole deflation" if deflate else ""),
                          "time_to_1e-6_s": lrun["time_to_solution_s"],
                          "iterations_to_1e-6": lrun["iterations"],
                          "ms_per_iteration_in_solve": lrun["solve_s"] * 1e3 / max(lrun["iterations"], 1),
                          "stats": lrun["learned_stats"],
                          "training": {k: record.get(k) for k in (
                              "net_nside", "pairs", "steps", "rounds", "round_pairs", "round_steps", "loss", "noise_gain",
                              "harvest_seeds", "on_policy_seeds", "holdout_block_jacobi", "holdout_trained",
                              "validation_block_jacobi_iterations", "validation_iterations", "kept_phase",
                              "total_seconds")}}
    if second_block is not None:
        out["block_jacobi_same_precision"] = {
            "solver": f"{args.precision} operator, block-Jacobi PCG" + (" + monopole deflation" if deflate else ""),
            "time_to_1e-6_s": second_block["time_to_solution_s"],
            "iterations_to_1e-6": second_block["iterations"],
            "ms_per_iteration_in_solve": second_block["solve_s"] * 1e3 / max(second_block["iterations"], 1)}
    if base is not None:
        out["baseline"] = {"solver": "FP64 operator, plain block-Jacobi PCG",
                           "time_to_1e-6_s": base["time_to_solution_s"],
                           "iterations_to_1e-6": base["iterations"],
                           "setup_s": base["upload_s"] + base["pixelize_s"] + base["setup_s"] + base["begin_s"],
                           "speedup_time_to_solution": (base["time_to_solution_s"] / tts
                                                        if tts and base["time_to_solution_s"] else None)}
    if not args.no_cpu:
        try:
            out["cpu_baseline"] = cpu_baseline(wl)
        except Exception as e:  # the GPU result stands on its own
            out["cpu_baseline"] = {"error": str(e)}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
