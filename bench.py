"""Benchmark: CG GLS map-making on one MI355X.

Workload (default): 16 detectors at angles k*pi/8, NSIDE 512, 2^26 samples per
detector in segments of 2^20 (1,073,741,824 TOD samples), spin/precession
scan, white + 1/f noise (fknee 0.1 Hz, index 1).  Synthetic inputs are
generated in host RAM first (untimed) and copied into pinned host buffers,
the form a caller hands to the library.

Complete solves are timed, each from inputs in pinned host memory to a true
relative residual |b - A x| / |b| <= 1e-6 (H2D, pointing, ang2pix, sort, b,
preconditioner, CG; file I/O excluded), after one small untimed solve that
loads the library's code onto the device and one untimed context of the
workload's size (created, loaded, destroyed) that takes the process's
first-context cost out of the first pipeline:
  headline   FP64 operator, plain 3x3 block-Jacobi PCG (the reference's
             arithmetic); after its time to solution it iterates on, untimed,
             to |r|^2/|b|^2 <= 1e-24 (a true residual near 1e-13) for the
             reference map of the error figures: at the reference's own stop
             (1e-15, a true residual near 3e-8) Q/U still carry ~5e-6 of error
  learned    FP32 operator with FP64 CG scalars, reductions, 3x3 inverses and
             FP64 residual refresh; flexible PCG with the learned U-Net
             preconditioner (trained offline by sanepic_amd.train_precond on
             disjoint seeds) and the monopole deflation
  block_jacobi_same_precision   the learned pipeline with block-Jacobi
  learned_literal_fallback      the learned pipeline whose network steps
             fall back to block-Jacobi whenever |r| does not decrease
Reported:
  value            TOD samples per second per CG iteration of the FP64 solve
                   (W warm-up iterations, then exactly K timed)
  time_to_1e-6_s   the FP64 solve's time to solution; each other pipeline has
                   its own, its ratio to the FP64 one, and its I/Q/U map error
                   against the FP64 reference map (relative RMS over seen
                   pixels, Q/U over unmasked ones, I monopole projected out)
  roofline         the FP64 operator's dominant kernel: algorithmic bytes per
                   launch over its average launch duration (HIP events on the
                   library stream); traffic from the hardware byte counters in
                   profiles/
  cpu_baseline     the CPU oracle (an FP64 restatement of the reference's
                   path) on bounded samples of the same workload: throughput
                   per CG iteration, and a complete block-Jacobi solve to 1e-6
                   beside the GPU's solve of the same sample
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
# the same classes when P^T runs over run records (the inverse pass sums runs)
CLASS_NAMES_RUNS = ["gather+column FFT", "row FFT + noise filter", "inverse column FFT + run sums",
                    "P^T over run records"]


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def class_bytes_per_sample(precision, rps=0.0, ang64=16, sorted64=16):
    """Algorithmic HBM bytes per TOD sample of each operator kernel class
    (reads + writes of its streams; the map gathers hit the caches).
    FP64: pixel int32, (cos, sin) 2psi as two doubles, complex FP64 pair
    buffer (8 bytes per sample), FP64 z.  FP32: pixel int32, 2psi as a 32-bit
    fraction of a turn, complex FP32 pair buffer (4 bytes per sample), FP32 z.
    rps > 0: P^T over run records, rps records per sample; the inverse pass
    reads the angles and one 8-byte word per 16-sample piece and writes three
    sums per record, P^T reads a 4-byte index and the record.  ang64: bytes of
    the FP64 angle in the gather and record passes (8 as a 64-bit turn, 16 as
    a pair); sorted64: bytes of the sorted angle that the FP64 P^T over
    samples reads (8 as a 64-bit turn, 16 as a pair)."""
    fp64 = precision == "fp64"
    ang = ang64 if fp64 else 4
    w = 8 if fp64 else 4
    z = 8 if fp64 else 4
    if rps:
        return [4 + ang + w, 2 * w, w + ang + 8 / 16 + 3 * z * rps, (4 + 3 * z) * rps]
    return [4 + ang + w,   # gather + column FFT: pixel, angle in; pair buffer out
            2 * w,         # row FFT + filter: pair buffer in and out
            w + z,         # inverse column FFT: pair buffer in, z out
            4 + z + (sorted64 if fp64 else 4)]  # pixel-sorted P^T: permutation, z (gathered), sorted angle


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


def cpu_baseline(wl, budget_s=20.0, settle=5.0):
    """The oracle on bounded samples of the same workload (same NSIDE and
    segment length), on every host core:
    * throughput: one detector, a few segments, timed over full
      A-applications (the work of one CG iteration);
    * time to solution: two detectors (one alone leaves Q and U of most
      pixels undetermined), 4 segments, block-Jacobi PCG from inputs in host
      memory to a true residual of 1e-6, beside the GPU's FP64 solve of the
      same inputs in this run."""
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
    del P, ds
    out = {"value": nsamp / dt, "unit": "samples/s per CG iteration (one A-application)",
           "cores": ncores, "kind": "port",
           "sample": f"1 detector x {nseg} segments of {cfg.ns} samples, NSIDE {cfg.nside}, "
                     f"{n} A-applications timed ({dt * 1e3:.1f} ms each)"}
    # time to solution, CPU and GPU, on one sample
    import sanepic_amd
    cfg2 = synth.Config(nside=wl["nside"], ndet=2, nseg=nseg, ns=ns, det_angles_deg=[0.0, 112.5])
    ds2 = synth.make_dataset(cfg2, workers=min(8, ncores))
    t0 = time.perf_counter()
    P = oracle.Problem(cfg2.nside, 2, nseg, ns, nthreads=ncores)
    for k in range(2):
        P.set_detector(k, 0.0, 0.0, cfg2.det_angles_deg[k], ds2["ispf"][k])
    P.pixelize(ds2["ra"], ds2["dec"], ds2["psi"])
    P.load_tod(ds2["tod"])
    sol = P.solve_pcg(kind=1, deflate=False, stop=1e-12, maxiter=600)
    npix = P.npix
    ax = P.apply_A(sol["S"])
    m = sol["hits"][npix:] < 4
    ax[npix:2 * npix][m] = 0.0
    ax[2 * npix:][m] = 0.0
    res = float(np.linalg.norm(sol["b"] - ax) / np.linalg.norm(sol["b"]))
    cpu_s = time.perf_counter() - t0
    del P
    pin = sanepic_amd.PinnedInputs(cfg2, ds2)
    time.sleep(settle)
    g = run_solve(cfg2, pin, "fp64", "block", False, None, 0, 0, 600, profile=False)
    pin.close()
    out["time_to_1e-6"] = {
        "sample": f"2 detectors x {nseg} segments of {ns} samples, NSIDE {cfg2.nside}, FP64 block-Jacobi PCG",
        "cpu_s": cpu_s, "cpu_iterations": sol["niter"], "cpu_true_residual": res, "cpu_cores": ncores,
        "gpu_s": g["time_to_solution_s"], "gpu_iterations": g["iterations"],
        "gpu_true_residual": g["final_true_residual"],
        "gpu_speedup": (cpu_s / g["time_to_solution_s"]) if g["time_to_solution_s"] else None}
    return out


def read_traffic(workload, precision, kernel):
    """HBM bytes per launch of a kernel class on a workload, measured with
    the hardware byte counters (rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE,
    each scaled by a device copy of known size) and recorded in
    profiles/*.md as lines
    'traffic <workload> <precision> <kernel class>: <bytes> bytes per launch'."""
    import glob
    import re
    pat = re.compile(r"^traffic (\S+) (\S+) (.+): (\d+) bytes per launch", re.M)
    for fn in sorted(glob.glob(os.path.join(ROOT, "profiles", "*.md"))):
        for wl, prec, k, v in pat.findall(open(fn).read()):
            if wl == workload and prec == precision and k.strip() == kernel:
                return int(v), os.path.relpath(fn, ROOT)
    return None, None


def warm_runtime():
    """One small untimed solve, so that loading the library's code onto the
    device (the first kernel launch of a process) is in no pipeline's time."""
    import sanepic_amd
    from sanepic_amd import synth
    cfg = synth.Config(nside=8, ndet=1, nseg=2, ns=1024)
    ds = synth.make_dataset(cfg, workers=1)
    for prec in (sanepic_amd.FP64, sanepic_amd.MIXED):
        with sanepic_amd.Context(cfg.nside, cfg.ndet, cfg.nseg, cfg.ns, precision=prec) as ctx:
            ctx.load(cfg, ds)
            ctx.pixelize()
            ctx.setup()
            ctx.solve(precond=sanepic_amd.PRECOND_BLOCK3, maxiter=3)
    log("runtime warmed up (small untimed solve)")


def map_errors(x, ref, cnt):
    """Relative RMS error of each component of map x (block layout) against
    ref over the seen pixels; Q and U over the pixels they are solved for
    (count >= 2), I with the monopole projected out (the DC bin's 1e-6 weight
    leaves it nearly free, sanepic.cc:537)."""
    npix = cnt.size
    ok = cnt >= 2
    out = {}
    for c, name in enumerate("IQU"):
        a, b = x[c * npix:(c + 1) * npix], ref[c * npix:(c + 1) * npix]
        if c == 0:
            a, b = a - a.mean(), b - b.mean()
        else:
            a, b = a[ok], b[ok]
        out[name] = float(np.linalg.norm(a - b) / np.linalg.norm(b))
    return out


def run_solve(cfg, pin, precision, precond, deflate, weights, steps, warmup, maxiter, profile, net_nside=None,
              refresh_fp32=False, refresh=10, fallback_growth=0.0, keep_map=False, reference=False,
              track=None, track_self=True):
    """One timed pipeline: create -> uploads -> pixelize -> setup -> solve to a
    true residual <= 1e-6.  Returns timings, per-iteration time and kernel
    times of `steps` iterations measured after `warmup`; with keep_map the
    map at the time of solution, with reference also the map after iterating
    on (untimed) to |r|^2/|b|^2 <= 1e-24 (and that map tracked as below).
    With track = (reference map,
    counts), an untimed second solve reads the map back after every
    iteration and records the first iteration whose I, Q and U are all
    within 1e-6 relative RMS of the reference (map_errors)."""
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
    ctx.solve_begin(pc, refresh, None, deflate, fallback_growth=fallback_growth, refresh_fp32=refresh_fp32)
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
    x_sol = ctx.map() if keep_map or reference else None
    x_ref, ref_iters, ref_res = None, None, None
    if reference:  # untimed: on well past the reference's own stop rule (1e-15, 1037, 1415)
        n, crit = 0, 1.0
        while n < 300 and crit > 1e-24:
            n += 1
            crit = ctx.step()["crit"]
        x_ref, ref_iters, ref_res = ctx.map(), len(hist) + n, ctx.true_residual()
        cnt_ref = ctx.get_setup()[0][:ctx.npix] // 2
        if track is None and track_self:
            track = (x_ref, cnt_ref)
    map_its, map_err = None, None
    if track is not None:
        ctx.solve_begin(pc, refresh, None, deflate, fallback_growth=fallback_growth, refresh_fp32=refresh_fp32)
        for k in range(1, min(maxiter, 200) + 1):
            ctx.step()
            map_err = map_errors(ctx.map(), track[0], track[1])
            if max(map_err.values()) <= 1e-6:
                map_its = k
                break
    # per-iteration time: a fresh start, `warmup` iterations, then exactly
    # `steps` iterations with the kernel timers on
    kt, ms_per_step = None, None
    if profile:
        ctx.solve_begin(pc, refresh, None, deflate, fallback_growth=fallback_growth, refresh_fp32=refresh_fp32)
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
               device_bytes=ctx.device_bytes(), ms_per_step=ms_per_step, kernel_times=kt,
               run_records=ctx.run_records(), fp64_angle_bytes=ctx.fp64_angle_bytes(),
               fp64_sorted_angle_bytes=ctx.fp64_sorted_angle_bytes())
    if un is not None:
        out["learned_stats"] = lstats
    out["map"] = x_sol
    out["map_error_iterations"], out["map_error_last"] = map_its, map_err
    if reference:
        out.update(ref_map=x_ref, ref_iterations=ref_iters, ref_true_residual=ref_res, cnt=cnt_ref)
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
    ap.add_argument("--workload", default="nside512_16det")
    ap.add_argument("--no-learned", action="store_true", help="only the FP64 solve")
    ap.add_argument("--no-literal", action="store_true",
                    help="skip the learned pipeline with the literal fallback rule (growth 1)")
    ap.add_argument("--no-deflate", action="store_true", help="mixed pipelines without the monopole deflation")
    ap.add_argument("--mixed-refresh", type=int, default=40,
                    help="FP64 explicit residual every this many iterations in the mixed-precision pipelines "
                         "(the FP64 solve keeps the reference's 10); the solve is confirmed by an explicit "
                         "FP64 residual in any case (measured: 20 -> 40 leaves the iteration counts and "
                         "map errors unchanged and saves one FP64 operator per solve)")
    ap.add_argument("--maxiter", type=int, default=600)
    ap.add_argument("--settle", type=float, default=5.0,
                    help="idle seconds before each pipeline after the first, outside the timed regions: a "
                         "context created right after another released ~100 GB of device memory has stalled "
                         "for up to 3 s inside its set-up, so every pipeline starts from an idle device")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-map-track", action="store_true",
                    help="skip the untimed re-solves that find the iteration where the maps reach 1e-6 of the "
                         "FP64 reference")
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
    warm_runtime()
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
    # Untimed: the process's first context of this size pays ~50 ms more in
    # creation and upload than every later one (measured: not the device
    # allocations, not the pinned pages); one created, loaded and destroyed
    # here gives every pipeline the same start.
    tw = time.perf_counter()
    if uneven:
        with sanepic_amd.Context.uneven(cfg.nside, cfg.seglen, cfg.nsamp) as w:
            w.load_uneven(cfg, pin, queued=True)
    else:
        with sanepic_amd.Context(cfg.nside, cfg.ndet, cfg.nseg, cfg.ns) as w:
            w.load_async(cfg, pin)
    log(f"first-context warm-up (create, upload, destroy) {time.perf_counter() - tw:.2f} s")

    weights, record = None, None
    path = args.weights or train_precond.weights_path(args.workload)
    if os.path.exists(path):
        weights, record = train_precond.load_weights(path)
    deflate = not args.no_deflate
    rf = args.mixed_refresh

    # the headline: FP64 block-Jacobi, the reference's arithmetic
    head = run_solve(cfg, pin, "fp64", "block", False, None, args.steps, args.warmup, args.maxiter,
                     profile=True, reference=True, track_self=not args.no_map_track)
    ref_map, cnt = head["ref_map"], head["cnt"]
    runs = {}
    if not args.no_learned:
        plan = [("block_jacobi_same_precision", "block", 0.0)]
        if weights is not None:
            plan.insert(0, ("learned", "learned", 0.0))
            if not args.no_literal:
                plan.append(("learned_literal_fallback", "learned", 1.0))
        else:
            log(f"no trained network at {path}: no learned pipeline")
        for key, pc, growth in plan:
            time.sleep(args.settle)
            runs[key] = run_solve(cfg, pin, "mixed", pc, deflate, weights if pc == "learned" else None, 2, 1,
                                  args.maxiter, profile=False, net_nside=(record or {}).get("net_nside"),
                                  refresh=rf, fallback_growth=growth, keep_map=True,
                                  track=None if args.no_map_track else (ref_map, cnt))
    pin.close()

    kt = head["kernel_times"]
    ms_per_step = head["ms_per_step"]
    per = [(kt[k][0], kt[k][1], k) for k in range(4) if kt[k][1] > 0]
    tot_ms, cnt_launch, dom = max(per)
    # per operator application: with per-detector lengths one application
    # launches each kernel class once per detector group
    n_ops = max(kt[4][1], 1)
    avg_ms = tot_ms / n_ops
    nrec, runs_on = head["run_records"]
    rps = nrec / ntot if runs_on else 0.0
    names = CLASS_NAMES_RUNS if runs_on else CLASS_NAMES
    cbs = class_bytes_per_sample("fp64", rps, head["fp64_angle_bytes"],
                                 head["fp64_sorted_angle_bytes"] or 16)
    alg = ntot * cbs[dom]  # algorithmic bytes of the dominant kernel class per operator application
    achieved = alg / (avg_ms * 1e-3)
    for k in range(4):
        if kt[k][1]:
            log(f"  {names[k]}: {kt[k][0] / n_ops:.3f} ms per operator ({kt[k][1]} launches, "
                f"{ntot * cbs[k] / (kt[k][0] / n_ops * 1e-3) / 1e12:.2f} TB/s of algorithmic bytes)")
    apply_ms = kt[4][0] / max(kt[4][1], 1)
    lens = cfg.seglen if uneven else [cfg.ns] * cfg.ndet
    npix = head["npix"]

    def b_alg(s):
        # whole-operator algorithmic bytes: the per-(detector, sample) stream
        # read once (int32 pixel + cos 2psi + sin 2psi), the per-detector
        # noise weights, ~12 map-vector passes of the CG in FP64
        return ntot * (4 + 2 * s) + sum(L // 2 + 1 for L in lens) * s + 12 * 3 * npix * 8

    traffic, traffic_src = read_traffic(args.workload, "fp64", names[dom])
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
        "dtype": "f64",
        "data": "synthetic (spin/precession scan, Gaussian I/Q/U sky, white + 1/f noise)",
        "config": {"workload": (f"{cfg.ndet} det, {ntot} samples, NSIDE {cfg.nside}, per-detector "
                                f"segment lengths m 2^l (m in 1,3,5), FP64 block-Jacobi PCG" if uneven else
                                f"{cfg.ndet} det x {cfg.nseg * cfg.ns} samples, NSIDE {cfg.nside}, "
                                f"segments of {cfg.ns}, FP64 block-Jacobi PCG"),
                   "nside": cfg.nside, "ndet": cfg.ndet, "samples": ntot,
                   "segment": sorted(set(lens)) if uneven else cfg.ns, "npix": npix, "precond": "block",
                   "deflate": False, "precision": "fp64", "residual_refresh_every": 10},
        "time_to_1e-6_s": tts,
        "iterations_to_1e-6": head["iterations"] if head["converged"] else None,
        "final_true_residual": head["final_true_residual"],
        "setup_s": head["upload_s"] + head["pixelize_s"] + head["setup_s"] + head["begin_s"],
        "phases_s": {k: head[k] for k in ("upload_s", "pixelize_s", "setup_s", "begin_s")},
        "idle_before_each_pipeline_s": args.settle,
        "operator_ms": apply_ms,
        "operator_alg_bytes": b_alg(8),
        "operator_frac_hbm": b_alg(8) / (apply_ms * 1e-3) / HBM_PEAK,
        "device_bytes": head["device_bytes"],
        "run_records": {"records": nrec, "per_sample": nrec / ntot, "used": runs_on},
        "fp64_angle_bytes": head["fp64_angle_bytes"],
        "fp64_sorted_angle_bytes": head["fp64_sorted_angle_bytes"],
        "roofline": {"bound": "hbm", "kernel": names[dom], "achieved": achieved / 1e9,
                     "peak": HBM_PEAK / 1e9, "unit": "GB/s", "frac": achieved / HBM_PEAK,
                     "traffic": traffic, "traffic_source": traffic_src, "bytes_per_launch": alg,
                     "ms_per_launch": avg_ms, "launches_per_operator": cnt_launch // n_ops},
        "reference_map": {"solver": "the FP64 solve continued to |r|^2/|b|^2 <= 1e-24",
                          "iterations": head["ref_iterations"], "true_residual": head["ref_true_residual"]},
        "map_error_at_1e-6": map_errors(head["map"], ref_map, cnt),
    }
    head_ms_it = head["solve_s"] * 1e3 / max(head["iterations"], 1)
    head_setup = head["upload_s"] + head["pixelize_s"] + head["setup_s"] + head["begin_s"]
    t64_maps = None
    if head["map_error_iterations"] is not None or head["map_error_last"] is not None:
        k = head["map_error_iterations"]
        t64_maps = (head_setup + k * head_ms_it * 1e-3) if k else None
        out["maps_within_1e-6_of_fp64_reference"] = {
            "iterations": k, "last_error": head["map_error_last"], "time_s_estimate": t64_maps}
    for key, r in runs.items():
        sub = {"solver": ("mixed operator, " + ("flexible PCG with the trained network" if r["precond"] == "learned"
                                                else "block-Jacobi PCG")
                          + (" + monopole deflation" if deflate else "")
                          + (", fallback whenever |r| does not decrease" if key == "learned_literal_fallback"
                             else "")),
               "time_to_1e-6_s": r["time_to_solution_s"],
               "iterations_to_1e-6": r["iterations"] if r["converged"] else None,
               "final_true_residual": r["final_true_residual"],
               "ms_per_iteration_in_solve": r["solve_s"] * 1e3 / max(r["iterations"], 1),
               "setup_s": r["upload_s"] + r["pixelize_s"] + r["setup_s"] + r["begin_s"],
               "speedup_vs_fp64_time_to_solution": (tts / r["time_to_solution_s"]
                                                    if tts and r["time_to_solution_s"] else None),
               "map_error_vs_fp64_reference": map_errors(r["map"], ref_map, cnt),
               "fp64_refresh_every": rf}
        if r["map_error_iterations"] is not None or r["map_error_last"] is not None:
            k = r["map_error_iterations"]
            tm = (sub["setup_s"] + k * sub["ms_per_iteration_in_solve"] * 1e-3) if k else None
            sub["maps_within_1e-6_of_fp64_reference"] = {
                "iterations": k, "last_error": r["map_error_last"],
                # setup + k iterations at the solve's own time per iteration
                "time_s_estimate": tm,
                "speedup_vs_fp64": (t64_maps / tm) if (t64_maps and tm) else None}
        if "learned_stats" in r:
            sub["stats"] = r["learned_stats"]
        if key == "learned":
            sub["training"] = {k: record.get(k) for k in (
                "net_nside", "pairs", "steps", "rounds", "round_pairs", "round_steps", "loss", "noise_gain",
                "harvest_seeds", "on_policy_seeds", "holdout_block_jacobi", "holdout_trained",
                "validation_block_jacobi_iterations", "validation_iterations", "kept_phase", "total_seconds")}
            sub["training"]["source"] = ("the weights file's own record, written by the offline run of "
                                         "sanepic_amd.train_precond that made it; no training runs here")
        out[key] = sub
    if not args.no_cpu:
        try:
            time.sleep(args.settle)
            out["cpu_baseline"] = cpu_baseline(wl, settle=args.settle)
        except Exception as e:  # the GPU result stands on its own
            out["cpu_baseline"] = {"error": str(e)}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
