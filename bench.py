"""Benchmark: CG GLS map-making on one MI355X.

Workload (default): 16 detectors at angles k*pi/8, NSIDE 512, 2^26 samples per
detector in segments of 2^20 (1,073,741,824 TOD samples), spin/precession
scan, white + 1/f noise (fknee 0.1 Hz, index 1), FP64 operator, 3x3
block-Jacobi PCG with the monopole deflated, residual refresh every 10
iterations.  Synthetic inputs are generated in host RAM first (untimed).

Reported:
  value            TOD samples per second per CG iteration (whole job = 1 GPU)
  time_to_1e-6_s   wall time from inputs in host RAM to a true relative
                   residual |b - A x| / |b| <= 1e-6 (H2D, pointing, ang2pix,
                   sort, b, preconditioner, CG), file I/O excluded
  roofline         the dominant kernel's algorithmic bytes per launch over its
                   average launch duration (HIP events on the library stream)
  cpu_baseline     the CPU oracle (an FP64 restatement of the reference's
                   path) on a bounded sample of the same workload
One step = one CG iteration.  W warm-up iterations, then exactly K timed.
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
WORKLOADS = {
    # name: nside, ndet, samples per detector, segment length
    "c1": dict(nside=512, ndet=16, nsamp=2 ** 26, ns=2 ** 20),
    "c0": dict(nside=64, ndet=1, nsamp=2 ** 22, ns=2 ** 19),
    "small": dict(nside=64, ndet=4, nsamp=2 ** 21, ns=2 ** 18),
}
CLASS_NAMES = ["gather+column FFT", "row FFT + noise filter", "inverse column FFT",
               "pixel-sorted P^T reduction"]


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def stream_bytes_per_sample(precision):
    # compulsory per-sample stream of the operator: compressed pixel (int32)
    # plus cos 2psi and sin 2psi in the operator's precision
    s = 8 if precision == "fp64" else 4
    return 4 + 2 * s


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
    nseg = 4 if wl["ns"] >= 2 ** 20 else 8
    cfg = synth.Config(nside=wl["nside"], ndet=1, nseg=nseg, ns=wl["ns"])
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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="c1", choices=sorted(WORKLOADS))
    ap.add_argument("--precision", default="fp64", choices=["fp64", "mixed"])
    ap.add_argument("--precond", default="block", choices=["block", "diag"])
    ap.add_argument("--no-deflate", action="store_true")
    ap.add_argument("--maxiter", type=int, default=600)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--calib-copy", action="store_true",
                    help="run a 1 GiB device copy first (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path "
              "(the sharding design is described in DESIGN.md)", file=sys.stderr)
        sys.exit(2)

    import sanepic_amd
    from sanepic_amd import synth

    if args.calib_copy:
        calibration_copy()
    wl = WORKLOADS[args.workload]
    nseg = wl["nsamp"] // wl["ns"]
    cfg = synth.Config(nside=wl["nside"], ndet=wl["ndet"], nseg=nseg, ns=wl["ns"])
    tg = time.perf_counter()
    ds = synth.make_dataset(cfg, workers=min(16, os.cpu_count() or 1))
    log(f"inputs generated in {time.perf_counter() - tg:.1f} s "
        f"({cfg.ndet} det x {cfg.nsamp} samples, NSIDE {cfg.nside})")
    ntot = cfg.ndet * cfg.nsamp
    prec = sanepic_amd.FP64 if args.precision == "fp64" else sanepic_amd.MIXED
    pc = sanepic_amd.PRECOND_BLOCK3 if args.precond == "block" else sanepic_amd.PRECOND_DIAG
    deflate = not args.no_deflate

    # ---- time to solution: inputs in host RAM -> true residual <= 1e-6
    t0 = time.perf_counter()
    ctx = sanepic_amd.Context(cfg.nside, cfg.ndet, cfg.nseg, cfg.ns, precision=prec)
    ctx.load(cfg, ds)
    t_load = time.perf_counter()
    npix, ndef = ctx.pixelize()
    t_pix = time.perf_counter()
    ctx.setup()
    t_setup = time.perf_counter()
    ctx.solve_begin(pc, 10, None, deflate)
    t_begin = time.perf_counter()
    log(f"load {t_load - t0:.2f} s, pixelize {t_pix - t_load:.2f} s (npix {npix}, "
        f"{ndef} deferred), setup {t_setup - t_pix:.2f} s, solve_begin {t_begin - t_setup:.2f} s, "
        f"device memory {ctx.device_bytes() / 1e9:.1f} GB")

    hist = []
    for _ in range(args.warmup):
        hist.append(ctx.step())
    ctx.set_profiling(True)
    tk0 = time.perf_counter()
    for _ in range(args.steps):
        hist.append(ctx.step())
    tk1 = time.perf_counter()
    kt = [ctx.kernel_time(k) for k in range(5)]
    ctx.set_profiling(False)
    ms_per_step = (tk1 - tk0) * 1e3 / args.steps

    converged, true_res = False, None
    while len(hist) < args.maxiter:
        if hist[-1]["crit"] <= 1e-12:
            true_res = ctx.true_residual()
            if true_res <= 1e-6:
                converged = True
                break
        hist.append(ctx.step())
    t_end = time.perf_counter()
    if true_res is None or not converged:
        true_res = ctx.true_residual()
        converged = true_res <= 1e-6
    tts = t_end - t0
    log(f"{len(hist)} iterations, final crit {hist[-1]['crit']:.3e}, true residual {true_res:.3e}, "
        f"time to solution {tts:.2f} s ({'reached' if converged else 'NOT reached'} 1e-6)")

    # ---- roofline of the dominant kernel (live HIP events)
    per = [(kt[k][0], kt[k][1], k) for k in range(4) if kt[k][1] > 0]
    tot_ms, cnt, dom = max(per)
    avg_ms = tot_ms / cnt
    bps = stream_bytes_per_sample(args.precision)
    alg = ntot * bps  # compulsory stream bytes of one launch
    achieved = alg / (avg_ms * 1e-3)
    for k in range(4):
        if kt[k][1]:
            log(f"  {CLASS_NAMES[k]}: {kt[k][0] / kt[k][1]:.3f} ms x {kt[k][1]}")
    apply_ms = kt[4][0] / max(kt[4][1], 1)
    b_alg = ntot * bps + cfg.ndet * (cfg.ns // 2 + 1) * (8 if args.precision == "fp64" else 4) \
        + 12 * 3 * npix * 8
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
        "config": {"workload": f"{cfg.ndet} det x {cfg.nsamp} samples, NSIDE {cfg.nside}, "
                               f"segments of {cfg.ns}, {args.precond} PCG"
                               f"{' + monopole deflation' if deflate else ''}",
                   "nside": cfg.nside, "ndet": cfg.ndet, "samples_per_det": cfg.nsamp,
                   "segment": cfg.ns, "npix": npix, "precond": args.precond,
                   "deflate": deflate, "precision": args.precision},
        "time_to_1e-6_s": tts if converged else None,
        "iterations_to_1e-6": len(hist) if converged else None,
        "final_true_residual": true_res,
        "setup_s": t_begin - t0,
        "operator_ms": apply_ms,
        "operator_frac_hbm": b_alg / (apply_ms * 1e-3) / HBM_PEAK,
        "roofline": {"bound": "hbm", "kernel": CLASS_NAMES[dom], "achieved": achieved / 1e9,
                     "peak": HBM_PEAK / 1e9, "unit": "GB/s", "frac": achieved / HBM_PEAK,
                     "traffic": None, "bytes_per_launch": alg, "avg_launch_ms": avg_ms},
    }
    ctx.close()
    del ds
    if not args.no_cpu:
        try:
            out["cpu_baseline"] = cpu_baseline(wl)
        except Exception as e:  # the GPU result stands on its own
            out["cpu_baseline"] = {"error": str(e)}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
