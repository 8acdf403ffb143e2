"""Train the learned preconditioner for one scan strategy.

    python -m sanepic_amd.train_precond --workload c1 --pairs 2000 --steps 10000

Builds the named synthetic workload (the generator bench.py uses: same
pointing, detectors and noise spectra), then on the GPU:

1. harvests training pairs (r_k, A^-1 r_k) from block-Jacobi PCG solves of
   synthetic right-hand sides b' = A m' + P^T N^-1 n' drawn with seeds
   10000, 10001, ... (disjoint from the data's sky seed 1 and noise seeds
   1000 + det),
2. trains the U-Net with Adam (cosine learning-rate schedule) on the
   residual loss |A z - r|^2 / |r|^2 (default) or the A-norm loss
   |z - e|_A^2 / |e|_A^2, holding out the last pairs,
3. writes the parameters (float32, little endian, the library's layout) and a
   JSON record next to them (sizes, seeds, losses, timings, SHA-256).

The network depends on the scan and the noise model, not on the sky or the
noise realisation, so the weights serve any data taken with the same scan.
"""
import argparse
import hashlib
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

import sanepic_amd  # noqa: E402
from sanepic_amd import synth  # noqa: E402

WEIGHTS_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "weights")


def weights_path(workload):
    return os.path.join(WEIGHTS_DIR, f"unet_{workload}.f32")


def load_weights(path):
    """(params float32, record dict) of a trained network."""
    with open(path + ".json") as f:
        rec = json.load(f)
    p = np.fromfile(path, dtype="<f4")
    if p.size != rec["nparams"]:
        raise ValueError(f"{path}: {p.size} parameters, record says {rec['nparams']}")
    if hashlib.sha256(p.tobytes()).hexdigest() != rec["sha256"]:
        raise ValueError(f"{path}: checksum mismatch")
    return p, rec


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--workload", default="c1", choices=sorted(synth.WORKLOADS))
    ap.add_argument("--pairs", type=int, default=2000)
    ap.add_argument("--holdout", type=int, default=100)
    ap.add_argument("--steps", type=int, default=10000)
    ap.add_argument("--chunk", type=int, default=500, help="steps per progress line")
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--lr-final", type=float, default=1e-5)
    ap.add_argument("--seed0", type=int, default=10000)
    ap.add_argument("--init-seed", type=int, default=1)
    ap.add_argument("--precision", default="mixed", choices=["fp64", "mixed"])
    ap.add_argument("--loss", default="residual", choices=["residual", "anorm"])
    ap.add_argument("--out", default=None)
    a = ap.parse_args(argv)
    out = a.out or weights_path(a.workload)
    t0 = time.time()
    cfg = synth.workload_config(a.workload)
    ds = synth.make_dataset(cfg, workers=min(16, os.cpu_count() or 1))
    log(f"workload {a.workload}: {cfg.ndet} det x {cfg.nseg * cfg.ns} samples, NSIDE {cfg.nside} "
        f"({time.time() - t0:.1f} s)")
    prec = sanepic_amd.FP64 if a.precision == "fp64" else sanepic_amd.MIXED
    ctx = sanepic_amd.Context(cfg.nside, cfg.ndet, cfg.nseg, cfg.ns, precision=prec)
    ctx.load(cfg, ds)
    del ds
    npix, _ = ctx.pixelize()
    ctx.setup()
    un = sanepic_amd.UNet(cfg.nside, seed=a.init_seed)
    ctx.learn_attach(un)
    log(f"set-up done, {npix} pixels ({time.time() - t0:.1f} s)")

    th = time.time()
    seed = a.seed0
    while ctx.npairs() < a.pairs:
        added, iters = ctx.harvest(seed, capacity=a.pairs)
        log(f"harvest seed {seed}: {added} pairs from {iters} iterations, total {ctx.npairs()}")
        seed += 1
    harvest_s = time.time() - th
    n = ctx.npairs()
    ntrain = n - a.holdout
    hold = range(ntrain, n)
    kind = sanepic_amd.LOSS_RESIDUAL if a.loss == "residual" else sanepic_amd.LOSS_ANORM

    def held_out():
        return {name: float(np.mean([ctx.eval_pair(k, kk) for k in hold]))
                for name, kk in (("residual", sanepic_amd.LOSS_RESIDUAL), ("anorm", sanepic_amd.LOSS_ANORM))}

    base = held_out()
    log(f"{n} pairs in {harvest_s:.1f} s; held-out losses of block-Jacobi alone {base}")

    tt = time.time()
    losses = []
    done = 0
    while done < a.steps:
        k = min(a.chunk, a.steps - done)
        lo = ctx.train(k, lr=a.lr, lr_final=a.lr_final, seed=a.init_seed * 1000003 + done,
                       first=0, count=ntrain, sched_offset=done, sched_total=a.steps, loss=kind)
        losses.extend(float(v) for v in lo)
        done += k
        log(f"step {done}/{a.steps}: mean training loss {np.mean(lo):.4f} ({time.time() - tt:.0f} s)")
    train_s = time.time() - tt
    held = held_out()
    log(f"held-out losses: block-Jacobi {base}, trained {held}")

    p = un.params()
    os.makedirs(os.path.dirname(out), exist_ok=True)
    p.astype("<f4").tofile(out)
    rec = dict(workload=a.workload, nside=cfg.nside, ndet=cfg.ndet, samples_per_det=cfg.nseg * cfg.ns,
               segment=cfg.ns, npix=npix, nparams=int(p.size), pairs=n, holdout=a.holdout,
               harvest_seeds=[a.seed0, seed - 1], steps=a.steps, lr=a.lr, lr_final=a.lr_final,
               loss=a.loss, init_seed=a.init_seed, precision=a.precision,
               holdout_block_jacobi=base, holdout_trained=held,
               train_loss_first_100=float(np.mean(losses[:100])),
               train_loss_last_100=float(np.mean(losses[-100:])),
               harvest_seconds=harvest_s, train_seconds=train_s,
               sha256=hashlib.sha256(p.astype("<f4").tobytes()).hexdigest())
    with open(out + ".json", "w") as f:
        json.dump(rec, f, indent=1)
    log(f"wrote {out} ({p.size} parameters) in {time.time() - t0:.1f} s")
    ctx.learn_attach(None)
    ctx.close()
    un.close()


if __name__ == "__main__":
    main()
