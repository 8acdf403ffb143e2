"""Train the learned preconditioner for one scan strategy.

    python -m sanepic_amd.train_precond --workload nside512_16det --pairs 2000 --steps 10000

Builds the named synthetic workload (the generator bench.py uses: same
pointing, detectors and noise spectra), then on the GPU:

1. harvests training pairs (r_k, A^-1 r_k) from block-Jacobi PCG solves of
   synthetic right-hand sides b' = A m' + g P^T N^-1 n' drawn with seeds
   10000, 10001, ... (disjoint from the data's sky seed 1 and noise seeds
   1000 + det).  The noise gain g (default 1024) makes them noise-dominated:
   the network then learns the stripe modes that block-Jacobi leaves to CG
   (at 16 det x 2^26, NSIDE 512: 32 iterations against 42; trained on
   right-hand sides with the data's own noise level, g = 1, it needed 76),
2. trains the U-Net with Adam (cosine learning-rate schedule) on the
   residual loss |A z - r|^2 / |r|^2 (default) or the A-norm loss
   |z - e|_A^2 / |e|_A^2, holding out the last pairs,
3. then, for --rounds rounds, harvests pairs from the trained network's own
   flexible-PCG iterates (the residuals the learned solver actually meets,
   which differ from block-Jacobi ones: trained on the latter alone the
   solver can lose to plain block-Jacobi) and fine-tunes on the union,
4. keeps the parameters with the fewest iterations to |r|^2/|b|^2 <= 1e-12
   on a validation right-hand side (seed --val-seed, never trained on),
5. writes them (float32, little endian, the library's layout) and a JSON
   record next to them (sizes, seeds, losses, iterations, timings, SHA-256).

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
    ap.add_argument("--workload", default="nside512_16det", choices=sorted(synth.WORKLOADS) + sorted(synth.UNEVEN_WORKLOADS))
    ap.add_argument("--pairs", type=int, default=1500, help="block-Jacobi pairs")
    ap.add_argument("--holdout", type=int, default=100)
    ap.add_argument("--steps", type=int, default=6000)
    ap.add_argument("--rounds", type=int, default=2, help="on-policy harvest + fine-tune rounds")
    ap.add_argument("--round-pairs", type=int, default=1000)
    ap.add_argument("--round-steps", type=int, default=3000)
    ap.add_argument("--round-lr", type=float, default=3e-4)
    ap.add_argument("--chunk", type=int, default=500, help="steps per progress line")
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--lr-final", type=float, default=1e-5)
    ap.add_argument("--seed0", type=int, default=10000)
    ap.add_argument("--val-seed", type=int, default=20000)
    ap.add_argument("--noise-gain", type=float, default=1024.0,
                    help="noise amplitude of the training right-hand sides relative to the detectors' own "
                         "(> 1: noise-dominated, the stripe modes the solver finds hardest)")
    ap.add_argument("--init-seed", type=int, default=1)
    ap.add_argument("--net-nside", type=int, default=64,
                    help="network resolution (capped at the map's nside): the network sees the residual "
                         "averaged over (nside / net-nside)^2 face pixels")
    ap.add_argument("--precision", default="mixed", choices=["fp64", "mixed"])
    ap.add_argument("--loss", default="residual", choices=["residual", "anorm"])
    ap.add_argument("--out", default=None)
    a = ap.parse_args(argv)
    out = a.out or weights_path(a.workload)
    t0 = time.time()
    cfg = synth.workload_config(a.workload)
    uneven = a.workload in synth.UNEVEN_WORKLOADS
    workers = min(16, os.cpu_count() or 1)
    ds = synth.make_uneven_dataset(cfg, workers) if uneven else synth.make_dataset(cfg, workers)
    nsamp = sum(cfg.used(d) for d in range(cfg.ndet)) if uneven else cfg.ndet * cfg.nseg * cfg.ns
    log(f"workload {a.workload}: {cfg.ndet} det, {nsamp} samples, NSIDE {cfg.nside} ({time.time() - t0:.1f} s)")
    prec = sanepic_amd.FP64 if a.precision == "fp64" else sanepic_amd.MIXED
    if uneven:  # per-detector segment lengths
        ctx = sanepic_amd.Context.uneven(cfg.nside, cfg.seglen, cfg.nsamp, precision=prec)
        ctx.load_uneven(cfg, ds)
    else:
        ctx = sanepic_amd.Context(cfg.nside, cfg.ndet, cfg.nseg, cfg.ns, precision=prec)
        ctx.load(cfg, ds)
    del ds
    npix, _ = ctx.pixelize()
    ctx.setup()
    net_nside = min(a.net_nside, cfg.nside)
    un = sanepic_amd.UNet(net_nside, seed=a.init_seed)
    ctx.learn_attach(un)
    log(f"set-up done, {npix} pixels ({time.time() - t0:.1f} s)")
    capacity = a.pairs + (a.round_pairs if a.rounds > 0 else 0)
    kind = sanepic_amd.LOSS_RESIDUAL if a.loss == "residual" else sanepic_amd.LOSS_ANORM

    def validate(policy):
        # iterations to |r|^2/|b|^2 <= 1e-12 on the validation right-hand side
        # (keep_crit above any criterion: no pair is stored)
        return ctx.harvest(a.val_seed, capacity=capacity, keep_crit=1e300, stop_crit=1e-12,
                           policy=policy)[1]

    def train(steps, lr, first, count, tag):
        losses = []
        done = 0
        tt = time.time()
        while done < steps:
            k = min(a.chunk, steps - done)
            lo = ctx.train(k, lr=lr, lr_final=a.lr_final, seed=a.init_seed * 1000003 + done + 7919 * len(history),
                           first=first, count=count, sched_offset=done, sched_total=steps, loss=kind)
            losses.extend(float(v) for v in lo)
            done += k
            log(f"{tag} step {done}/{steps}: mean training loss {np.mean(lo):.4f} ({time.time() - tt:.0f} s)")
        return losses, time.time() - tt

    history = []
    val_bj = validate(sanepic_amd.HARVEST_BLOCK3)
    log(f"validation right-hand side: block-Jacobi {val_bj} iterations")

    th = time.time()
    seed = a.seed0
    while ctx.npairs() < a.pairs:
        added, iters = ctx.harvest(seed, capacity=capacity, noise_gain=a.noise_gain)
        log(f"harvest seed {seed}: {added} pairs from {iters} iterations, total {ctx.npairs()}")
        seed += 1
    ctx.truncate_pairs(min(ctx.npairs(), a.pairs))
    harvest_s = time.time() - th
    n = ctx.npairs()
    ntrain = n - a.holdout
    hold = range(ntrain, n)

    def held_out():
        return {name: float(np.mean([ctx.eval_pair(k, kk) for k in hold]))
                for name, kk in (("residual", sanepic_amd.LOSS_RESIDUAL), ("anorm", sanepic_amd.LOSS_ANORM))}

    base = held_out()
    log(f"{n} pairs in {harvest_s:.1f} s; held-out losses of block-Jacobi alone {base}")
    losses, train_s = train(a.steps, a.lr, 0, ntrain, "block-Jacobi pairs")
    held = held_out()
    v = validate(sanepic_amd.HARVEST_LEARNED)
    log(f"held-out losses: block-Jacobi {base}, trained {held}; validation {v} iterations (block-Jacobi {val_bj})")
    history.append(dict(phase="block-Jacobi pairs", pairs=n, steps=a.steps, val_iterations=v,
                        train_loss_first_100=float(np.mean(losses[:100])),
                        train_loss_last_100=float(np.mean(losses[-100:])), train_seconds=train_s))
    best = (v, un.params(), 0)
    rseed = a.seed0 + 5000
    for rnd in range(1, a.rounds + 1):
        th = time.time()
        ctx.truncate_pairs(a.pairs)
        while ctx.npairs() < capacity:
            added, iters = ctx.harvest(rseed, capacity=capacity, policy=sanepic_amd.HARVEST_LEARNED,
                                       noise_gain=a.noise_gain)
            log(f"round {rnd} on-policy harvest seed {rseed}: {added} pairs from {iters} iterations, "
                f"total {ctx.npairs()}")
            rseed += 1
        hs = time.time() - th
        losses, ts = train(a.round_steps, a.round_lr, 0, ctx.npairs(), f"round {rnd}")
        v = validate(sanepic_amd.HARVEST_LEARNED)
        log(f"round {rnd}: validation {v} iterations (block-Jacobi {val_bj})")
        history.append(dict(phase=f"on-policy round {rnd}", pairs=ctx.npairs(), steps=a.round_steps,
                            val_iterations=v, harvest_seconds=hs,
                            train_loss_first_100=float(np.mean(losses[:100])),
                            train_loss_last_100=float(np.mean(losses[-100:])), train_seconds=ts))
        if v < best[0]:
            best = (v, un.params(), rnd)

    p = best[1]
    os.makedirs(os.path.dirname(out), exist_ok=True)
    p.astype("<f4").tofile(out)
    rec = dict(workload=a.workload, nside=cfg.nside, net_nside=net_nside, ndet=cfg.ndet,
               samples=nsamp, segment=sorted(set(cfg.seglen)) if uneven else cfg.ns, npix=npix,
               nparams=int(p.size), pairs=n, holdout=a.holdout,
               harvest_seeds=[a.seed0, seed - 1], on_policy_seeds=[a.seed0 + 5000, rseed - 1],
               steps=a.steps, rounds=a.rounds, round_pairs=a.round_pairs, round_steps=a.round_steps,
               lr=a.lr, round_lr=a.round_lr, lr_final=a.lr_final,
               loss=a.loss, init_seed=a.init_seed, precision=a.precision, noise_gain=a.noise_gain,
               holdout_block_jacobi=base, holdout_trained=held,
               validation_seed=a.val_seed, validation_block_jacobi_iterations=val_bj,
               validation_iterations=best[0], kept_phase=best[2], history=history,
               harvest_seconds=harvest_s, total_seconds=time.time() - t0,
               sha256=hashlib.sha256(p.astype("<f4").tobytes()).hexdigest())
    with open(out + ".json", "w") as f:
        json.dump(rec, f, indent=1)
    log(f"wrote {out} ({p.size} parameters, phase {best[2]}, {best[0]} validation iterations) "
        f"in {time.time() - t0:.1f} s")
    ctx.learn_attach(None)
    ctx.close()
    un.close()


if __name__ == "__main__":
    main()
