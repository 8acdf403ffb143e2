"""Shared builders for parity tests: one synthetic run fed to both the GPU
library (through the C ABI) and the CPU oracle."""
import numpy as np

import oracle
from sanepic_amd import synth


def make_case(nside=8, ndet=2, nseg=3, ns=1024, fknee=0.2, offsets=None, seed=1, start=0,
              fsamp=19.0, spins=40.0, precessions=5.0, dipole=False):
    # short test streams: the scan is sped up so the stream spans `spins` spin
    # periods and `precessions` precession periods, which gives every pixel
    # revisits at several polarisation angles (well-conditioned 3x3 blocks)
    tmin = nseg * ns / fsamp / 60.0
    cfg = synth.Config(nside=nside, ndet=ndet, nseg=nseg, ns=ns, fknee=fknee, fsamp=fsamp,
                       spin_min=tmin / spins, prec_min=tmin / precessions,
                       offsets_deg=offsets, sky_seed=seed, noise_seed0=1000 + 17 * seed,
                       start=start, dipole=dipole)
    ds = synth.make_dataset(cfg, workers=1)
    return cfg, ds


def oracle_problem(cfg, ds, nthreads=4, hwp=0, omega=0.0):
    P = oracle.Problem(cfg.nside, cfg.ndet, cfg.nseg, cfg.ns, hwp=hwp, omega=omega,
                       nthreads=nthreads)
    for k in range(cfg.ndet):
        lel, xel = cfg.offsets_deg[k]
        P.set_detector(k, lel, xel, cfg.det_angles_deg[k], ds["ispf"][k])
    P.pixelize(ds["ra"], ds["dec"], ds["psi"])
    P.load_tod(ds["tod"])
    if ds.get("odip") is not None:
        P.set_dipole(ds["odip"])
    return P


def rel(a, b):
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)
