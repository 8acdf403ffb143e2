"""Synthetic SANEPIC inputs: scan pointing, sky, TODs, iSpf spectra, segments.

Writes exactly the file tree and binary formats that SANEPIC reads
(README.md:19-33 and src/sanepic.cc of patanch/SANEPIC):

  segments/segment_list_<n>_<ns>.bi   int64 (start, end) pairs, end inclusive
                                      (helpers/generate_segfile.c:22-27)
  pointings_list/pointings_{ra,dec,psi}.bi   float64 radians, shared by all
                                      detectors (sanepic.cc:680-692)
  tod_list/data_<det><ext>.bi         float64 TOD per detector (777-778)
  bolo_info/iSpf..._<det>             float64 inverse noise spectrum (524-527)
  bolo_info/bolometer_info.txt        "num pnum lel xel angle tel name"
                                      (parsed at sanepic.cc:3173-3179)

The scan is a spin/precession strategy; the precession axis is the
anti-solar direction in the ecliptic plane, moving at 2 pi per year.
Angles follow SANEPIC's conventions: dec in [-pi/2, pi/2], ra in (0, 2 pi],
psi in (-pi, pi].  Noise is white plus 1/f, generated per segment as a
circulant in Fourier space (the reference's per-segment noise model).
"""
import math
import mmap
import os
import struct
from concurrent.futures import ProcessPoolExecutor
import multiprocessing as mp

import numpy as np

YEAR_S = 365.25 * 86400.0


def scan_pointing(nsamp, fsamp=19.0, prec_deg=45.0, spin_deg=50.0, spin_min=20.0,
                  prec_min=192.0, start=0, dtype=np.float64):
    """Boresight (ra, dec, psi) for samples [start, start + nsamp)."""
    t = (np.arange(start, start + nsamp, dtype=np.float64)) / fsamp
    ws = 2.0 * np.pi / (spin_min * 60.0)
    wp = 2.0 * np.pi / (prec_min * 60.0)
    beta = np.deg2rad(spin_deg)
    alpha = np.deg2rad(prec_deg)
    # boresight around the spin axis (z)
    ps = ws * t
    bx = math.sin(beta) * np.cos(ps)
    by = math.sin(beta) * np.sin(ps)
    bz = np.full_like(t, math.cos(beta))
    # tilt spin axis by alpha about y
    ca, sa = math.cos(alpha), math.sin(alpha)
    x1 = ca * bx + sa * bz
    y1 = by
    z1 = -sa * bx + ca * bz
    # precess about z
    pp = wp * t
    cp, sp = np.cos(pp), np.sin(pp)
    x2 = cp * x1 - sp * y1
    y2 = sp * x1 + cp * y1
    z2 = z1
    # precession axis (z) -> anti-solar axis in the ecliptic plane: R_y(pi/2)
    x3 = z2
    y3 = y2
    z3 = -x2
    lam = 2.0 * np.pi * t / YEAR_S
    cl, sl = np.cos(lam), np.sin(lam)
    x4 = cl * x3 - sl * y3
    y4 = sl * x3 + cl * y3
    z4 = z3
    theta = np.arccos(np.clip(z4, -1.0, 1.0))
    phi = np.mod(np.arctan2(y4, x4), 2.0 * np.pi)
    ra = np.where(phi <= 0.0, 2.0 * np.pi, phi)
    dec = 0.5 * np.pi - theta
    psi = np.angle(np.exp(1j * ps))  # spin phase wrapped to (-pi, pi]
    psi = np.where(psi <= -np.pi, np.pi, psi)
    return ra.astype(dtype), dec.astype(dtype), psi.astype(dtype)


def ang2pix_ring(nside, theta, phi):
    """Vectorised HEALPix RING ang2pix (Gorski et al. 2005) for signal
    synthesis; the product's bit-exact chain lives in the HIP library."""
    theta = np.asarray(theta, dtype=np.float64)
    phi = np.mod(np.asarray(phi, dtype=np.float64), 2.0 * np.pi)
    z = np.cos(theta)
    za = np.abs(z)
    tt = phi / (0.5 * np.pi)
    ncap = 2 * nside * (nside - 1)
    npix = 12 * nside * nside
    out = np.empty(theta.shape, dtype=np.int64)
    eq = za <= 2.0 / 3.0
    # equatorial belt
    zq, tq = z[eq], tt[eq]
    jp = np.floor(nside * (0.5 + tq - zq * 0.75)).astype(np.int64)
    jm = np.floor(nside * (0.5 + tq + zq * 0.75)).astype(np.int64)
    ir = nside + 1 + jp - jm
    kshift = 1 - (ir & 1)
    ip = (jp + jm - nside + kshift + 1) // 2 + 1
    ip = np.where(ip > 4 * nside, ip - 4 * nside, ip)
    out[eq] = ncap + 4 * nside * (ir - 1) + ip - 1
    # polar caps
    pc = ~eq
    zp, tp_, zap = z[pc], tt[pc], za[pc]
    tp = tp_ - np.floor(tp_)
    tmp = np.sqrt(3.0 * (1.0 - zap))
    jp = np.floor(nside * tp * tmp).astype(np.int64)
    jm = np.floor(nside * (1.0 - tp) * tmp).astype(np.int64)
    ir = jp + jm + 1
    ip = np.floor(tp_ * ir).astype(np.int64) + 1
    ip = np.where(ip > 4 * ir, ip - 4 * ir, ip)
    north = 2 * ir * (ir - 1) + ip - 1
    south = npix - 2 * ir * (ir + 1) + ip - 1
    out[pc] = np.where(zp > 0.0, north, south)
    return out


def orbital_dipole(ra, dec, start=0, fsamp=19.0, amp_k=2.725 * 1.0e-4):
    """Orbital dipole seen along (ra, dec): T_CMB v/c (v = 30 km/s) times the
    cosine between the line of sight and the orbital velocity, which lies in
    the ecliptic plane a quarter turn ahead of the anti-solar direction the
    scan's precession axis follows."""
    t = np.arange(start, start + ra.size, dtype=np.float64) / fsamp
    lam = 2.0 * np.pi * t / YEAR_S + 0.5 * np.pi
    cd = np.cos(dec)
    return amp_k * (np.cos(lam) * cd * np.cos(ra) + np.sin(lam) * cd * np.sin(ra))


def make_sky(nside, seed=1, sigma_uK=(100.0, 10.0, 10.0)):
    rng = np.random.default_rng(seed)
    n = 12 * nside * nside
    return np.stack([rng.standard_normal(n) * s * 1e-6 for s in sigma_uK])


def ispf_spectrum(ns, fsamp, sigma, fknee, alpha):
    """iSpf[k] = 1/(sigma^2 (1 + (fknee/f_k)^alpha)), f_k = k fsamp/ns for
    k <= ns/2, mirrored above; entry 0 (ignored by the reference, 537) gets
    the value of entry 1."""
    k = np.arange(ns, dtype=np.float64)
    kk = np.minimum(k, ns - k)
    kk[0] = 1.0
    f = kk * fsamp / ns
    return 1.0 / (sigma ** 2 * (1.0 + (fknee / f) ** alpha))


def noise_segment(ns, fsamp, sigma, fknee, alpha, seed):
    """White + 1/f noise for one segment as a circulant in Fourier space."""
    rng = np.random.default_rng(seed)
    w = rng.standard_normal(ns) * sigma
    k = np.arange(ns // 2 + 1, dtype=np.float64)
    k[0] = 1.0
    f = k * fsamp / ns
    amp = np.sqrt(1.0 + (fknee / f) ** alpha)
    return np.fft.irfft(np.fft.rfft(w) * amp, ns)


# Named workloads: nside, detectors, samples per detector, segment length.
WORKLOADS = {
    # 16 detectors at k pi/8, NSIDE 512, 2^26 samples each in segments of 2^20
    "nside512_16det": dict(nside=512, ndet=16, nsamp=2 ** 26, ns=2 ** 20),
    # 1 detector, NSIDE 64, 2^22 samples in 8 segments of 2^19
    "nside64_1det": dict(nside=64, ndet=1, nsamp=2 ** 22, ns=2 ** 19),
    "small": dict(nside=64, ndet=4, nsamp=2 ** 21, ns=2 ** 18),
}


# Batched uneven workload: 64 detectors at NSIDE 1024, each with its own
# segment length m 2^l (m in 1, 3, 5; l in 16..21), knee frequency
# (0.02-0.5 Hz) and spectral index (0.8-2.0), on a timeline of 2^24 samples
# (1.07e9 samples in all: 2^26 per detector would need ~390 GB of resident
# per-sample state, more than one GPU's 288 GB).
UNEVEN_WORKLOADS = {
    "nside1024_64det_uneven": dict(nside=1024, ndet=64, nsamp=2 ** 24, seed=7),
}


def workload_config(name, **kw):
    """Config of a named workload (same scan, sky and noise recipe)."""
    if name in UNEVEN_WORKLOADS:
        w = UNEVEN_WORKLOADS[name]
        return draw_uneven(w["nside"], w["ndet"], w["nsamp"], seed=w["seed"], **kw)
    w = WORKLOADS[name]
    return Config(nside=w["nside"], ndet=w["ndet"], nseg=w["nsamp"] // w["ns"], ns=w["ns"], **kw)


class Config:
    """One synthetic run.  Defaults reproduce the 1-detector NSIDE-64 case."""

    def __init__(self, nside=64, ndet=1, nseg=8, ns=2 ** 19, fsamp=19.0, sigma=50e-6,
                 fknee=0.1, alpha=1.0, det_angles_deg=None, offsets_deg=None,
                 sky_seed=1, noise_seed0=1000, start=0, spin_min=20.0, prec_min=192.0,
                 dipole=False):
        self.nside, self.ndet, self.nseg, self.ns = nside, ndet, nseg, ns
        self.dipole = dipole  # TODs carry the orbital dipole, given as ds["odip"]
        self.spin_min, self.prec_min = spin_min, prec_min
        self.fsamp, self.sigma, self.fknee, self.alpha = fsamp, sigma, fknee, alpha
        if det_angles_deg is None:
            det_angles_deg = [22.5 * k for k in range(ndet)]
        self.det_angles_deg = list(det_angles_deg)
        self.offsets_deg = offsets_deg or [(0.0, 0.0)] * ndet
        self.sky_seed, self.noise_seed0, self.start = sky_seed, noise_seed0, start
        self.det_names = [f"det{k}" for k in range(ndet)]

    @property
    def nsamp(self):
        return self.nseg * self.ns


_SHARED = {}


def _alloc_shared(nbytes):
    m = mmap.mmap(-1, max(nbytes, 1))
    return m


def _tod_worker(args):
    idet, seg_lo, seg_hi = args
    cfg = _SHARED["cfg"]
    tod = np.frombuffer(_SHARED["tod_buf"], dtype=np.float64).reshape(cfg.ndet, cfg.nsamp)
    pix = _SHARED["pix"]
    psi = _SHARED["psi"]
    sky = _SHARED["sky"]
    ang = math.radians(cfg.det_angles_deg[idet])
    ns = cfg.ns
    for iseg in range(seg_lo, seg_hi):
        sl = slice(iseg * ns, (iseg + 1) * ns)
        p = pix[sl]
        a2 = 2.0 * (psi[sl] + ang)
        sig = sky[0][p] + sky[1][p] * np.cos(a2) + sky[2][p] * np.sin(a2)
        n = noise_segment(ns, cfg.fsamp, cfg.sigma, cfg.fknee, cfg.alpha,
                          (cfg.noise_seed0 + idet, iseg))
        tod[idet, sl] = sig + n
        if _SHARED["odip"] is not None:
            tod[idet, sl] += _SHARED["odip"][sl]
    return idet


def make_dataset(cfg, workers=None, log=None):
    """Returns dict(ra, dec, psi [nsamp], tod [ndet, nsamp], ispf [ndet, ns],
    sky [3, 12 nside^2], odip [ndet, nsamp] or None).  TOD generation runs in
    forked workers writing into one shared buffer, deterministic for any
    worker count.  With cfg.dipole every TOD carries the orbital dipole of
    the boresight direction, which odip holds for the -o input files."""
    ra, dec, psi = scan_pointing(cfg.nsamp, cfg.fsamp, spin_min=cfg.spin_min,
                                 prec_min=cfg.prec_min, start=cfg.start)
    sky = make_sky(cfg.nside, cfg.sky_seed)
    pix = ang2pix_ring(cfg.nside, 0.5 * np.pi - dec, ra)
    od = orbital_dipole(ra, dec, cfg.start, cfg.fsamp) if getattr(cfg, "dipole", False) else None
    buf = _alloc_shared(cfg.ndet * cfg.nsamp * 8)
    _SHARED.update(cfg=cfg, tod_buf=buf, pix=pix, psi=psi, sky=sky, odip=od)
    chunk = max(1, min(cfg.nseg, 8))
    jobs = [(d, s, min(s + chunk, cfg.nseg)) for d in range(cfg.ndet)
            for s in range(0, cfg.nseg, chunk)]
    if workers is None:
        workers = min(16, os.cpu_count() or 1)
    if workers > 1 and len(jobs) > 1:
        ctx = mp.get_context("fork")
        with ProcessPoolExecutor(max_workers=workers, mp_context=ctx) as ex:
            for _ in ex.map(_tod_worker, jobs):
                pass
    else:
        for j in jobs:
            _tod_worker(j)
    tod = np.frombuffer(buf, dtype=np.float64).reshape(cfg.ndet, cfg.nsamp)
    ispf = np.stack([ispf_spectrum(cfg.ns, cfg.fsamp, cfg.sigma, cfg.fknee, cfg.alpha)
                     for _ in range(cfg.ndet)])
    _SHARED.clear()
    odip = None if od is None else np.broadcast_to(od, (cfg.ndet, cfg.nsamp))
    return dict(ra=ra, dec=dec, psi=psi, tod=tod, ispf=ispf, sky=sky, pix=pix, odip=odip)


class UnevenConfig:
    """Detectors with their own segment length, knee frequency and spectral
    index on one boresight timeline of nsamp samples; detector d uses
    floor(nsamp / seglen[d]) whole segments."""

    def __init__(self, nside, seglen, nsamp, fknee=None, alpha=None, fsamp=19.0, sigma=50e-6,
                 det_angles_deg=None, sky_seed=1, noise_seed0=1000, start=0, spin_min=20.0,
                 prec_min=192.0):
        self.nside, self.nsamp = nside, nsamp
        self.seglen = [int(v) for v in seglen]
        self.ndet = ndet = len(self.seglen)
        self.fknee = list(fknee) if fknee is not None else [0.1] * ndet
        self.alpha = list(alpha) if alpha is not None else [1.0] * ndet
        self.fsamp, self.sigma = fsamp, sigma
        self.det_angles_deg = list(det_angles_deg) if det_angles_deg is not None else \
            [22.5 * (k % 8) + 2.8125 * (k // 8) for k in range(ndet)]
        self.offsets_deg = [(0.0, 0.0)] * ndet
        self.sky_seed, self.noise_seed0, self.start = sky_seed, noise_seed0, start
        self.spin_min, self.prec_min = spin_min, prec_min

    def nseg(self, d):
        return self.nsamp // self.seglen[d]

    def used(self, d):
        return self.nseg(d) * self.seglen[d]


def draw_uneven(nside, ndet, nsamp, seed=7, log2_range=(16, 21), radices=(1, 3, 5),
                fknee_range=(0.02, 0.5), alpha_range=(0.8, 2.0), **kw):
    """A batched uneven case: per detector a segment length m 2^l (m from
    radices, l uniform in log2_range, at most nsamp), fknee and alpha uniform
    in their ranges, drawn from seed."""
    rng = np.random.default_rng(seed)
    seglen, fk, al = [], [], []
    for _ in range(ndet):
        while True:
            L = int(rng.choice(radices)) * 2 ** int(rng.integers(log2_range[0], log2_range[1] + 1))
            if L <= nsamp:
                break
        seglen.append(L)
        fk.append(float(rng.uniform(*fknee_range)))
        al.append(float(rng.uniform(*alpha_range)))
    return UnevenConfig(nside, seglen, nsamp, fknee=fk, alpha=al, **kw)


def _uneven_worker(args):
    idet, seg_lo, seg_hi = args
    cfg = _SHARED["cfg"]
    off = _SHARED["off"]
    tod = np.frombuffer(_SHARED["tod_buf"], dtype=np.float64)
    pix, psi, sky = _SHARED["pix"], _SHARED["psi"], _SHARED["sky"]
    ang = math.radians(cfg.det_angles_deg[idet])
    L = cfg.seglen[idet]
    for iseg in range(seg_lo, seg_hi):
        sl = slice(iseg * L, (iseg + 1) * L)
        p = pix[sl]
        a2 = 2.0 * (psi[sl] + ang)
        sig = sky[0][p] + sky[1][p] * np.cos(a2) + sky[2][p] * np.sin(a2)
        n = noise_segment(L, cfg.fsamp, cfg.sigma, cfg.fknee[idet], cfg.alpha[idet],
                          (cfg.noise_seed0 + idet, iseg))
        tod[off[idet] + iseg * L:off[idet] + (iseg + 1) * L] = sig + n
    return idet


def make_uneven_dataset(cfg, workers=None):
    """Returns dict(ra, dec, psi [nsamp], tod: list of per-detector arrays
    [used(d)], ispf: list [seglen(d)], sky, pix, off)."""
    ra, dec, psi = scan_pointing(cfg.nsamp, cfg.fsamp, spin_min=cfg.spin_min,
                                 prec_min=cfg.prec_min, start=cfg.start)
    sky = make_sky(cfg.nside, cfg.sky_seed)
    pix = ang2pix_ring(cfg.nside, 0.5 * np.pi - dec, ra)
    off = np.zeros(cfg.ndet + 1, dtype=np.int64)
    for d in range(cfg.ndet):
        off[d + 1] = off[d] + cfg.used(d)
    buf = _alloc_shared(int(off[-1]) * 8)
    _SHARED.update(cfg=cfg, tod_buf=buf, pix=pix, psi=psi, sky=sky, off=off)
    jobs = []
    for d in range(cfg.ndet):
        ns_d = cfg.nseg(d)
        chunk = max(1, min(ns_d, 8))
        jobs += [(d, s, min(s + chunk, ns_d)) for s in range(0, ns_d, chunk)]
    if workers is None:
        workers = min(16, os.cpu_count() or 1)
    if workers > 1 and len(jobs) > 1:
        ctx = mp.get_context("fork")
        with ProcessPoolExecutor(max_workers=workers, mp_context=ctx) as ex:
            for _ in ex.map(_uneven_worker, jobs):
                pass
    else:
        for j in jobs:
            _uneven_worker(j)
    flat = np.frombuffer(buf, dtype=np.float64)
    tod = [flat[off[d]:off[d + 1]] for d in range(cfg.ndet)]
    ispf = [ispf_spectrum(cfg.seglen[d], cfg.fsamp, cfg.sigma, cfg.fknee[d], cfg.alpha[d])
            for d in range(cfg.ndet)]
    _SHARED.clear()
    return dict(ra=ra, dec=dec, psi=psi, tod=tod, ispf=ispf, sky=sky, pix=pix, off=off)


def segment_list(nseg, ns, start=0):
    """int64 (start, end) pairs, end inclusive (generate_segfile.c:22-27)."""
    out = np.empty(2 * nseg, dtype=np.int64)
    out[0::2] = start + np.arange(nseg, dtype=np.int64) * ns
    out[1::2] = start + (np.arange(nseg, dtype=np.int64) + 1) * ns - 1
    return out


def ispf_prefix(cfg):
    return f"iSpf{cfg.fknee}_{cfg.fsamp / 2}_index{cfg.alpha}__"


def write_tree(base, cfg, ds, suffix="output"):
    """Writes the README's directory tree; returns the argument dict that
    run_sanepic.sh would build (paths and scalar parameters)."""
    for sub in ("bolo_info", "out_dir", "pointings_list", "segments", "tod_list"):
        os.makedirs(os.path.join(base, sub), exist_ok=True)
    with open(os.path.join(base, "bolo_info", "bolometer_info.txt"), "w") as fh:
        for k, name in enumerate(cfg.det_names):
            lel, xel = cfg.offsets_deg[k]
            fh.write(f"{k} 0 {lel!r} {xel!r} {cfg.det_angles_deg[k]!r} LFT {name}\n")
    pre = ispf_prefix(cfg)
    for k, name in enumerate(cfg.det_names):
        ds["ispf"][k].astype("<f8").tofile(os.path.join(base, "bolo_info", pre + name))
        np.ascontiguousarray(ds["tod"][k]).astype("<f8").tofile(
            os.path.join(base, "tod_list", f"data_{name}{suffix}.bi"))
        if ds.get("odip") is not None:  # read with -o from <-F>/orbdip_<bolo> (702-708)
            np.ascontiguousarray(ds["odip"][k]).astype("<f8").tofile(
                os.path.join(base, "tod_list", f"orbdip_{name}"))
    for key in ("ra", "dec", "psi"):
        ds[key].astype("<f8").tofile(os.path.join(base, "pointings_list", f"pointings_{key}.bi"))
    seg = os.path.join(base, "segments", f"segment_list_{cfg.nseg}_{cfg.ns}.bi")
    segment_list(cfg.nseg, cfg.ns, cfg.start).astype("<i8").tofile(seg)
    return dict(
        tod_dir=os.path.join(base, "tod_list") + "/",
        suffix=suffix,
        segfile=seg,
        ndet=cfg.ndet,
        dets=cfg.det_names,
        ispf_prefix=os.path.join(base, "bolo_info", pre),
        ns=cfg.ns,
        nseg=cfg.nseg,
        nside=cfg.nside,
        bolofile=os.path.join(base, "bolo_info", "bolometer_info.txt"),
        pointing=os.path.join(base, "pointings_list", "pointings_"),
        out=os.path.join(base, "out_dir", "map_"),
        ext=f"{suffix}_N{cfg.nside}_.bin",
    )


def sanepic_argv(args, exe="sanepic", hwp=0, omega=1.123456789123, pol=1, write_iter=0, odip=0,
                 leakage_file=None):
    """The flag list of run_sanepic.sh:55-73 for the given tree."""
    argv = [exe, "-F", args["tod_dir"], "-B", args["suffix"], "-f", args["segfile"],
            "-d", str(args["ndet"])]
    for d in args["dets"]:
        argv += ["-C", d]
    argv += ["-e", args["ext"], "-h", str(hwp), "-w", repr(omega), "-k", args["ispf_prefix"],
             "-u", str(args["ns"]), "-O", args["out"], "-p", str(pol), "-l", str(args["ns"]),
             "-n", str(args["nseg"]), "-N", str(args["nside"]), "-X", args["bolofile"],
             "-Z", args["pointing"], "-i", str(write_iter)]
    if odip:
        argv += ["-o", str(odip)]
    if leakage_file:  # -b 1 with the coefficients file (551-575)
        argv += ["-b", "1", "-K", leakage_file]
    return argv
