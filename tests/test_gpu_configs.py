"""GPU parity at the benchmark workloads' own shapes, through the C ABI.

The small parity cases (test_gpu_parity.py) run single-pass and short
four-step segments; these run the kernels the benchmark workloads use:

* the 1-detector NSIDE-64 workload at full size (8 segments of 2^19):
  pixels, hits, crosslinks, Mp, b, A x, the exact 3x3 blocks, the first
  diagonal-PCG and block-Jacobi (+ monopole deflation) iterates, and a
  150-iteration block-Jacobi solve whose map solves the oracle's system as
  well as the oracle's own does;
* NSIDE 512 with segments of 2^20 (the 16-detector workload's four-step
  split, column length 256), FP64 and mixed precision;
* one detector's full 2^26-sample pointing at NSIDE 512, pixel by pixel;
* per-detector segment lengths m 2^l (m = 1, 3, 5; l = 16..21), the
  batched uneven workload's lengths, FP64 and mixed precision;
* pixels hit more than 4096 times, whose P^T work is split over several
  items and recombined (uniform and uneven contexts);
* detector offsets of +-0.5 deg and a half-wave plate;
* the exact blocks and the block-Jacobi / A-DEF2 iterates on the small cases.

The checker is the CPU oracle (oracle/), run with all host cores."""
import os

import numpy as np
import pytest

import oracle
import sanepic_amd
from sanepic_amd import synth
from tests.helpers import make_case, oracle_problem, rel

pytestmark = pytest.mark.gpu

NT = max(1, min(16, os.cpu_count() or 1))


def _gpu(cfg, ds, precision=sanepic_amd.FP64, hwp=0, omega=0.0):
    ctx = sanepic_amd.Context(cfg.nside, cfg.ndet, cfg.nseg, cfg.ns, hwp=hwp, omega=omega,
                              precision=precision)
    ctx.load(cfg, ds)
    npix, ndef = ctx.pixelize()
    ctx.setup()
    return ctx, npix, ndef


def _run(ctx, n, **kw):
    ctx.solve_begin(**kw)
    return [ctx.step() for _ in range(n)], ctx.map()


def _check_setup(ctx, P, npix):
    dp, _, _, ind = P.pixels()
    assert npix == P.npix
    assert np.array_equal(ctx.all_pixels(), dp)
    assert np.array_equal(ctx.indpix(), ind)
    _, hits, mp, sc = P.setup_pass()
    ghits, gmp, gsc = ctx.get_setup()
    assert np.array_equal(ghits, hits)
    assert rel(gsc, sc) < 1e-12
    assert rel(gmp, mp) < 1e-12
    return hits


# --------------------------------------------------------------------------
# 1 detector, NSIDE 64, 2^22 samples in 8 segments of 2^19 (full size)
# --------------------------------------------------------------------------

@pytest.fixture(scope="module")
def one_det():
    cfg = synth.workload_config("nside64_1det")
    print(f"full single-detector case: {cfg.ndet} det x {cfg.nseg} x {cfg.ns}, NSIDE {cfg.nside}", flush=True)
    ds = synth.make_dataset(cfg, workers=NT)
    P = oracle_problem(cfg, ds, nthreads=NT)
    return cfg, ds, P


def test_one_detector_full_size_operator(one_det):
    cfg, ds, P = one_det
    ctx, npix, ndef = _gpu(cfg, ds)
    print(f"  npix {npix}, deferred samples {ndef}", flush=True)
    _check_setup(ctx, P, npix)
    x = np.random.default_rng(5).standard_normal(3 * npix)
    e_b, e_a = rel(ctx.rhs(), P.rhs()), rel(ctx.apply(x), P.apply_A(x))
    c_or = P.chi2(x)
    e_c = abs(ctx.chi2(x) - c_or) / c_or
    e_bl = rel(ctx.blocks(), P.exact_blocks())
    print(f"  b {e_b:.1e}, A x {e_a:.1e}, chi2 {e_c:.1e}, 3x3 blocks {e_bl:.1e}", flush=True)
    assert e_b < 1e-11 and e_a < 1e-11 and e_c < 1e-10 and e_bl < 1e-10
    ctx.close()


def test_one_detector_full_size_iterates(one_det):
    """First 8 iterates of the reference's diagonal PCG and of block-Jacobi
    with and without the monopole deflation, against the oracle's, within
    the oracle's own 1-vs-NT-thread reordering spread (block-Jacobi: at least
    1e-6, since half of this workload's 3x3 blocks are within rounding of
    singular and their inverses amplify the GPU's different summation order;
    an indexing or sign error moves the iterates by O(1))."""
    cfg, ds, P = one_det
    P1 = oracle_problem(cfg, ds, nthreads=1)
    ctx, npix, _ = _gpu(cfg, ds)
    K = 8
    ref = P.solve(stop=0.0, maxiter=K)
    spread = rel(P1.solve(stop=0.0, maxiter=K)["S"], ref["S"])
    h, S = _run(ctx, K, precond=sanepic_amd.PRECOND_DIAG)
    e = rel(S, ref["S"])
    print(f"  diagonal PCG: {K} iterates, gpu-oracle {e:.2e}, oracle spread {spread:.2e}", flush=True)
    assert e < max(100 * spread, 1e-10)
    assert np.allclose([v["crit"] for v in h], ref["crit"], rtol=max(100 * spread, 1e-8), atol=0)
    for deflate in (False, True):
        ref = P.solve_pcg(kind=1, deflate=deflate, stop=0.0, maxiter=K)
        spread = rel(P1.solve_pcg(kind=1, deflate=deflate, stop=0.0, maxiter=K)["S"], ref["S"])
        h, S = _run(ctx, K, precond=sanepic_amd.PRECOND_BLOCK3, deflate=deflate)
        e = rel(S, ref["S"])
        print(f"  block-Jacobi deflate={deflate}: gpu-oracle {e:.2e}, oracle spread {spread:.2e}",
              flush=True)
        assert e < max(100 * spread, 1e-6)
        assert np.allclose([v["crit"] for v in h], ref["crit"], rtol=max(100 * spread, 1e-6), atol=0)
    ctx.close()


def test_one_detector_full_size_block_jacobi_solve(one_det):
    """Block-Jacobi (+ deflation) PCG, 150 iterations.  With one detector
    whose angle is the spin phase, a pixel is crossed at nearly one angle:
    more than half of the pixels' 3x3 blocks have a reciprocal condition
    number below 1e-6, so Q and U are nearly undetermined there and the
    solve stalls near |r|/|b| ~ 1e-4 (the oracle's own PCG stalls the same
    way, and never reaches 1e-6 in 4000 iterations).  Parity: the GPU's
    150-iteration map solves the oracle's system as well as the oracle's
    own 150-iteration map does, both residuals measured with the oracle's
    operator."""
    cfg, ds, P = one_det
    ctx, npix, _ = _gpu(cfg, ds)
    K = 150
    h, S = _run(ctx, K, precond=sanepic_amd.PRECOND_BLOCK3, deflate=True)
    ref = P.solve_pcg(kind=1, deflate=True, stop=0.0, maxiter=K)
    hits = ref["hits"]
    m = hits[npix:] < 4

    def resid(x):
        ax = P.apply_A(x)
        ax[npix:2 * npix][m] = 0.0
        ax[2 * npix:][m] = 0.0
        return np.linalg.norm(ref["b"] - ax) / np.linalg.norm(ref["b"])

    rg, ro = resid(S), resid(ref["S"])
    print(f"  {K} iterations: oracle residual of the GPU map {rg:.2e}, of the oracle's map {ro:.2e}; "
          f"GPU true residual {ctx.true_residual():.2e}, crit {h[-1]['crit']:.2e} "
          f"(oracle {ref['crit'][-1]:.2e})", flush=True)
    assert rg <= 1.5 * ro and ro <= 1.5 * rg
    assert 0.5 <= h[-1]["crit"] / ref["crit"][-1] <= 2.0
    ctx.close()


# --------------------------------------------------------------------------
# NSIDE 512, segments of 2^20 (the 16-detector workload's kernels)
# --------------------------------------------------------------------------

@pytest.fixture(scope="module")
def n512():
    cfg = synth.Config(nside=512, ndet=2, nseg=2, ns=2 ** 20, det_angles_deg=[0.0, 112.5],
                       noise_seed0=1003)
    print(f"NSIDE 512: {cfg.ndet} det x {cfg.nseg} x 2^20 (workload scan)", flush=True)
    ds = synth.make_dataset(cfg, workers=NT)
    P = oracle_problem(cfg, ds, nthreads=NT)
    return cfg, ds, P


def test_nside512_segments_2e20_fp64(n512):
    cfg, ds, P = n512
    ctx, npix, ndef = _gpu(cfg, ds)
    print(f"  npix {npix}, deferred samples {ndef}", flush=True)
    _check_setup(ctx, P, npix)
    x = np.random.default_rng(6).standard_normal(3 * npix)
    e_b, e_a = rel(ctx.rhs(), P.rhs()), rel(ctx.apply(x), P.apply_A(x))
    c_or = P.chi2(x)
    e_c = abs(ctx.chi2(x) - c_or) / c_or
    print(f"  FP64: b {e_b:.1e}, A x {e_a:.1e}, chi2 {e_c:.1e}", flush=True)
    assert e_b < 1e-11 and e_a < 1e-11 and e_c < 1e-10
    # the headline's kernels: P^T over run records, 2psi as 64-bit turns
    assert ctx.run_records()[1] and ctx.fp64_angle_bytes() == 8 and ctx.fp64_sorted_angle_bytes() == 0
    ctx.close()


def test_nside512_segments_2e20_mixed(n512):
    cfg, ds, P = n512
    ctx, npix, _ = _gpu(cfg, ds, precision=sanepic_amd.MIXED)
    x = np.random.default_rng(6).standard_normal(3 * npix)
    e_a = rel(ctx.apply(x), P.apply_A(x))
    e_b = rel(ctx.rhs(), P.rhs())
    print(f"  mixed: A x {e_a:.2e} (FP32 operator), b {e_b:.1e} (FP64)", flush=True)
    assert 1e-9 < e_a < 2e-6 and e_b < 1e-11
    assert ctx.run_records()[1]
    ctx.close()


def test_pixels_of_one_detector_at_nside512():
    """One detector of the 16-detector workload (angle 112.5 deg): its full
    2^26-sample pointing at NSIDE 512 in 64 segments of 2^20, every RING
    pixel bit-exact and the seen-pixel compaction identical."""
    nside, nseg, ns = 512, 64, 2 ** 20
    ra, dec, psi = synth.scan_pointing(nseg * ns)
    ispf = synth.ispf_spectrum(ns, 19.0, 50e-6, 0.1, 1.0)
    print("one detector's pixels: 1 det x 64 x 2^20 at NSIDE 512", flush=True)
    ctx = sanepic_amd.Context(nside, 1, nseg, ns)
    ctx.set_detector(0, 0.0, 0.0, 112.5, ispf)
    for s in range(nseg):
        sl = slice(s * ns, (s + 1) * ns)
        ctx.set_pointing(s, ra[sl], dec[sl], psi[sl])
    npix, ndef = ctx.pixelize()
    P = oracle.Problem(nside, 1, nseg, ns, nthreads=NT)
    P.set_detector(0, 0.0, 0.0, 112.5, ispf)
    P.pixelize(ra, dec, psi)
    dp, _, _, ind = P.pixels()
    print(f"  npix {npix}, {ndef} samples resolved on the host", flush=True)
    assert npix == P.npix
    assert np.array_equal(ctx.indpix(), ind)
    for s in range(nseg):
        assert np.array_equal(ctx.pixels(0, s), dp[s * ns:(s + 1) * ns]), f"segment {s}"
    ctx.close()


# --------------------------------------------------------------------------
# per-detector segment lengths m 2^l, l = 16..21 (the uneven workload's)
# --------------------------------------------------------------------------

def _uneven_oracle(cfg, ds):
    groups = {}
    for d, L in enumerate(cfg.seglen):
        groups.setdefault(L, []).append(d)
    probs = []
    for L, dets in groups.items():
        nseg = cfg.nsamp // L
        n = nseg * L
        P = oracle.Problem(cfg.nside, len(dets), nseg, L, nthreads=NT)
        for j, d in enumerate(dets):
            P.set_detector(j, 0.0, 0.0, cfg.det_angles_deg[d], ds["ispf"][d])
        P.pixelize(ds["ra"][:n], ds["dec"][:n], ds["psi"][:n])
        P.load_tod(np.stack([ds["tod"][d] for d in dets]))
        probs.append((L, dets, P))
    ring = np.unique(np.concatenate([np.flatnonzero(P.pixels()[3] >= 0) for _, _, P in probs]))
    return probs, ring


def _gidx(P, ring):
    ind = P.pixels()[3]
    mine = np.flatnonzero(ind >= 0)
    mine = mine[np.argsort(ind[mine])]
    return np.searchsorted(ring, mine)


def _sum_groups(probs, ring, fn, x=None):
    npix = ring.size
    out = np.zeros(3 * npix)
    for _, _, P in probs:
        gi = _gidx(P, ring)
        xg = None if x is None else np.concatenate([x[c * npix + gi] for c in range(3)])
        yg = fn(P, xg)
        for c in range(3):
            out[c * npix + gi] += yg[c * P.npix:(c + 1) * P.npix]
    return out


@pytest.mark.parametrize("radix", [1, 3, 5])
def test_uneven_workload_lengths(radix):
    seglen = [radix * 2 ** l for l in range(16, 22)]
    nsamp = radix * 2 ** 21
    rng = np.random.default_rng(20 + radix)
    cfg = synth.UnevenConfig(256, seglen, nsamp, fknee=list(rng.uniform(0.02, 0.5, len(seglen))),
                             alpha=list(rng.uniform(0.8, 2.0, len(seglen))))
    print(f"uneven lengths {seglen} on {nsamp} samples, NSIDE {cfg.nside}", flush=True)
    ds = synth.make_uneven_dataset(cfg, workers=NT)
    probs, ring = _uneven_oracle(cfg, ds)
    x = np.random.default_rng(7).standard_normal(3 * ring.size)
    ax_ref = _sum_groups(probs, ring, lambda P, xg: P.apply_A(xg), x)
    b_ref = _sum_groups(probs, ring, lambda P, xg: P.rhs())
    for prec in (sanepic_amd.FP64, sanepic_amd.MIXED):
        ctx = sanepic_amd.Context.uneven(cfg.nside, cfg.seglen, cfg.nsamp, precision=prec)
        ctx.load_uneven(cfg, ds)
        ctx.pixelize()
        ctx.setup()
        assert ctx.npix == ring.size
        assert np.array_equal(np.flatnonzero(ctx.indpix() >= 0), ring)
        if prec == sanepic_amd.FP64:
            for L, dets, P in probs:
                dp = P.pixels()[0].reshape(len(dets), -1)
                for j, d in enumerate(dets):
                    got = np.concatenate([ctx.pixels(d, s) for s in range(nsamp // L)])
                    assert np.array_equal(got, dp[j]), f"pixels of length {L}"
            href = np.zeros(ring.size, dtype=np.int64)
            for _, _, P in probs:
                dp, _, _, ind = P.pixels()
                href[_gidx(P, ring)] += 2 * np.bincount(ind[dp], minlength=P.npix)
            assert np.array_equal(ctx.get_setup()[0][:ring.size], href)
        e_a, e_b = rel(ctx.apply(x), ax_ref), rel(ctx.rhs(), b_ref)
        print(f"  precision {prec}: A x {e_a:.2e}, b {e_b:.1e}", flush=True)
        assert e_b < 1e-11
        assert e_a < (1e-11 if prec == sanepic_amd.FP64 else 2e-6)
        ctx.close()


# --------------------------------------------------------------------------
# pixels with more than 4096 samples: P^T items split and recombined
# --------------------------------------------------------------------------

def test_split_pixels_uniform():
    cfg, ds = make_case(nside=4, ndet=2, nseg=4, ns=2 ** 16)
    P = oracle_problem(cfg, ds, nthreads=NT)
    ctx, npix, _ = _gpu(cfg, ds)
    hits = _check_setup(ctx, P, npix)
    print(f"  most samples in one pixel: {hits[:npix].max() // 2}", flush=True)
    assert hits[:npix].max() // 2 > 3 * 4096
    x = np.random.default_rng(8).standard_normal(3 * npix)
    assert rel(ctx.rhs(), P.rhs()) < 1e-11
    assert rel(ctx.apply(x), P.apply_A(x)) < 1e-11
    assert rel(ctx.blocks(), P.exact_blocks()) < 1e-10
    ctx.close()


def test_split_pixels_uneven():
    seglen = [2 ** 15, 3 * 2 ** 13, 2 ** 16, 5 * 2 ** 13]
    nsamp = 2 ** 17 + 2 ** 16
    tmin = nsamp / 19.0 / 60.0
    cfg = synth.UnevenConfig(4, seglen, nsamp, spin_min=tmin / 40, prec_min=tmin / 5)
    ds = synth.make_uneven_dataset(cfg, workers=1)
    probs, ring = _uneven_oracle(cfg, ds)
    ctx = sanepic_amd.Context.uneven(cfg.nside, cfg.seglen, cfg.nsamp)
    ctx.load_uneven(cfg, ds)
    ctx.pixelize()
    ctx.setup()
    hits, mp, sc = ctx.get_setup()
    npix = ring.size
    print(f"  most samples in one pixel: {hits[:npix].max() // 2}", flush=True)
    assert hits[:npix].max() // 2 > 3 * 4096
    x = np.random.default_rng(9).standard_normal(3 * npix)
    assert rel(ctx.apply(x), _sum_groups(probs, ring, lambda P, xg: P.apply_A(xg), x)) < 1e-11
    assert rel(ctx.rhs(), _sum_groups(probs, ring, lambda P, xg: P.rhs())) < 1e-11
    sc_ref = np.zeros(3 * npix)
    mp_ref = np.zeros(3 * npix)
    for _, _, Pg in probs:
        _, _, mpg, scg = Pg.setup_pass()
        gi = _gidx(Pg, ring)
        for c in range(3):
            sc_ref[c * npix + gi] += scg[c * Pg.npix:(c + 1) * Pg.npix]
            mp_ref[c * npix + gi] += mpg[c * Pg.npix:(c + 1) * Pg.npix]
    assert rel(sc, sc_ref) < 1e-12
    assert rel(mp[:npix], mp_ref[:npix]) < 1e-12
    ctx.close()


# --------------------------------------------------------------------------
# detector offsets and the half-wave plate
# --------------------------------------------------------------------------

OFFSETS = [(0.5, -0.5), (-0.5, 0.25), (0.0, 0.5)]


@pytest.mark.parametrize("ns", [1024, 2 ** 14])
@pytest.mark.parametrize("hwp", [0, 1])
def test_offsets_and_hwp(ns, hwp):
    omega = 0.0123456789 if hwp else 0.0
    cfg, ds = make_case(nside=16, ndet=3, nseg=3, ns=ns, offsets=OFFSETS)
    print(f"offsets {OFFSETS}, hwp {hwp}, ns {ns}", flush=True)
    P = oracle_problem(cfg, ds, nthreads=NT, hwp=hwp, omega=omega)
    ctx, npix, _ = _gpu(cfg, ds, hwp=hwp, omega=omega)
    _check_setup(ctx, P, npix)   # sumCross uses the angle without the plate (psipFP)
    x = np.random.default_rng(10).standard_normal(3 * npix)
    assert rel(ctx.rhs(), P.rhs()) < 1e-11
    assert rel(ctx.apply(x), P.apply_A(x)) < 1e-11
    c_or = P.chi2(x)
    assert abs(ctx.chi2(x) - c_or) / c_or < 1e-10
    ctx.close()


# --------------------------------------------------------------------------
# block-Jacobi against the oracle on the small cases
# --------------------------------------------------------------------------

SMALL = [dict(nside=8, ndet=2, nseg=3, ns=1024), dict(nside=16, ndet=2, nseg=2, ns=2 ** 13),
         dict(nside=32, ndet=1, nseg=3, ns=2 ** 16)]


@pytest.mark.parametrize("case", SMALL, ids=lambda c: "n{nside}_d{ndet}_s{nseg}_l{ns}".format(**c))
def test_block_jacobi_against_oracle(case):
    """Exact blocks to 1e-10; 8 block-Jacobi iterates within 100x the
    oracle's reordering spread, at least 1e-6 (nearly singular blocks of
    one-detector cases amplify summation-order differences)."""
    cfg, ds = make_case(**case)
    P1 = oracle_problem(cfg, ds, nthreads=1)
    Pn = oracle_problem(cfg, ds, nthreads=4)
    ctx, npix, _ = _gpu(cfg, ds)
    assert rel(ctx.blocks(), Pn.exact_blocks()) < 1e-10
    K = 8
    for deflate in (False, True):
        ref = P1.solve_pcg(kind=1, deflate=deflate, stop=0.0, maxiter=K)
        spread = rel(Pn.solve_pcg(kind=1, deflate=deflate, stop=0.0, maxiter=K)["S"], ref["S"])
        h, S = _run(ctx, K, precond=sanepic_amd.PRECOND_BLOCK3, deflate=deflate)
        e = rel(S, ref["S"])
        print(f"  deflate={deflate}: gpu-oracle {e:.2e}, oracle spread {spread:.2e}", flush=True)
        assert e < max(100 * spread, 1e-6)
        assert np.allclose([v["crit"] for v in h], ref["crit"], rtol=max(100 * spread, 1e-6), atol=0)
    ctx.close()


# --------------------------------------------------------------------------
# orbital dipole (-o)
# --------------------------------------------------------------------------

@pytest.mark.parametrize("ns", [1024, 2 ** 14])
def test_dipole_subtraction(ns):
    """The dipole leaves the data before b and chi2, after the TOD's NaN
    samples are filled (the oracle: fdatas0 -= r2c(orbdip), 700-720,
    1055-1067)."""
    cfg, ds = make_case(nside=16, ndet=2, nseg=3, ns=ns, dipole=True)
    tod = ds["tod"].copy()
    tod[0, [10, 11, ns + 5]] = np.nan
    ds = dict(ds, tod=tod)
    P = oracle_problem(cfg, ds, nthreads=NT)
    ctx, npix, _ = _gpu(cfg, ds)
    x = np.random.default_rng(12).standard_normal(3 * npix) * 1e-4
    e_b = rel(ctx.rhs(), P.rhs())
    c_or = P.chi2(x)
    e_c = abs(ctx.chi2(x) - c_or) / c_or
    print(f"  dipole: b {e_b:.1e}, chi2 {e_c:.1e}", flush=True)
    assert e_b < 1e-11 and e_c < 1e-10
    ctx.close()
    # a dipole for some blocks only is refused
    ctx = sanepic_amd.Context(cfg.nside, cfg.ndet, cfg.nseg, cfg.ns)
    ctx.load(cfg, dict(ds, odip=None))
    ctx.set_dipole(0, 0, ds["odip"][0, :ns])
    ctx.pixelize()
    with pytest.raises(sanepic_amd.SanepicError, match="dipole block"):
        ctx.setup()
    ctx.close()


def test_dipole_subtraction_uneven():
    seglen = [8192, 3 * 8192, 16384]
    nsamp = 3 * 16384
    tmin = nsamp / 19.0 / 60.0
    cfg = synth.UnevenConfig(16, seglen, nsamp, spin_min=tmin / 40, prec_min=tmin / 5)
    ds = synth.make_uneven_dataset(cfg, workers=1)
    od = synth.orbital_dipole(ds["ra"], ds["dec"])
    for d in range(cfg.ndet):
        ds["tod"][d][:] += od[:cfg.used(d)]
    probs, ring = _uneven_oracle(cfg, ds)
    for L, dets, P in probs:
        P.set_dipole(np.concatenate([od[:P.nseg * L] for _ in dets]))
    ctx = sanepic_amd.Context.uneven(cfg.nside, cfg.seglen, cfg.nsamp)
    ctx.load_uneven(cfg, ds)
    for d in range(cfg.ndet):
        h = cfg.used(d) // 2
        ctx.set_dipole_range(d, 0, od[:h])
        ctx.set_dipole_range(d, h, od[h:cfg.used(d)])
    ctx.pixelize()
    ctx.setup()
    assert rel(ctx.rhs(), _sum_groups(probs, ring, lambda P, xg: P.rhs())) < 1e-11
    ctx.close()
