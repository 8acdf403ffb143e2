"""The GalBP leakage component (-b 1; sanepic.cc:551-575, 2860-2870,
2591-2607) on the GPU against the oracle (itself pinned by
test_oracle_leakage.py): maps [I|Q|U|L], detector d sees L times its
normalised coefficient."""
import os
import re
import subprocess

import numpy as np
import pytest

import oracle
import sanepic_amd
from sanepic_amd import synth
from tests.helpers import make_case, oracle_problem, rel

pytestmark = pytest.mark.gpu

RAW = np.array([1.7, 0.4, -0.9])


def normalised(raw):
    """(r - mean) / population std, the reference's 561-570."""
    m = raw.mean()
    return (raw - m) / np.sqrt(((raw - m) ** 2).mean())


def _pair(ns, precision=sanepic_amd.FP64):
    cfg, ds = make_case(nside=16, ndet=3, nseg=3, ns=ns)
    rel_ = normalised(RAW)
    P = oracle_problem(cfg, ds)
    P.set_leakage(rel_)
    ctx = sanepic_amd.Context(cfg.nside, cfg.ndet, cfg.nseg, cfg.ns, precision=precision)
    ctx.load(cfg, ds)
    ctx.set_leakage(rel_)
    ctx.pixelize()
    ctx.setup()
    return cfg, ds, P, ctx


@pytest.mark.parametrize("ns", [1024, 2 ** 14])
def test_leakage_operator_rhs_setup(ns):
    print("leakage component, ns", ns, flush=True)
    cfg, ds, P, ctx = _pair(ns)
    npix = ctx.npix
    assert npix == P.npix and ctx.nc == 4
    _, hits, mp, sc = P.setup_pass()
    ghits, gmp, gsc = ctx.get_setup()
    assert np.array_equal(ghits, hits)
    assert rel(gmp, mp) < 1e-12 and gmp.size == 4 * npix
    assert rel(gsc, sc[:3 * npix]) < 1e-12
    x = np.random.default_rng(4).standard_normal(4 * npix)
    e_a, e_b = rel(ctx.apply(x), P.apply_A(x)), rel(ctx.rhs(), P.rhs())
    c_or = P.chi2(x)
    e_c = abs(ctx.chi2(x) - c_or) / c_or
    print(f"  A x {e_a:.1e}, b {e_b:.1e}, chi2 {e_c:.1e}", flush=True)
    assert e_a < 1e-11 and e_b < 1e-11 and e_c < 1e-10
    assert np.array_equal(ctx.apply_noalloc(x), ctx.apply(x))
    ctx.close()
    _, _, P, c32 = _pair(ns, sanepic_amd.MIXED)
    e32 = rel(c32.apply(x), P.apply_A(x))
    print(f"  mixed: A x {e32:.2e}", flush=True)
    assert 1e-9 < e32 < 2e-6
    c32.close()


def test_leakage_pcg_iterates_and_nvar():
    cfg, ds, P, ctx = _pair(2 ** 13)
    P1 = oracle_problem(cfg, ds, nthreads=1)
    P1.set_leakage(normalised(RAW))
    K = 8
    ref = P.solve(stop=0.0, maxiter=K)
    spread = rel(P1.solve(stop=0.0, maxiter=K)["S"], ref["S"])
    S, h = ctx.solve(stop=0.0, maxiter=K)
    e = rel(S, ref["S"])
    print(f"  leakage PCG: {K} iterates, gpu-oracle {e:.2e}, oracle spread {spread:.2e}", flush=True)
    assert e < max(100 * spread, 1e-9)
    assert rel(ctx.nvar(), ref["Mptot"]) < 1e-12
    with pytest.raises(sanepic_amd.SanepicError):
        ctx.solve_begin(sanepic_amd.PRECOND_BLOCK3)
    ctx.close()


def test_leakage_rejections():
    cfg, ds = make_case(nside=8, ndet=2, nseg=2, ns=1024)
    ctx = sanepic_amd.Context(cfg.nside, cfg.ndet, cfg.nseg, cfg.ns)
    ctx.load(cfg, ds)
    with pytest.raises(sanepic_amd.SanepicError):
        ctx.set_leakage(np.array([np.nan, 1.0]))
    ctx.pixelize()
    with pytest.raises(sanepic_amd.SanepicError):  # after pixelisation
        ctx.set_leakage(np.array([1.0, -1.0]))
    ctx.close()
    u = sanepic_amd.Context.uneven(8, [1024, 2048], 4096)
    with pytest.raises(sanepic_amd.SanepicError):
        u.set_leakage(np.array([1.0, -1.0]))
    u.close()


def test_drop_in_leakage_matches_oracle(tmp_path):
    """-b 1 -K file: the coefficients file read and normalised, the PCG on
    [I|Q|U|L] following the oracle's, the _Leak_ maps written."""
    tmin = 3 * 2048 / 19.0 / 60.0
    cfg = synth.Config(nside=8, ndet=3, nseg=3, ns=2048, fknee=0.2, spin_min=tmin / 40, prec_min=tmin / 5)
    ds = synth.make_dataset(cfg, workers=1)
    args = synth.write_tree(str(tmp_path), cfg, ds)
    kfile = os.path.join(str(tmp_path), "bolo_info", "relfGal.bi")
    RAW.astype("<f8").tofile(kfile)
    argv = synth.sanepic_argv(args, exe=sanepic_amd.EXE_PATH, leakage_file=kfile, write_iter=1) + ["-j", "2"]
    r = subprocess.run(argv, capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, r.stderr[-2000:]
    got = [float(v) for v in re.findall(r"relfGal\[\d+\] = (\S+)", r.stdout)]
    assert np.allclose(got, normalised(RAW), rtol=1e-14)
    crit = np.array([float(m) for m in re.findall(r"^iter = \d+, crit = (\S+),", r.stdout, re.M)])
    P = oracle_problem(cfg, ds)
    P.set_leakage(normalised(RAW))
    ref = P.solve(stop=0.0, maxiter=5)
    assert rel(crit[:5], ref["crit"]) < 1e-8
    npix = P.npix
    _, _, _, ind = P.pixels()
    O, E = args["out"], args["ext"]
    A = np.fromfile(f"{O}_A_{E}_1", dtype="<f8")
    assert A.size == 4 * npix
    assert np.array_equal(np.fromfile(f"{O}_Leak_{E}_1", dtype="<f8"), oracle.expand_map(cfg.nside, ind, A[3 * npix:]))
    nv = np.fromfile(f"{O}_Leak_NVAR{E}", dtype="<f8")
    assert rel(nv, oracle.expand_map(cfg.nside, ind, ref["Mptot"][3 * npix:])) < 1e-12
    it2 = P.solve(stop=0.0, maxiter=3)["S"]
    assert rel(np.fromfile(f"{O}_Leak{E}_2", dtype="<f8"), oracle.expand_map(cfg.nside, ind, it2[3 * npix:])) < 1e-9
    # the written map solves the oracle's 4-component system
    full = P.solve(stop=1e-15, maxiter=2000)
    hits = full["hits"]
    ax = P.apply_A(A)
    ax[npix:2 * npix][hits[npix:] < 4] = 0.0
    ax[2 * npix:3 * npix][hits[npix:] < 4] = 0.0
    bres = full["b"]
    rg = np.linalg.norm(bres - ax) / np.linalg.norm(bres)
    ao = P.apply_A(full["S"])
    ao[npix:2 * npix][hits[npix:] < 4] = 0.0
    ao[2 * npix:3 * npix][hits[npix:] < 4] = 0.0
    ro = np.linalg.norm(bres - ao) / np.linalg.norm(bres)
    print(f"  drop-in with leakage: {crit.size} iterations, oracle residual of its map {rg:.2e} "
          f"(oracle's own {ro:.2e})", flush=True)
    assert rg <= 2.0 * ro + 1e-9
