"""Relative calibration (-r; sanepic.cc:1041-1067, 2114-2316, 2543-2545,
3206-3255) on the GPU against the oracle (itself pinned by
test_oracle_recal.py): b, A x and chi2 with per-block gains, the gain-fit
sums, a warm-started solve with the first pass's |b|^2, and the drop-in's
whole loop (maps of every pass, relCalib.dat<pass>, chi2 lines)."""
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


def _pair(ns, nseg=4, precision=sanepic_amd.FP64, dipole=True):
    cfg, ds = make_case(nside=8, ndet=2, nseg=nseg, ns=ns, dipole=dipole)
    P = oracle_problem(cfg, ds)
    ctx = sanepic_amd.Context(cfg.nside, cfg.ndet, cfg.nseg, cfg.ns, precision=precision)
    ctx.load(cfg, ds)
    ctx.pixelize()
    ctx.set_recalib(True)
    ctx.setup()
    return cfg, ds, P, ctx


def _gains(cfg, seed=3):
    return np.random.default_rng(seed).uniform(0.8, 1.25, (cfg.ndet, cfg.nseg))


@pytest.mark.parametrize("ns", [1024, 3 * 2 ** 13, 2 ** 14])
def test_gains_rhs_operator_chi2_sums(ns):
    print("recalibration, ns", ns, flush=True)
    cfg, ds, P, ctx = _pair(ns)
    npix = ctx.npix
    x = np.random.default_rng(4).standard_normal(3 * npix) * 1e-4
    # unit gains: the plain problem
    assert rel(ctx.rhs(), P.rhs()) < 1e-11
    g = _gains(cfg)
    ctx.set_gains(g)
    P.set_gains(g)
    e_b, e_a = rel(ctx.rhs(), P.rhs()), rel(ctx.apply(x), P.apply_A(x))
    c_or = P.chi2(x)
    e_c = abs(ctx.chi2(x) - c_or) / c_or
    s_g, s_o = ctx.recal_sums(x), P.recal_sums(x)
    e_s = np.abs(s_g - s_o).max() / np.abs(s_o).max()
    print(f"  b {e_b:.1e}, A x {e_a:.1e}, chi2 {e_c:.1e}, fit sums {e_s:.1e}", flush=True)
    assert e_b < 1e-11 and e_a < 1e-11 and e_c < 1e-10 and e_s < 1e-10
    # each sum on its own scale (zeta can be small against the data term)
    for k in range(3):
        assert rel(s_g[..., k], s_o[..., k]) < 1e-10
    ctx.close()


def test_gains_mixed_precision_operator():
    cfg, ds, P, c32 = _pair(2 ** 13, precision=sanepic_amd.MIXED)
    g = _gains(cfg, seed=5)
    c32.set_gains(g)
    P.set_gains(g)
    x = np.random.default_rng(6).standard_normal(3 * c32.npix)
    e = rel(c32.apply(x), P.apply_A(x))
    print(f"  mixed A x with gains {e:.2e}", flush=True)
    assert 1e-9 < e < 2e-6
    assert rel(c32.rhs(), P.rhs()) < 1e-11  # b is built in FP64
    c32.close()


def test_warm_solve_with_first_pass_norm():
    """A pass after the first: gains set, start from the previous map, crit
    divided by the first pass's |b|^2 (1377-1385)."""
    cfg, ds, P, ctx = _pair(2 ** 12)
    first = P.solve_pcg(kind=0, stop=1e-15, maxiter=2000)
    var00 = float(first["b"] @ first["b"])
    g = _gains(cfg, seed=7)
    P.set_gains(g)
    ctx.set_gains(g)
    K = 8
    ref = P.solve_pcg(kind=0, x0=first["S"], stop=0.0, maxiter=K, var00=var00)
    ctx.solve_begin(sanepic_amd.PRECOND_DIAG, 10, first["S"], var00=var00)
    crit = [ctx.step()["crit"] for _ in range(K)]
    e = rel(ctx.map(), ref["S"])
    print(f"  warm pass: crit {crit[0]:.3e} vs {ref['crit'][0]:.3e}, map after {K}: {e:.2e}", flush=True)
    assert np.allclose(crit, ref["crit"], rtol=1e-7)
    assert e < 1e-9
    ctx.close()


def _oracle_loop(P, npass, global_gain):
    """The drop-in's loop on the oracle: solve (block-Jacobi PCG, as the
    drop-in runs with SANEPIC_PRECOND=block; refresh 10, stop 1e-15 or 2000
    iterations), fit, next pass from the map."""
    g = np.ones((P.ndet, P.nseg))
    x, var00, out = None, 0.0, []
    for k in range(1, npass + 1):
        P.set_gains(g)
        sol = P.solve_pcg(kind=1, x0=x, stop=1e-15, maxiter=2000, var00=var00)
        if k == 1:
            var00 = float(sol["b"] @ sol["b"])
        x = sol["S"]
        s = P.recal_sums(x)
        before = P.recal_chi2(x, g)
        if global_gain:
            g = np.repeat((s[..., 0].sum(1) / s[..., 1].sum(1))[:, None], P.nseg, axis=1)
        else:
            g = s[..., 0] / s[..., 1]
        out.append(dict(S=x, g=g.copy(), before=before, after=P.recal_chi2(x, g), b=sol["b"]))
    return out


@pytest.mark.parametrize("global_gain", [0, 1])
def test_drop_in_recalibration_loop(tmp_path, global_gain):
    tmin = 4 * 2048 / 19.0 / 60.0
    cfg = synth.Config(nside=8, ndet=2, nseg=4, ns=2048, fknee=0.2, spin_min=tmin / 40,
                       prec_min=tmin / 5, dipole=True)
    ds = synth.make_dataset(cfg, workers=1)
    # calibration errors of up to 10% per (det, seg) block in the timelines
    g_true = np.random.default_rng(21).uniform(0.9, 1.1, (cfg.ndet, cfg.nseg))
    tod = np.asarray(ds["tod"]).reshape(cfg.ndet, cfg.nseg, cfg.ns) * g_true[:, :, None]
    ds = dict(ds, tod=tod.reshape(cfg.ndet, cfg.nseg * cfg.ns))
    args = synth.write_tree(str(tmp_path), cfg, ds)
    npass = 3
    argv = synth.sanepic_argv(args, exe=sanepic_amd.EXE_PATH, odip=1) + ["-r", "1", "-R", str(global_gain)]
    print("running", " ".join(argv), flush=True)
    # block-Jacobi: this case needs ~1400 iterations of the diagonal PCG and
    # ~20 of block-Jacobi to reach crit 1e-15, so the maps of the two solvers
    # agree closely only with the latter
    env = dict(os.environ, SANEPIC_MAX_RLOOP=str(npass), SANEPIC_PRECOND="block")
    r = subprocess.run(argv, capture_output=True, text=True, env=env, timeout=240, cwd=str(tmp_path))
    assert r.returncode == 0, r.stderr[-2000:]
    before = [float(v) for v in re.findall(r"^Before recal chi2 = (\S+)", r.stdout, re.M)]
    after = [float(v) for v in re.findall(r"^After recal chi2 = (\S+)", r.stdout, re.M)]
    assert len(before) == npass and len(after) == npass, r.stdout[-3000:]
    P = oracle_problem(cfg, ds)
    ref = _oracle_loop(P, npass, global_gain)
    npix = P.npix
    O, E = args["out"], args["ext"]
    for k in range(1, npass + 1):
        A = np.fromfile(f"{O}_A_{E}_{k}", dtype="<f8")
        rc = np.fromfile(str(tmp_path / f"relCalib.dat{k}"), dtype="<f8")
        assert A.size == 3 * npix and rc.size == cfg.ndet * cfg.nseg
        g_or = ref[k - 1]["g"]
        g_gpu = rc.reshape(cfg.nseg, cfg.ndet).T  # relCalib[idet + iseg ndet]
        e_m, e_g = rel(A, ref[k - 1]["S"]), np.abs(g_gpu / g_or - 1).max()
        e_c = abs(before[k - 1] - ref[k - 1]["before"]) / ref[k - 1]["before"]
        print(f"  pass {k}: map {e_m:.2e}, gains {e_g:.2e}, chi2 before {e_c:.2e} "
              f"({before[k - 1]:.6e} -> {after[k - 1]:.6e})", flush=True)
        assert e_m < 1e-6 and e_g < 1e-8 and e_c < 1e-8
        assert after[k - 1] <= before[k - 1] * (1 + 1e-12)
        if global_gain:
            assert np.allclose(g_gpu, g_gpu[:, :1], rtol=0, atol=0)
        assert os.path.exists(f"{O}_I_{E}_{k}") and os.path.exists(f"{O}_U_{E}_{k}")
    if not global_gain:  # the per-block fit finds the injected gains (to the noise)
        assert np.abs(g_gpu / g_true - 1).max() < 0.05
        assert np.abs(g_true - 1).max() > 0.05
    # passes after the first start from the previous map and print var0 / var00
    v0 = [float(v) for v in re.findall(r"^var0 = (\S+)", r.stdout, re.M)]
    v00 = [float(v) for v in re.findall(r"^var00 = (\S+)", r.stdout, re.M)]
    assert len(v0) == npass and len(set(v00)) == 1
    assert v0[0] == v00[0] and v0[1] < 1e-3 * v00[0]


def test_recalibration_error_paths():
    """Gains that are zero or not finite, fits without the resident TOD,
    a negative crit denominator and per-detector lengths are refused with an
    error, not a wrong map."""
    cfg, ds = make_case(nside=8, ndet=2, nseg=2, ns=1024, dipole=True)
    ctx = sanepic_amd.Context(cfg.nside, cfg.ndet, cfg.nseg, cfg.ns)
    ctx.load(cfg, ds)
    ctx.pixelize()
    ctx.setup()
    x = np.zeros(3 * ctx.npix)
    with pytest.raises(sanepic_amd.SanepicError, match="relative calibration is off"):
        ctx.recal_sums(x)
    with pytest.raises(sanepic_amd.SanepicError, match="relative calibration is off"):
        ctx.set_gains(np.ones((cfg.ndet, cfg.nseg)))
    with pytest.raises(sanepic_amd.SanepicError, match="before sanepic_setup"):
        ctx.set_recalib(True)
    with pytest.raises(sanepic_amd.SanepicError, match="var00"):
        ctx.solve_begin(sanepic_amd.PRECOND_DIAG, 10, var00=-1.0)
    ctx.close()
    cfg, ds, P, ctx = _pair(1024, nseg=2)
    b0 = ctx.rhs()
    for bad in (0.0, np.nan, np.inf):
        g = np.ones((cfg.ndet, cfg.nseg))
        g[1, 0] = bad
        with pytest.raises(sanepic_amd.SanepicError, match="zero or not finite"):
            ctx.set_gains(g)
    assert np.array_equal(ctx.rhs(), b0)  # a refused call leaves b as it was
    ctx.close()
    u = synth.UnevenConfig(8, [1024, 2048], 4096, fknee=[0.2, 0.2], alpha=[1.0, 1.0])
    cu = sanepic_amd.Context.uneven(u.nside, u.seglen, u.nsamp)
    with pytest.raises(sanepic_amd.SanepicError, match="one segment length"):
        cu.set_recalib(True)
    cu.close()
