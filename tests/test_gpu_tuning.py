"""Launch-geometry choices must not change results: every four-step split
gives the oracle's operator and right-hand side, the autotuner's cache is
honoured, and a HIP-graph replay of the PCG iteration reproduces the eager
launches bit for bit."""
import os

import numpy as np
import pytest

import sanepic_amd
from sanepic_amd import synth
from tests.helpers import make_case, oracle_problem, rel

pytestmark = pytest.mark.gpu

CASE = dict(nside=32, ndet=2, nseg=3, ns=2 ** 16)


@pytest.fixture(scope="module")
def data():
    cfg, ds = make_case(**CASE)
    P = oracle_problem(cfg, ds)
    return cfg, ds, P


def _ctx(cfg, ds, split=None, precision=sanepic_amd.FP64):
    ctx = sanepic_amd.Context(cfg.nside, cfg.ndet, cfg.nseg, cfg.ns, precision=precision)
    ctx.load(cfg, ds)
    ctx.pixelize()
    if split is not None:
        ctx.set_fft_split(split)
    ctx.setup()
    return ctx


@pytest.mark.parametrize("split", [9, 10, 11, 12])
def test_every_split_matches_oracle(data, split):
    cfg, ds, P = data
    print("case", CASE, "log2 N1", split, flush=True)
    ctx = _ctx(cfg, ds, split)
    x = np.random.default_rng(7).standard_normal(3 * ctx.npix)
    assert rel(ctx.apply(x), P.apply_A(x)) < 1e-11
    assert rel(ctx.rhs(), P.rhs()) < 1e-11
    c_or = P.chi2(x)
    assert abs(ctx.chi2(x) - c_or) / c_or < 1e-10
    ctx.close()


def test_split_out_of_range_is_rejected(data):
    cfg, ds, _ = data
    ctx = sanepic_amd.Context(cfg.nside, cfg.ndet, cfg.nseg, cfg.ns)
    ctx.load(cfg, ds)
    ctx.pixelize()
    with pytest.raises(sanepic_amd.SanepicError):
        ctx.set_fft_split(13)
    with pytest.raises(sanepic_amd.SanepicError):
        ctx.set_fft_split(8)
    ctx.close()


def test_autotune_and_cache(data, tmp_path):
    cfg, ds, P = data
    cache = str(tmp_path / "tune.txt")
    ctx = sanepic_amd.Context(cfg.nside, cfg.ndet, cfg.nseg, cfg.ns)
    ctx.load(cfg, ds)
    ctx.pixelize()
    l1, ms = ctx.autotune(cache)
    print(f"  autotune: log2 N1 = {l1}, {ms:.3f} ms per operator", flush=True)
    assert 9 <= l1 <= 12 and ms > 0
    ctx.setup()
    x = np.random.default_rng(8).standard_normal(3 * ctx.npix)
    assert rel(ctx.apply(x), P.apply_A(x)) < 1e-11
    ctx.close()
    text = open(cache).read()
    assert f"log2_n1={l1}" in text and "ns=65536" in text
    # a second run takes the cached choice
    ctx = sanepic_amd.Context(cfg.nside, cfg.ndet, cfg.nseg, cfg.ns)
    ctx.load(cfg, ds)
    ctx.pixelize()
    assert ctx.autotune(cache)[0] == l1
    ctx.close()
    assert open(cache).read() == text


@pytest.mark.parametrize("precision", [sanepic_amd.FP64, sanepic_amd.MIXED])
def test_graph_replay_matches_eager(data, precision):
    cfg, ds, _ = data
    print("graph replay, precision", precision, flush=True)
    ctx = _ctx(cfg, ds, precision=precision)
    y = np.random.default_rng(9).standard_normal(3 * ctx.npix)
    runs = []
    for graph in (False, True):
        ctx.solve_begin(sanepic_amd.PRECOND_BLOCK3, 10, None, True, graph=graph)
        h = []
        for k in range(25):  # 3 refresh iterations, many replays of each graph
            h.append(ctx.step())
            if graph and k % 4 == 1:
                ctx.apply(y)  # eager work on the same stream between replays
        runs.append((np.array([v["crit"] for v in h]), np.array([v["alpha"] for v in h]), ctx.map()))
    (c0, a0, x0), (c1, a1, x1) = runs
    assert np.array_equal(c0, c1) and np.array_equal(a0, a1)
    assert np.array_equal(x0, x1)
    ctx.close()


@pytest.mark.parametrize("uneven", [False, True])
@pytest.mark.parametrize("pregather", ["0", "1"])
def test_graph_solve_on_a_fresh_context(data, uneven, pregather, monkeypatch):
    """A solve whose very first iteration is captured (graph mode, no
    deflation, no starting map, so no operator application has run before
    the capture): every buffer the operator uses exists after set-up, and
    the replayed iterates equal the eager ones bit for bit."""
    cfg, ds, _ = data
    monkeypatch.setenv("SANEPIC_PREGATHER", pregather)
    print(f"graph from a fresh context, uneven {uneven}, pre-gather {pregather}", flush=True)
    runs = []
    for graph in (False, True):
        if uneven:
            L = [cfg.ns] * (cfg.ndet - 1) + [cfg.ns // 2]
            ctx = sanepic_amd.Context.uneven(cfg.nside, L, cfg.nseg * cfg.ns)
            ctx.set_pointing_range(0, ds["ra"], ds["dec"], ds["psi"])
            for k in range(cfg.ndet):
                ctx.set_detector(k, 0.0, 0.0, cfg.det_angles_deg[k], synth.ispf_spectrum(
                    L[k], cfg.fsamp, cfg.sigma, cfg.fknee, cfg.alpha))
                ctx.set_tod_range(k, 0, ds["tod"][k])
            ctx.pixelize()
            ctx.setup()
        else:
            ctx = _ctx(cfg, ds)
        ctx.solve_begin(sanepic_amd.PRECOND_BLOCK3, 10, None, False, graph=graph)
        h = [ctx.step() for _ in range(22)]
        runs.append((np.array([v["crit"] for v in h]), ctx.map()))
        ctx.close()
    assert np.array_equal(runs[0][0], runs[1][0]) and np.array_equal(runs[0][1], runs[1][1])


@pytest.mark.parametrize("env", ["SANEPIC_ZSORT", "SANEPIC_XCD_BLOCKS", "SANEPIC_NESTED_MAP", "SANEPIC_ITEM_ORDER",
                                 "SANEPIC_PREGATHER"])
def test_layout_options_are_bit_identical(data, env, monkeypatch):
    """Writing the filter output in pixel-sorted order (read contiguously by
    P^T), the XCD-contiguous block numbering, the nested-order gather map and
    items, and gathering P x in a pass of its own change where values are
    stored and which L2 serves a block, not any sum:
    A x, b and chi2 are bit-identical with the option off, and the per-sample
    pixels reported back are the same."""
    cfg, ds, P = data
    print("layout option", env, flush=True)
    if env == "SANEPIC_ZSORT":  # sorted z replaces the run records: compare both on samples
        monkeypatch.setenv("SANEPIC_RUNS", "0")
    out = []
    for v in ("0", "1"):
        monkeypatch.setenv(env, v)
        for prec in (sanepic_amd.FP64, sanepic_amd.MIXED):
            ctx = _ctx(cfg, ds, precision=prec)
            x = np.random.default_rng(11).standard_normal(3 * ctx.npix)
            out.append((ctx.apply(x), ctx.rhs(), ctx.chi2(x) if prec == sanepic_amd.FP64 else 0.0,
                        ctx.pixels(1, 2)))
            ctx.close()
    for a, b in zip(out[:2], out[2:]):
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]) and a[2] == b[2]
        assert np.array_equal(a[3], b[3])
    assert rel(out[2][0], P.apply_A(np.random.default_rng(11).standard_normal(3 * out[2][0].size // 3))) < 1e-11


@pytest.mark.parametrize("split", [9, 12])
def test_run_records_match_samples(data, split, monkeypatch):
    """P^T over run records (the inverse column pass sums each run of
    consecutive samples in one pixel, P^T sums the runs) against P^T over
    samples and the oracle: the same sums in another order, so A x agrees to
    rounding; b, built from samples, agrees to rounding (P^T over samples
    reads the sorted angles as 64-bit turns where the operator uses it, as
    pairs where only the set-up does)."""
    cfg, ds, P = data
    x = np.random.default_rng(12).standard_normal(3 * P.npix)
    want = P.apply_A(x)
    for prec, tol in ((sanepic_amd.FP64, 1e-11), (sanepic_amd.MIXED, 2e-6)):
        got = {}
        for v in ("0", "1"):
            monkeypatch.setenv("SANEPIC_RUNS", v)
            ctx = _ctx(cfg, ds, split, precision=prec)
            got[v] = (ctx.apply(x), ctx.rhs(), ctx.device_bytes())
            nrec, used = ctx.run_records()
            assert used == (v == "1") and (nrec > 0) == (v == "1")
            if v == "1":  # every piece of 16 samples holds at least one run, at most 16
                assert cfg.ndet * cfg.nseg * cfg.ns // 16 <= nrec <= cfg.ndet * cfg.nseg * cfg.ns
            ctx.close()
        e_runs, e_samples = rel(got["1"][0], want), rel(got["0"][0], want)
        d = rel(got["1"][0], got["0"][0])
        print(f"  log2 N1 {split} {prec}: runs {e_runs:.2e}, samples {e_samples:.2e}, apart {d:.2e}, "
              f"records {(got['1'][2] - got['0'][2]) / 2**20:.1f} MiB", flush=True)
        assert e_runs < tol and e_samples < tol
        assert 0 < d < (1e-11 if prec == sanepic_amd.FP64 else 2e-6)
        assert rel(got["1"][1], got["0"][1]) < 1e-14


def test_monopole_image_without_fft(data, monkeypatch):
    """The deflated solver's A e_I taken as P^T of the constant N SpN[0] per
    detector (P e_I = 1, and the filter of a constant keeps only its DC bin)
    against a full FP64 operator application: the same iterates."""
    cfg, ds, P = data
    crit = {}
    for v in ("0", "1"):
        monkeypatch.setenv("SANEPIC_MONO_FFT", v)
        ctx = _ctx(cfg, ds)
        ctx.solve_begin(sanepic_amd.PRECOND_BLOCK3, 10, deflate=True)
        crit[v] = [ctx.step()["crit"] for _ in range(12)]
        ctx.close()
    print("  crit", crit["0"][:3], crit["1"][:3], flush=True)
    assert np.allclose(crit["0"], crit["1"], rtol=1e-9, atol=0)


def test_fp64_angles_as_turns(data, monkeypatch):
    """The FP64 operator reading 2psi as 64-bit turns (decoded by integer
    quarter-turn reduction and Taylor series) against reading the stored
    (cos, sin) pairs: the decoded pairs are within a few 1e-16, so A x agrees
    to rounding, and both match the oracle."""
    cfg, ds, P = data
    x = np.random.default_rng(13).standard_normal(3 * P.npix)
    want = P.apply_A(x)
    got = {}
    for v in ("0", "1"):
        monkeypatch.setenv("SANEPIC_TURN64", v)
        ctx = _ctx(cfg, ds)
        got[v] = (ctx.apply(x), ctx.chi2(x), ctx.device_bytes())
        assert ctx.fp64_angle_bytes() == (8 if v == "1" else 16)
        ctx.close()
    d = rel(got["1"][0], got["0"][0])
    print(f"  turns vs pairs {d:.2e}, vs oracle {rel(got['1'][0], want):.2e}", flush=True)
    assert d < 1e-14 and rel(got["1"][0], want) < 1e-11
    assert abs(got["1"][1] - got["0"][1]) <= 1e-12 * abs(got["0"][1])
    assert got["1"][2] > got["0"][2]  # the turns were made
