"""Detectors with different (power-of-two and radix-3/5) segment lengths.

The reference requires one length; its operator for such data is the sum
over detectors of each detector's own segment filter, so the checker is the
sum of uniform oracle problems (one per length, over the timeline's first
nseg L samples) mapped into the union of seen pixels:
A x = sum_g S_g A_g S_g^T x and b = sum_g S_g b_g."""
import numpy as np
import pytest

import oracle
import sanepic_amd
from sanepic_amd import synth
from tests.helpers import rel

pytestmark = pytest.mark.gpu

SEGLEN = [8192, 3 * 8192, 5 * 8192, 16384, 8192]
NSAMP = 3 * 16384


def _cfg():
    tmin = NSAMP / 19.0 / 60.0
    return synth.UnevenConfig(32, SEGLEN, NSAMP, fknee=[0.2, 0.05, 0.3, 0.1, 0.4], alpha=[1.0, 1.5, 0.9, 2.0, 1.2],
                              spin_min=tmin / 40, prec_min=tmin / 5)


@pytest.fixture(scope="module")
def case():
    cfg = _cfg()
    ds = synth.make_uneven_dataset(cfg, workers=1)
    groups = {}
    for d, L in enumerate(cfg.seglen):
        groups.setdefault(L, []).append(d)
    probs = []
    for L, dets in groups.items():
        nseg = NSAMP // L
        n = nseg * L
        P = oracle.Problem(cfg.nside, len(dets), nseg, L)
        for j, d in enumerate(dets):
            P.set_detector(j, 0.0, 0.0, cfg.det_angles_deg[d], ds["ispf"][d])
        P.pixelize(ds["ra"][:n], ds["dec"][:n], ds["psi"][:n])
        P.load_tod(np.stack([ds["tod"][d] for d in dets]))
        probs.append((L, dets, P))
    ring = np.unique(np.concatenate([np.flatnonzero(P.pixels()[3] >= 0) for _, _, P in probs]))
    return cfg, ds, probs, ring


def _group_index(P, ring):
    ind = P.pixels()[3]
    mine = np.flatnonzero(ind >= 0)
    mine = mine[np.argsort(ind[mine])]
    return np.searchsorted(ring, mine)  # group pixel -> union pixel


def _sum_groups(probs, ring, fn, x=None):
    npix = ring.size
    out = np.zeros(3 * npix)
    for _, _, P in probs:
        gi = _group_index(P, ring)
        xg = None if x is None else np.concatenate([x[c * npix + gi] for c in range(3)])
        yg = fn(P, xg)
        for c in range(3):
            out[c * npix + gi] += yg[c * P.npix:(c + 1) * P.npix]
    return out


def _ctx(cfg, ds, precision=sanepic_amd.FP64, queued=False):
    ctx = sanepic_amd.Context.uneven(cfg.nside, cfg.seglen, cfg.nsamp, precision=precision)
    ctx.load_uneven(cfg, ds, queued=queued)
    ctx.pixelize()
    ctx.setup()
    return ctx


@pytest.mark.parametrize("queued", [False, True])
def test_uneven_pixels_hits_operator(case, queued):
    cfg, ds, probs, ring = case
    print("uneven lengths", SEGLEN, "queued", queued, flush=True)
    ctx = _ctx(cfg, ds, queued=queued)
    assert ctx.npix == ring.size
    ind = ctx.indpix()
    assert np.array_equal(np.flatnonzero(ind >= 0), ring)
    for L, dets, P in probs:
        dp = P.pixels()[0].reshape(len(dets), -1)
        for j, d in enumerate(dets):
            assert ctx.seglen(d) == (L, NSAMP // L)
            got = np.concatenate([ctx.pixels(d, s) for s in range(NSAMP // L)])
            assert np.array_equal(got, dp[j])
    hits, mp, sc = ctx.get_setup()
    href = np.zeros(ring.size, dtype=np.int64)
    for _, _, P in probs:
        href[_group_index(P, ring)] += P.setup_pass()[1][:P.npix]
    assert np.array_equal(hits[:ring.size], href)
    x = np.random.default_rng(3).standard_normal(3 * ring.size)
    e_op = rel(ctx.apply(x), _sum_groups(probs, ring, lambda P, xg: P.apply_A(xg), x))
    e_b = rel(ctx.rhs(), _sum_groups(probs, ring, lambda P, xg: P.rhs()))
    c_ref = sum(P.chi2(np.concatenate([x[c * ring.size + _group_index(P, ring)] for c in range(3)]))
                for _, _, P in probs)
    e_c = abs(ctx.chi2(x) - c_ref) / c_ref
    print(f"  A x {e_op:.2e}, b {e_b:.2e}, chi2 {e_c:.2e}", flush=True)
    assert e_op < 1e-11 and e_b < 1e-11 and e_c < 1e-10
    ctx.close()


def test_uneven_mixed_precision_solve_with_graphs(case):
    cfg, ds, probs, ring = case
    ctx = _ctx(cfg, ds, precision=sanepic_amd.MIXED)
    x = np.random.default_rng(4).standard_normal(3 * ring.size)
    e = rel(ctx.apply(x), _sum_groups(probs, ring, lambda P, xg: P.apply_A(xg), x))
    assert e < 2e-6
    runs = []
    for graph in (False, True):
        ctx.solve_begin(sanepic_amd.PRECOND_BLOCK3, 10, None, True, graph=graph)
        h = [ctx.step() for _ in range(25)]
        runs.append((np.array([v["crit"] for v in h]), ctx.map()))
    assert np.array_equal(runs[0][0], runs[1][0]) and np.array_equal(runs[0][1], runs[1][1])
    ctx.solve_begin(sanepic_amd.PRECOND_BLOCK3, 10, None, True)
    n = 0
    while n < 300:
        n += 1
        if ctx.step()["crit"] <= 1e-12:
            break
    res = ctx.true_residual()
    print(f"  mixed-precision block-Jacobi solve: {n} iterations, true residual {res:.2e}", flush=True)
    assert res < 1.5e-6
    ctx.close()


def test_uneven_rejects_uniform_uploads(case):
    cfg, ds, _, _ = case
    ctx = sanepic_amd.Context.uneven(cfg.nside, cfg.seglen, cfg.nsamp)
    with pytest.raises(sanepic_amd.SanepicError):
        ctx.set_tod(0, 0, np.zeros(8192))
    with pytest.raises(sanepic_amd.SanepicError):
        ctx.set_tod_range(0, 0, np.zeros(NSAMP + 1))
    with pytest.raises(sanepic_amd.SanepicError):  # pointing missing
        for k in range(ctx.ndet):
            ctx.set_detector(k, 0.0, 0.0, 0.0, ds["ispf"][k])
        ctx.pixelize()
    ctx.close()
    with pytest.raises(sanepic_amd.SanepicError):
        sanepic_amd.Context.uneven(32, [8192, 7 * 8192], NSAMP * 3)
    # the first half of detector 0 uploaded twice leaves its second half
    # unset, although as many samples arrived as the detector has
    ctx = sanepic_amd.Context.uneven(cfg.nside, cfg.seglen, cfg.nsamp)
    for k in range(ctx.ndet):
        ctx.set_detector(k, 0.0, 0.0, 0.0, ds["ispf"][k])
    ctx.set_pointing_range(0, ds["ra"], ds["dec"], ds["psi"])
    for k in range(ctx.ndet):
        t = ds["tod"][k]
        h = t.size // 2
        ctx.set_tod_range(k, 0, t[:h])
        ctx.set_tod_range(k, 0 if k == 0 else h, t[:h] if k == 0 else t[h:])
    ctx.pixelize()
    with pytest.raises(sanepic_amd.SanepicError, match="not set from sample"):
        ctx.setup()
    ctx.close()


@pytest.mark.parametrize("env", ["SANEPIC_NESTED_MAP", "SANEPIC_ITEM_ORDER"])
def test_uneven_layout_options_are_bit_identical(case, env, monkeypatch):
    """The groups' NESTED-order gather maps and P^T items (a group reads its
    pixels of the parent's x into its own map and adds its P^T into the
    parent's q) move where values are stored, not any sum: A x, b, chi2 and
    the pixels read back are bit-identical in RING order."""
    cfg, ds, probs, ring = case
    print("uneven layout option", env, flush=True)
    out = []
    for v in ("0", "1"):
        monkeypatch.setenv(env, v)
        for prec in (sanepic_amd.FP64, sanepic_amd.MIXED):
            ctx = _ctx(cfg, ds, precision=prec)
            x = np.random.default_rng(11).standard_normal(3 * ctx.npix)
            out.append((ctx.apply(x), ctx.rhs(), ctx.chi2(x) if prec == sanepic_amd.FP64 else 0.0,
                        ctx.pixels(1, 0)))
            ctx.close()
    for a, b in zip(out[:2], out[2:]):
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]) and a[2] == b[2]
        assert np.array_equal(a[3], b[3])


def test_union_pt_matches_per_group_pt(case, monkeypatch):
    """P^T once over the union of the groups' samples (the default) against
    one P^T per group adding into the union's q in turn: the same sums in
    another order, so A x and chi2 agree to rounding; b (built per group at
    set-up either way) is bit-identical; both match the oracle."""
    cfg, ds, probs, ring = case
    out = {}
    for v in ("1", "0"):
        monkeypatch.setenv("SANEPIC_UNION_PT", v)
        for prec in (sanepic_amd.FP64, sanepic_amd.MIXED):
            ctx = _ctx(cfg, ds, precision=prec)
            x = np.random.default_rng(12).standard_normal(3 * ctx.npix)
            out[(v, prec)] = (ctx.apply(x), ctx.rhs(), ctx.chi2(x) if prec == sanepic_amd.FP64 else 0.0)
            ctx.close()
    ref = _sum_groups(probs, ring, lambda P, xg: P.apply_A(xg), np.random.default_rng(12).standard_normal(3 * ring.size))
    for prec, tol in ((sanepic_amd.FP64, 1e-13), (sanepic_amd.MIXED, 1e-6)):
        u, g = out[("1", prec)], out[("0", prec)]
        e = rel(u[0], g[0])
        print(f"  precision {prec}: union vs per-group A x {e:.2e}", flush=True)
        assert e < tol and np.array_equal(u[1], g[1])
    u, g = out[("1", sanepic_amd.FP64)], out[("0", sanepic_amd.FP64)]
    assert abs(u[2] - g[2]) <= 1e-13 * abs(g[2])
    assert rel(u[0], ref) < 1e-11


def test_uneven_fp64_turns_match_pairs(case, monkeypatch):
    """The FP64 operator of per-detector lengths reading 2psi as 64-bit turns
    (every group's gather, radix 3 and 5 included, and the union's P^T over
    its pixel-sorted entries) against reading (cos, sin) pairs: A x and chi2
    agree to rounding, and both match the oracle's sum of groups."""
    cfg, ds, probs, ring = case
    got = {}
    for v in ("0", "1"):
        monkeypatch.setenv("SANEPIC_TURN64", v)
        ctx = _ctx(cfg, ds)
        x = np.random.default_rng(14).standard_normal(3 * ctx.npix)
        got[v] = (ctx.apply(x), ctx.chi2(x))
        assert ctx.fp64_angle_bytes() == (8 if v == "1" else 16)
        assert ctx.fp64_sorted_angle_bytes() == (8 if v == "1" else 16)
        ctx.close()
    ref = _sum_groups(probs, ring, lambda P, xg: P.apply_A(xg),
                      np.random.default_rng(14).standard_normal(3 * ring.size))
    d = rel(got["1"][0], got["0"][0])
    print(f"  turns vs pairs {d:.2e}, vs oracle {rel(got['1'][0], ref):.2e}", flush=True)
    assert d < 1e-14 and rel(got["1"][0], ref) < 1e-11 and rel(got["0"][0], ref) < 1e-11
    assert abs(got["1"][1] - got["0"][1]) <= 1e-12 * abs(got["0"][1])
