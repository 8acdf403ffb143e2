"""GPU tests of the learned preconditioner (through the C ABI).

The network has no counterpart in the reference; its checker is the FP64
torch restatement in oracle/unet.py (itself pinned by test_oracle_unet.py).
A bf16 ReLU network is sensitive to rounding flips: the oracle moves by ~1%
(forward) and ~5% (weight gradients) under a 1e-7 relative change of its
parameters.  The GPU is therefore required to stay within a small multiple of
that intrinsic sensitivity, measured in the test itself; a wrong tap order,
transposition or pooling factor moves the result by O(1) (see the planted
mistake in test_oracle_unet.py).
"""
import numpy as np
import pytest
import torch

import sanepic_amd
from oracle import unet as ou
from tests.helpers import make_case, rel
from tests.test_oracle_unet import random_params

pytestmark = pytest.mark.gpu

S = 64


def _x16(x):
    xb = ou.bf16(torch.as_tensor(x.astype(np.float64))).numpy()
    x16 = np.zeros(x.shape[:3] + (16,))
    x16[..., :10] = xb
    return x16


def _layers(g):
    o = 0
    for ci, co, _ in ou.SPEC:
        nw = 9 * ci * co
        yield g[o:o + nw], g[o + nw:o + nw + co]
        o += nw + co


def test_unet_forward_backward_match_oracle():
    print("unet nside", S, flush=True)
    rng = np.random.default_rng(0)
    p = random_params(rng).astype(np.float32)
    x = rng.standard_normal((12, S, S, 10)).astype(np.float32)
    dy = rng.standard_normal((12, S, S, 3)).astype(np.float32)
    u = sanepic_amd.UNet(S)
    u.set_params(p)
    assert np.array_equal(u.params(), p)
    y = u.forward(x)
    y2, g = u.backward(x, dy)
    assert np.array_equal(y, y2)
    p64 = p.astype(np.float64)
    x16 = _x16(x)
    yo, go = ou.grads(p64, x16, dy.astype(np.float64))
    pp = p64 * (1 + 1e-7 * rng.standard_normal(p64.size))
    yp, gp = ou.grads(pp, x16, dy.astype(np.float64))
    ef, ef0 = rel(y, yo[..., :3]), rel(yp[..., :3], yo[..., :3])
    print(f"  forward: GPU {ef:.2e}, oracle sensitivity {ef0:.2e}", flush=True)
    assert ef < 3 * ef0 + 1e-4 and ef < 3e-2
    for i, ((gw, gb), (ow, ob), (pw, pb)) in enumerate(zip(_layers(g), _layers(go), _layers(gp))):
        ew, ew0 = rel(gw, ow), rel(pw, ow)
        eb, eb0 = rel(gb, ob), rel(pb, ob)
        print(f"  conv {i:2d}: dW {ew:.2e} (sensitivity {ew0:.2e})  db {eb:.2e} ({eb0:.2e})", flush=True)
        assert ew < 3 * ew0 + 1e-3 and ew < 0.3
        assert eb < 3 * eb0 + 1e-3 and eb < 0.3


def test_unet_rejects_bad_nside():
    with pytest.raises(sanepic_amd.SanepicError):
        sanepic_amd.UNet(48)
    with pytest.raises(sanepic_amd.SanepicError):
        sanepic_amd.UNet(32)


@pytest.fixture(scope="module")
def problem():
    print("learned-preconditioner case nside 64, 4 det x 8 x 2^16", flush=True)
    cfg, ds = make_case(nside=S, ndet=4, nseg=8, ns=2 ** 16, fknee=0.2)
    ctx = sanepic_amd.Context(cfg.nside, cfg.ndet, cfg.nseg, cfg.ns)
    ctx.load(cfg, ds)
    npix, _ = ctx.pixelize()
    ctx.setup()
    un = sanepic_amd.UNet(S, seed=1)
    ctx.learn_attach(un)
    yield ctx, un
    ctx.learn_attach(None)
    ctx.close()
    un.close()


def _solve(ctx, precond, maxit=300):
    ctx.solve_begin(precond, 10, None, True)
    h = []
    while len(h) < maxit:
        h.append(ctx.step())
        if h[-1]["crit"] <= 1e-12:
            break
    return h


def _mask(v, cnt, npix):
    v = v.copy()
    m = cnt < 2
    v[npix:2 * npix][m] = 0.0
    v[2 * npix:][m] = 0.0
    return v


def test_face_index_and_features(problem):
    ctx, un = problem
    npix = ctx.npix
    ind = ctx.indpix()
    ring = np.flatnonzero(ind >= 0)
    ring = ring[np.argsort(ind[ring])]
    pos = ctx.face_index()
    assert np.array_equal(pos, sanepic_amd.ring_to_face(S, ring))
    assert np.unique(pos).size == npix
    hits, _, _ = ctx.get_setup()
    cnt = hits[:npix] // 2
    r = np.random.default_rng(4).standard_normal(3 * npix)
    feat, sig = ctx.features(r)
    fo, so = ou.features(r, ctx.blocks(), cnt, pos, S)
    fo = ou.bf16(torch.as_tensor(fo)).numpy()
    assert abs(sig - so) <= 1e-12 * so
    assert np.allclose(feat, fo, rtol=2.0 ** -7, atol=1e-30)


def test_untrained_network_is_block_jacobi(problem):
    ctx, un = problem
    hb = _solve(ctx, sanepic_amd.PRECOND_BLOCK3)
    hl = _solve(ctx, sanepic_amd.PRECOND_LEARNED)
    st = ctx.learn_stats()
    print(f"  block-Jacobi {len(hb)} its, untrained learned {len(hl)} its, {st}", flush=True)
    assert st["fallback"] == 0 and st["rejected"] == 0
    for k in range(min(20, len(hb), len(hl))):
        assert abs(hl[k]["crit"] - hb[k]["crit"]) <= 1e-6 * hb[k]["crit"]
    assert abs(len(hl) - len(hb)) <= 1


@pytest.mark.parametrize("kind", ["nan", "garbage"])
def test_bad_network_falls_back_to_block_jacobi(problem, kind):
    """A network whose parameters are NaN, or random weights far from any
    trained scale, cannot derail the flexible PCG: NaN or non-descent
    directions are replaced by block-Jacobi (rejected), a kept network step
    that lets |r| grow sends the next application to block-Jacobi
    (fallback), and the solve reaches the same map as block-Jacobi.  Both the
    default growth factor and the literal rule (fallback whenever a network
    step fails to reduce |r|, growth 1) are run."""
    ctx, un = problem
    keep = un.params()
    hb = _solve(ctx, sanepic_amd.PRECOND_BLOCK3)
    xb = ctx.map()
    ab = xb @ ctx.apply(xb)
    rng = np.random.default_rng(12)
    bad = (np.full(keep.size, np.nan, np.float32) if kind == "nan"
           else (3.0 * rng.standard_normal(keep.size)).astype(np.float32))
    try:
        un.set_params(bad)
        for growth in (0.0, 1.0):
            ctx.solve_begin(sanepic_amd.PRECOND_LEARNED, 10, None, True, fallback_growth=growth)
            h = []
            while len(h) < 600:
                h.append(ctx.step())
                if h[-1]["crit"] <= 1e-12:
                    break
            st = ctx.learn_stats()
            res = ctx.true_residual()
            d = ctx.map() - xb
            dA = (d @ ctx.apply(d)) / ab
            print(f"  {kind} network, growth {growth or 2}: {len(h)} its (block-Jacobi {len(hb)}), {st}, "
                  f"true residual {res:.2e}, A-norm difference {dA:.1e}", flush=True)
            assert st["rejected"] + st["fallback"] > 0
            assert h[-1]["crit"] <= 1e-12 and res <= 1.5e-6
            assert dA <= 1e-6
            if kind == "nan":  # every network direction (one per step and the
                # first) is replaced: block-Jacobi's iterates
                assert st["rejected"] == len(h) + 1 and st["learned"] == 0
                assert abs(len(h) - len(hb)) <= 1
    finally:
        un.set_params(keep)
    assert np.array_equal(un.params(), keep)


def test_harvest_pairs_are_exact(problem):
    ctx, un = problem
    npix = ctx.npix
    hits, _, _ = ctx.get_setup()
    cnt = hits[:npix] // 2
    added, iters = ctx.harvest(10000, capacity=3000)
    print(f"  harvest: {added} pairs from a {iters}-iteration solve", flush=True)
    assert 10 < added <= iters
    # e_k = x* - x_k with x* solved to |r*|^2/|b|^2 <= 1e-17, so A e_k = r_k
    # holds up to |r*| / |r_k| <= sqrt(1e-17 / 1e-12) for the last kept pair
    # (|r_k|^2/|b|^2 >= 1e-12), plus the FP32 storage of the pair
    for k, tol in ((0, 1e-6), (1, 1e-6), (added - 1, 4e-3)):
        r, e = ctx.pair(k)
        err = rel(_mask(ctx.apply(e), cnt, npix), r)
        print(f"  pair {k}: |A e - r| / |r| = {err:.2e}", flush=True)
        assert err < tol


def test_loss_matches_numpy(problem):
    ctx, un = problem
    npix = ctx.npix
    hits, _, _ = ctx.get_setup()
    cnt = hits[:npix] // 2
    pos = ctx.face_index()
    r, e = ctx.pair(1)
    feat, sig = ctx.features(r)
    net = un.forward(feat).reshape(-1, 3)[pos]
    M = ou.block_jacobi(ctx.blocks(), cnt)
    z3 = ou.apply_m(M, r.reshape(3, npix).T) + sig * net
    z3[cnt < 2, 1:] = 0.0
    z = z3.T.reshape(-1)
    Az = _mask(ctx.apply(z), cnt, npix)
    ref = (z @ Az - 2 * z @ r + e @ r) / (e @ r)
    got = ctx.eval_pair(1, sanepic_amd.LOSS_ANORM)
    assert abs(got - ref) <= 1e-6 * abs(ref)
    ref_r = np.sum((Az - r) ** 2) / np.sum(r ** 2)
    got_r = ctx.eval_pair(1, sanepic_amd.LOSS_RESIDUAL)
    assert abs(got_r - ref_r) <= 1e-6 * abs(ref_r)


def test_training_gain_over_block_jacobi(problem):
    """Ablation: the same flexible PCG with and without the trained network.

    Training pairs come from block-Jacobi solves of synthetic right-hand sides
    (seeds 10001-10039, disjoint from the data); the test solves the data's
    own b."""
    ctx, un = problem
    for sd in range(10001, 10040):
        ctx.harvest(sd, capacity=3000)
    n = ctx.npairs()
    ntrain = int(0.9 * n)
    R = sanepic_amd.LOSS_RESIDUAL
    base = np.mean([ctx.eval_pair(k, R) for k in range(ntrain, n)])
    loss = ctx.train(6000, lr=1e-3, lr_final=1e-5, seed=0, first=0, count=ntrain, loss=R)
    held = np.mean([ctx.eval_pair(k, R) for k in range(ntrain, n)])
    print(f"  {n} pairs; held-out |A z - r|^2/|r|^2 of one application: block-Jacobi {base:.3f}, "
          f"trained {held:.3f}; training loss {loss[:100].mean():.3f} -> {loss[-100:].mean():.3f}",
          flush=True)
    assert held < 0.8 * base
    hb = _solve(ctx, sanepic_amd.PRECOND_BLOCK3)
    xb = ctx.map()
    hl = _solve(ctx, sanepic_amd.PRECOND_LEARNED)
    xl = ctx.map()
    res = ctx.true_residual()
    st = ctx.learn_stats()
    print(f"  iterations to |r|/|b| <= 1e-6: block-Jacobi {len(hb)}, learned {len(hl)} {st}, "
          f"true residual {res:.2e}", flush=True)
    assert res <= 1.5e-6
    assert len(hl) <= 0.8 * len(hb)
    # both maps solve the same system: their difference is small in the A-norm
    d = xl - xb
    dA = (d @ ctx.apply(d)) / (xb @ ctx.apply(xb))
    print(f"  relative A-norm difference of the two maps {dA:.2e}", flush=True)
    assert dA <= 1e-6
    # on-policy pairs: iterates of the trained network's own flexible PCG,
    # with the same exactness A e_k = r_k as block-Jacobi pairs
    ctx.clear_pairs()
    added, iters = ctx.harvest(20000, capacity=3000, policy=sanepic_amd.HARVEST_LEARNED)
    print(f"  on-policy harvest: {added} pairs from a {iters}-iteration solve", flush=True)
    assert 5 < added <= iters
    npix = ctx.npix
    cnt = ctx.get_setup()[0][:npix] // 2
    r0 = np.linalg.norm(ctx.pair(0)[0])  # ~ |b'|: the solve starts from x = 0
    for k in (0, 1, added // 2):
        r, e = ctx.pair(k)
        err = rel(_mask(ctx.apply(e), cnt, npix), r)
        # truncation bound |r*| / |r_k| (solved to |r*|^2/|b|^2 <= 1e-17)
        bound = np.sqrt(1e-17) * r0 / np.linalg.norm(r)
        print(f"  on-policy pair {k}: |A e - r| / |r| = {err:.2e} (bound {bound:.2e})", flush=True)
        assert err < 3 * bound + 1e-6
    ctx.truncate_pairs(1)
    assert ctx.npairs() == 1
    with pytest.raises(sanepic_amd.SanepicError):
        ctx.truncate_pairs(2)
    with pytest.raises(sanepic_amd.SanepicError):
        ctx.harvest(20001, capacity=3000, policy=7)
