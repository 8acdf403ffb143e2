"""A network coarser than the map (through the C ABI).

At map nside 128 a 64-nside network sees the residual's features averaged
over 2 x 2 face pixels and its output is added to every pixel of the block
(a learned coarse-scale correction to block-Jacobi); at map nside 512 the
same network pools 8 x 8 face pixels, the factor of the 16-detector
workload's network.  Checkers: the oracle's
features pooled with oracle.unet.pool (itself pinned against loops in
test_oracle_unet.py); the residual loss evaluated in numpy from the
network's own forward output, upsampled with oracle.unet.coarse_pos; and the
loss gradient through the pooling, which must equal the network's backward
applied to the block sums of the per-pixel output gradient computed in
numpy (a missing or transposed pooling moves it by O(1))."""
import numpy as np
import pytest
import torch

import sanepic_amd
from oracle import unet as ou
from tests.helpers import make_case, rel

pytestmark = pytest.mark.gpu

SN = 64


@pytest.fixture(scope="module", params=[128, 512], ids=lambda s: f"map{s}_net64")
def problem(request):
    S = request.param
    print(f"coarse network: map nside {S}, network nside {SN}, 4 det x 4 x 2^16", flush=True)
    cfg, ds = make_case(nside=S, ndet=4, nseg=4, ns=2 ** 16, fknee=0.2)
    ctx = sanepic_amd.Context(cfg.nside, cfg.ndet, cfg.nseg, cfg.ns)
    ctx.load(cfg, ds)
    ctx.pixelize()
    ctx.setup()
    un = sanepic_amd.UNet(SN, seed=1)
    ctx.learn_attach(un)
    yield ctx, un, S
    ctx.learn_attach(None)
    ctx.close()
    un.close()


def _cnt(ctx):
    return ctx.get_setup()[0][:ctx.npix] // 2


def _mask(v, cnt, npix):
    v = v.copy()
    v[npix:2 * npix][cnt < 2] = 0.0
    v[2 * npix:][cnt < 2] = 0.0
    return v


def test_network_finer_than_map_is_rejected(problem):
    ctx, _, S = problem
    fine = sanepic_amd.UNet(2 * S)
    with pytest.raises(sanepic_amd.SanepicError):
        ctx.learn_attach(fine)
    fine.close()


def test_pooled_features(problem):
    ctx, un, S = problem
    npix = ctx.npix
    pos = ctx.face_index()
    r = np.random.default_rng(4).standard_normal(3 * npix)
    feat, sig = ctx.features(r)
    assert feat.shape == (12, SN, SN, 10)
    fo, so = ou.features(r, ctx.blocks(), _cnt(ctx), pos, S)
    ref = ou.bf16(torch.as_tensor(ou.pool(fo, S // SN))).numpy()
    assert abs(sig - so) <= 1e-12 * so
    # GPU: per-pixel features in FP32, pooled in FP32, rounded to bf16 once
    assert np.allclose(feat, ref, rtol=2.0 ** -7, atol=1e-6)


def _trained_a_little(ctx):
    if ctx.npairs() == 0:
        for sd in range(10001, 10004):
            ctx.harvest(sd, capacity=400)
        ctx.train(200, lr=1e-3, lr_final=1e-4, seed=0, loss=sanepic_amd.LOSS_RESIDUAL)


def test_residual_loss_and_gradient_through_pooling(problem):
    ctx, un, S = problem
    _trained_a_little(ctx)  # a nonzero head, so the output matters
    npix = ctx.npix
    cnt = _cnt(ctx)
    cpos = ou.coarse_pos(ctx.face_index(), S, S // SN)
    r, e = ctx.pair(1)
    feat, sig = ctx.features(r)
    y = un.forward(feat).reshape(-1, 3)
    M = ou.block_jacobi(ctx.blocks(), cnt)
    z3 = ou.apply_m(M, r.reshape(3, npix).T) + sig * y[cpos]
    z3[cnt < 2, 1:] = 0.0
    z = z3.T.reshape(-1)
    u = _mask(ctx.apply(z), cnt, npix) - r
    rr = r @ r
    ref = (u @ u) / rr
    loss, g = ctx.grad_pair(1, sanepic_amd.LOSS_RESIDUAL)
    print(f"  residual loss GPU {loss:.6e}, numpy {ref:.6e}", flush=True)
    assert abs(loss - ref) <= 1e-6 * ref
    # dL/d(output at pixel p) = 2 sigma (A u)_p / |r|^2, summed over each block
    v = _mask(ctx.apply(u), cnt, npix)
    gp = (2.0 * sig / rr) * v.reshape(3, npix).T
    gp[cnt < 2, 1:] = 0.0
    dy = np.zeros((12 * SN * SN, 3))
    np.add.at(dy, cpos, gp)
    _, gref = un.backward(feat, dy.reshape(12, SN, SN, 3).astype(np.float32))
    e_g = rel(g, gref)
    print(f"  parameter gradient through the pooling: rel. difference {e_g:.2e}", flush=True)
    assert e_g < 2e-2
    # a wrong transpose (mean instead of sum over the block) is far off
    _, gbad = un.backward(feat, (dy / (S // SN) ** 2).reshape(12, SN, SN, 3).astype(np.float32))
    assert rel(gbad, gref) > 0.5


def test_coarse_learned_solve_converges(problem):
    ctx, un, S = problem
    _trained_a_little(ctx)
    ctx.solve_begin(sanepic_amd.PRECOND_BLOCK3, 10, None, True)
    nb = 0
    while nb < 300:
        nb += 1
        if ctx.step()["crit"] <= 1e-12:
            break
    xb = ctx.map()
    ctx.solve_begin(sanepic_amd.PRECOND_LEARNED, 10, None, True)
    nl = 0
    while nl < 300:
        nl += 1
        if ctx.step()["crit"] <= 1e-12:
            break
    res = ctx.true_residual()
    print(f"  block-Jacobi {nb} its, coarse learned {nl} its {ctx.learn_stats()}, true residual {res:.2e}",
          flush=True)
    assert res <= 1.5e-6
    d = ctx.map() - xb
    assert (d @ ctx.apply(d)) / (xb @ ctx.apply(xb)) <= 1e-6
