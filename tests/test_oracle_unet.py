"""Pins the learned preconditioner's CPU oracle (oracle/unet.py) and the
host-side face indexing, without a GPU.

* The oracle's convolution (torch conv2d with the library's W[cout][9][cin]
  layout) equals direct loops over taps: pins the tap order, the row/column
  offsets and the zero padding.
* The oracle's autograd gradient equals central finite differences of the
  un-rounded network: pins the wiring of pooling, upsampling and the skip
  concatenations.
* The nested face position of every RING pixel agrees with the nested ->
  ring renumbering of Gorski et al. (2005) written independently in
  test_oracle_healpix.py.
* The oracle's block-Jacobi inverse times the block is the identity.
"""
import numpy as np
import pytest
import torch

import sanepic_amd
from oracle import unet as ou
from tests.helpers import rel
from tests.test_oracle_healpix import xyf2ring


def random_params(rng, scale=1.0):
    p = []
    for ci, co, _ in ou.SPEC:
        p.append(rng.standard_normal(9 * ci * co) * scale * np.sqrt(2.0 / (9 * ci)))
        p.append(rng.standard_normal(co) * 0.1)
    return np.concatenate(p)


def test_conv_matches_loops():
    rng = np.random.default_rng(1)
    x = rng.standard_normal((12, 8, 8, 16))
    w = rng.standard_normal((32, 9, 16))
    b = rng.standard_normal(32)
    X = torch.as_tensor(x).permute(0, 3, 1, 2)
    got = ou._conv(X, torch.as_tensor(w), torch.as_tensor(b), relu=False, emulate=False)
    ref = ou.conv_loop(x, w, b)
    assert np.abs(got.permute(0, 2, 3, 1).numpy() - ref).max() < 1e-12


def test_parameter_count():
    assert ou.nparams() == sum(9 * ci * co + co for ci, co, _ in ou.SPEC)


def test_autograd_matches_finite_differences():
    rng = np.random.default_rng(2)
    S = 8
    p = random_params(rng)
    x = np.zeros((12, S, S, 16))
    x[..., :10] = rng.standard_normal((12, S, S, 10))
    dy = rng.standard_normal((12, S, S, 3))
    _, g = ou.grads(p, x, dy, emulate=False)

    def loss(q):
        out, _ = ou.forward(q, x, emulate=False)
        return float((out[..., :3].numpy() * dy).sum())

    idx = rng.choice(p.size, 40, replace=False)
    h = 1e-6
    bad = 0
    for i in idx:
        q1, q2 = p.copy(), p.copy()
        q1[i] += h
        q2[i] -= h
        fd = (loss(q1) - loss(q2)) / (2 * h)
        if abs(fd - g[i]) > 1e-5 * (1 + abs(g[i])):
            bad += 1
    # a finite difference straddling a ReLU kink may miss; the wiring may not
    assert bad <= 1, f"{bad} of {idx.size} parameter gradients differ from finite differences"


def test_bf16_rounding_is_nearest_even():
    v = torch.tensor([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, 1.0 + 2 ** -9], dtype=torch.float64)
    r = ou.bf16(v).numpy()
    assert r[0] == 1.0
    assert r[1] == 1.0  # tie -> even
    assert r[2] == 1.0 + 2 ** -6  # tie -> even (up)
    assert r[3] == 1.0


@pytest.mark.parametrize("nside", [64, 128])
def test_ring_to_face_is_nested_geometry(nside):
    S = nside
    faces = np.repeat(np.arange(12), S * S)
    iy = np.tile(np.repeat(np.arange(S), S), 12)
    ix = np.tile(np.arange(S), 12 * S)
    ring = np.array([xyf2ring(S, int(a), int(b), int(f)) for a, b, f in zip(ix, iy, faces)])
    assert np.array_equal(np.sort(ring), np.arange(12 * S * S))
    pos = sanepic_amd.ring_to_face(S, ring)
    assert np.array_equal(pos, (faces * S + iy) * S + ix)


def test_ring_to_face_rejects_bad_input():
    with pytest.raises(sanepic_amd.SanepicError):
        sanepic_amd.ring_to_face(48, np.arange(4))
    with pytest.raises(sanepic_amd.SanepicError):
        sanepic_amd.ring_to_face(64, np.array([12 * 64 * 64]))


def test_block_jacobi_inverse():
    rng = np.random.default_rng(3)
    n = 50
    A = rng.standard_normal((n, 3, 3))
    B = A @ A.transpose(0, 2, 1) + 0.1 * np.eye(3)
    blocks = np.stack([B[:, 0, 0], B[:, 0, 1], B[:, 0, 2], B[:, 1, 1], B[:, 1, 2], B[:, 2, 2]], 1)
    M = ou.block_jacobi(blocks, np.full(n, 10))
    Mm = np.stack([M[:, [0, 1, 2]], M[:, [1, 3, 4]], M[:, [2, 4, 5]]], 1)
    assert np.abs(Mm @ B - np.eye(3)).max() < 1e-9
    # masked pixels: I only
    M2 = ou.block_jacobi(blocks, np.full(n, 1))
    assert np.allclose(M2[:, 0], 1 / blocks[:, 0]) and np.all(M2[:, 1:] == 0)


def test_planted_tap_flip_is_detected():
    """A mistake of the kind the GPU comparison must catch (one layer's taps
    reversed, i.e. convolution instead of correlation) moves the output by
    far more than the bf16 network's intrinsic rounding sensitivity."""
    rng = np.random.default_rng(5)
    S = 16
    p = random_params(rng)
    x = np.zeros((12, S, S, 16))
    x[..., :10] = ou.bf16(torch.as_tensor(rng.standard_normal((12, S, S, 10)))).numpy()
    y, _ = ou.forward(p, x)
    y2, _ = ou.forward(p * (1 + 1e-7 * rng.standard_normal(p.size)), x)
    sens = rel(y2.numpy()[..., :3], y.numpy()[..., :3])
    q = p.copy()
    o = sum(9 * ci * co + co for ci, co, _ in ou.SPEC[:5])
    ci, co, _ = ou.SPEC[5]
    q[o:o + 9 * ci * co] = p[o:o + 9 * ci * co].reshape(co, 9, ci)[:, ::-1, :].reshape(-1)
    y3, _ = ou.forward(q, x)
    assert rel(y3.numpy()[..., :3], y.numpy()[..., :3]) > 20 * sens


def test_pool_and_coarse_pos_against_loops():
    """Pooling is the block mean written out with loops; nested pooling
    composes; coarse_pos agrees with the block a face pixel falls in."""
    rng = np.random.default_rng(5)
    S, c = 8, 4
    img = rng.standard_normal((12, S, S, 3))
    ref = np.zeros((12, S // c, S // c, 3))
    for f in range(12):
        for Y in range(S // c):
            for X in range(S // c):
                ref[f, Y, X] = img[f, Y * c:(Y + 1) * c, X * c:(X + 1) * c].sum(axis=(0, 1)) / (c * c)
    assert np.allclose(ou.pool(img, c), ref, rtol=1e-14, atol=0)
    assert np.allclose(ou.pool(ou.pool(img, 2), 2), ou.pool(img, 4), rtol=1e-13, atol=1e-15)
    pos = np.arange(12 * S * S)
    cp = ou.coarse_pos(pos, S, c)
    f, y, x = pos // (S * S), (pos // S) % S, pos % S
    assert np.array_equal(cp, (f * (S // c) + y // c) * (S // c) + x // c)
    # upsampling by coarse_pos is the transpose of the block sum (= c^2 pool)
    v = rng.standard_normal(12 * (S // c) ** 2)
    up = v[cp]
    lhs = up @ img[..., 0].reshape(-1)
    rhs = v @ (c * c * ou.pool(img[..., :1], c)).reshape(-1)
    assert abs(lhs - rhs) <= 1e-12 * abs(rhs)
