"""CPU oracle of the learned preconditioner (test infrastructure only).

Not part of the product: only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline may import this module.  The learned preconditioner has no
counterpart in the reference; these are FP64 restatements that check the HIP
kernels:

* ``forward(params, x)``: the U-Net in float64 torch on the CPU, with the
  HIP path's bf16 storage points emulated (weights, every stored activation,
  and every stored gradient are rounded to bf16; accumulation is float64
  here, float32 on the GPU).  ``torch.autograd`` gives the gradients the
  hand-written backward kernels must reproduce.
* ``features(...)``: the 10 input channels of a residual.
* ``block_jacobi(blocks, cnt)``: the guarded 3x3 inverse used as M.
* ``pool(img, c)`` / ``coarse_pos(pos, nside, c)``: a network coarser than
  the map by c sees the mean over c x c face pixels (unseen pixels zero) and
  its output at network position coarse_pos(pos) is added to each pixel.

The parameter layout is the library's: per convolution W[cout][9][cin]
(tap t = 3*dy + dx, offsets dy - 1 and dx - 1 on rows and columns) then
bias[cout].
"""
import numpy as np
import torch
import torch.nn.functional as F

# (cin, cout, level) of e0a e0b e1a e1b e2a e2b ba bb d2a d2b d1a d1b d0a d0b head
SPEC = [(16, 16, 0), (16, 16, 0), (16, 32, 1), (32, 32, 1), (32, 64, 2), (64, 64, 2),
        (64, 128, 3), (128, 128, 3), (192, 64, 2), (64, 64, 2), (96, 32, 1), (32, 32, 1),
        (48, 16, 0), (16, 16, 0), (16, 16, 0)]
NFEAT, NOUT, CIN0 = 10, 3, 16


def nparams():
    return sum(9 * ci * co + co for ci, co, _ in SPEC)


def unpack(p):
    """Flat parameter vector -> list of (W[cout][9][cin], bias[cout]) float64 tensors."""
    p = torch.as_tensor(np.asarray(p, dtype=np.float64))
    out, o = [], 0
    for ci, co, _ in SPEC:
        w = p[o:o + 9 * ci * co].reshape(co, 9, ci)
        o += 9 * ci * co
        b = p[o:o + co]
        o += co
        out.append((w, b))
    assert o == p.numel()
    return out


def bf16(t):
    """Round to the nearest bf16 (ties to even), kept in float64."""
    return t.to(torch.bfloat16).to(torch.float64)


class _BF(torch.autograd.Function):
    """bf16 storage: rounds the value going forward and the gradient going back."""

    @staticmethod
    def forward(ctx, x):
        return bf16(x)

    @staticmethod
    def backward(ctx, g):
        return bf16(g)


def _conv(x, w, b, relu, emulate):
    """3x3 'same' convolution of NCHW x with W[cout][9][cin] (zero padding)."""
    co, _, ci = w.shape
    wk = w.reshape(co, 3, 3, ci).permute(0, 3, 1, 2)  # [cout][cin][dy][dx]
    if emulate:
        wk = bf16(wk.detach()) + (wk - wk.detach())  # bf16 weights, gradient to the master copy
    y = F.conv2d(x, wk, b, padding=1)
    if relu:
        y = torch.relu(y)
        if emulate:
            y = _BF.apply(y)
    return y


def forward(params, x, emulate=True, requires_grad=False):
    """x[12][S][S][16] (feature channels then zeros) -> OUT[12][S][S][16] (float64).

    Returns (out, param_tensors); with requires_grad the parameter tensors
    are leaves whose .grad autograd fills."""
    layers = unpack(params)
    if requires_grad:
        layers = [(w.clone().requires_grad_(True), b.clone().requires_grad_(True)) for w, b in layers]
    B = lambda t: _BF.apply(t) if emulate else t  # noqa: E731
    X = torch.as_tensor(np.asarray(x, dtype=np.float64)).permute(0, 3, 1, 2)
    X = B(X)
    c = lambda i, t, relu=True: _conv(t, layers[i][0], layers[i][1], relu, emulate)  # noqa: E731
    pool = lambda t: B(F.avg_pool2d(t, 2))  # noqa: E731
    up = lambda t: F.interpolate(t, scale_factor=2, mode="nearest")  # noqa: E731
    y0 = c(0, X)
    y1 = c(1, y0)
    y2 = c(2, pool(y1))
    y3 = c(3, y2)
    y4 = c(4, pool(y3))
    y5 = c(5, y4)
    y6 = c(6, pool(y5))
    y7 = c(7, y6)
    y8 = c(8, B(torch.cat([up(y7), y5], 1)))
    y9 = c(9, y8)
    y10 = c(10, B(torch.cat([up(y9), y3], 1)))
    y11 = c(11, y10)
    y12 = c(12, B(torch.cat([up(y11), y1], 1)))
    y13 = c(13, y12)
    out = c(14, y13, relu=False)
    return out.permute(0, 2, 3, 1), layers


def grads(params, x, dy, emulate=True):
    """Gradient of sum(out[..., :3] * dy) w.r.t. the flat parameters."""
    out, layers = forward(params, x, emulate, requires_grad=True)
    dyt = torch.as_tensor(np.asarray(dy, dtype=np.float64))
    if emulate:
        dyt = bf16(dyt)
    (out[..., :NOUT] * dyt).sum().backward()
    g = []
    for w, b in layers:
        g.append(w.grad.reshape(-1).numpy())
        g.append(b.grad.numpy())
    return out.detach().numpy(), np.concatenate(g)


def conv_loop(x, w, b):
    """Direct loops for one 3x3 'same' convolution on NHWC x[F][S][S][cin]
    with W[cout][9][cin]: out[f,y,x,o] = b[o] + sum_t,c W[o,t,c] x[f,y+dy-1,x+dx-1,c]."""
    Fc, S, _, ci = x.shape
    co = w.shape[0]
    xp = np.zeros((Fc, S + 2, S + 2, ci))
    xp[:, 1:-1, 1:-1] = x
    out = np.broadcast_to(b, (Fc, S, S, co)).copy()
    for t in range(9):
        dy, dx = divmod(t, 3)
        out += np.einsum("fyxc,oc->fyxo", xp[:, dy:dy + S, dx:dx + S], w[:, t, :])
    return out


def block_jacobi(blocks, cnt):
    """M[npix][6] from blocks[npix][6] (b00 b01 b02 b11 b12 b22): inverse of the
    block, 1/b00 on I only where Q/U are masked (cnt < 2), and the diagonal
    inverse where the block is (near) singular, det <= 1e-10 b00 b11 b22."""
    b00, b01, b02, b11, b12, b22 = (blocks[:, k] for k in range(6))
    M = np.zeros_like(blocks)
    with np.errstate(divide="ignore", invalid="ignore"):
        c00 = b11 * b22 - b12 * b12
        c01 = b02 * b12 - b01 * b22
        c02 = b01 * b12 - b02 * b11
        c11 = b00 * b22 - b02 * b02
        c12 = b01 * b02 - b00 * b12
        c22 = b00 * b11 - b01 * b01
        det = b00 * c00 + b01 * c01 + b02 * c02
        good = (b00 > 0) & (b11 > 0) & (b22 > 0) & (det > 1e-10 * b00 * b11 * b22)
        inv = lambda v: np.where(v > 0, 1.0 / np.where(v > 0, v, 1.0), 0.0)  # noqa: E731
        full = np.stack([c00, c01, c02, c11, c12, c22], 1) / np.where(good, det, 1.0)[:, None]
        diag = np.stack([inv(b00), 0 * b00, 0 * b00, inv(b11), 0 * b00, inv(b22)], 1)
    M = np.where(good[:, None], full, diag)
    masked = cnt < 2
    M[masked] = 0.0
    M[masked, 0] = inv(b00[masked])
    return M


def apply_m(M, r3):
    """z = M r per pixel, r3[npix][3]."""
    m = M
    return np.stack([m[:, 0] * r3[:, 0] + m[:, 1] * r3[:, 1] + m[:, 2] * r3[:, 2],
                     m[:, 1] * r3[:, 0] + m[:, 3] * r3[:, 1] + m[:, 4] * r3[:, 2],
                     m[:, 2] * r3[:, 0] + m[:, 4] * r3[:, 1] + m[:, 5] * r3[:, 2]], 1)


def features(r_blk, blocks, cnt, pos, nside):
    """The network input of residual r (block layout [I|Q|U]): sigma and
    feat[12][S][S][10] = (M r / sigma, log(1 + cnt) / 4, log(b00 / mean b00),
    b01/b00, b02/b00, b11/b00, b12/b00, b22/b00) at each seen pixel's face
    position, zero elsewhere."""
    npix = blocks.shape[0]
    r3 = np.asarray(r_blk).reshape(3, npix).T
    M = block_jacobi(blocks, cnt)
    z = apply_m(M, r3)
    sigma = np.sqrt((z ** 2).sum() / (3 * npix))
    if not sigma > 0:
        sigma = 1.0
    b00 = blocks[:, 0]
    ref = b00.mean()
    f = np.zeros((npix, NFEAT))
    f[:, 0:3] = z / sigma
    f[:, 3] = 0.25 * np.log1p(cnt)
    with np.errstate(divide="ignore", invalid="ignore"):
        f[:, 4] = np.where(b00 > 0, np.log(b00 / ref), 0.0)
        for k in range(5):
            f[:, 5 + k] = np.where(b00 > 0, blocks[:, 1 + k] / np.where(b00 > 0, b00, 1.0), 0.0)
    out = np.zeros((12 * nside * nside, NFEAT))
    out[pos] = f
    return out.reshape(12, nside, nside, NFEAT), sigma


def pool(img, c):
    """Mean over c x c blocks of face images img[12][S][S][ch]."""
    f, S, _, ch = img.shape
    return img.reshape(f, S // c, c, S // c, c, ch).mean(axis=(2, 4))


def coarse_pos(pos, nside, c):
    """Network position (f, y // c, x // c) of face position f*S*S + y*S + x."""
    pos = np.asarray(pos, dtype=np.int64)
    f, y, x = pos // (nside * nside), (pos // nside) % nside, pos % nside
    Sn = nside // c
    return (f * Sn + y // c) * Sn + x // c
