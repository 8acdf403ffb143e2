"""GPU path (C ABI) against the CPU oracle on the same seeded inputs."""
import numpy as np
import pytest

import sanepic_amd
from tests.helpers import make_case, oracle_problem, rel

pytestmark = pytest.mark.gpu

CASES = [
    # nside, ndet, nseg, ns  -- single-pass FFT sizes and four-step sizes
    dict(nside=8, ndet=2, nseg=3, ns=1024),
    dict(nside=16, ndet=1, nseg=4, ns=256),
    dict(nside=16, ndet=2, nseg=2, ns=2 ** 13),
    dict(nside=32, ndet=1, nseg=3, ns=2 ** 16),
    dict(nside=32, ndet=2, nseg=2, ns=2 ** 17),
]


def _gpu(cfg, ds, precision=sanepic_amd.FP64):
    ctx = sanepic_amd.Context(cfg.nside, cfg.ndet, cfg.nseg, cfg.ns, precision=precision)
    ctx.load(cfg, ds)
    npix, ndef = ctx.pixelize()
    ctx.setup()
    return ctx, npix, ndef


@pytest.mark.parametrize("case", CASES, ids=lambda c: "n{nside}_d{ndet}_s{nseg}_l{ns}".format(**c))
def test_pixels_hits_operator(case):
    print("case", case, flush=True)
    cfg, ds = make_case(**case)
    P = oracle_problem(cfg, ds)
    ctx, npix, ndef = _gpu(cfg, ds)
    # pixel indices and the seen-pixel set: bit-exact
    dp, _, _, ind = P.pixels()
    assert npix == P.npix
    assert np.array_equal(ctx.all_pixels(), dp)
    assert np.array_equal(ctx.indpix(), ind)
    # hits (doubled), crosslinks, diagonal preconditioner
    ax, hits, mp, sc = P.setup_pass()
    ghits, gmp, gsc = ctx.get_setup()
    assert np.array_equal(ghits, hits)
    assert rel(gsc, sc) < 1e-12
    assert rel(gmp, mp) < 1e-12
    # b = P^T N^-1 d and A x on a random map
    assert rel(ctx.rhs(), P.rhs()) < 1e-11
    rng = np.random.default_rng(5)
    x = rng.standard_normal(3 * npix)
    assert rel(ctx.apply(x), P.apply_A(x)) < 1e-11
    # chi2 of a map
    c_or = P.chi2(x)
    assert abs(ctx.chi2(x) - c_or) / c_or < 1e-10


def _masked(v, hits, npix):
    v = v.copy()
    m = hits[npix:] < 4
    v[npix:2 * npix][m] = 0.0
    v[2 * npix:][m] = 0.0
    return v


def _dense_solution(P, b, hits, npix):
    """x* of the masked system from the oracle's dense A (tiny cases only)."""
    n3 = 3 * npix
    keep = np.ones(n3, bool)
    keep[npix:2 * npix][hits[npix:] < 4] = False
    keep[2 * npix:][hits[npix:] < 4] = False
    idx = np.flatnonzero(keep)
    A = np.zeros((idx.size, idx.size))
    for j, col in enumerate(idx):
        e = np.zeros(n3)
        e[col] = 1.0
        A[:, j] = P.apply_A(e)[idx]
    A = 0.5 * (A + A.T)
    w, V = np.linalg.eigh(A)
    x = np.zeros(n3)
    x[idx] = V @ ((V.T @ b[idx]) / w)
    return x, A, idx


@pytest.mark.parametrize("case", CASES[:3], ids=lambda c: "n{nside}_d{ndet}_s{nseg}_l{ns}".format(**c))
def test_solve_matches_oracle(case):
    """The GPU PCG follows the reference's PCG.

    The reference system is ill-conditioned by construction (the DC bin gets
    1e-6 of the first bin's weight, sanepic.cc:537, so segment offsets and the
    map monopole are nearly free), so two correct FP64 solvers that differ
    only in summation order drift apart after a few iterations: the oracle
    itself run with 1 and 4 threads is the yardstick.  Checks: (1) early
    iterates agree within that spread, (2) after a fixed number of iterations
    the GPU's A-norm error to the exact solution is as small as the
    oracle's, (3) the masked outputs NVAR and crosslinks match."""
    print("case", case, flush=True)
    cfg, ds = make_case(**case)
    P1 = oracle_problem(cfg, ds, nthreads=1)
    P4 = oracle_problem(cfg, ds, nthreads=4)
    ctx, npix, _ = _gpu(cfg, ds)
    K = 8
    r1 = P1.solve(stop=0.0, maxiter=K)
    r4 = P4.solve(stop=0.0, maxiter=K)
    SK, hK = ctx.solve(stop=0.0, maxiter=K)
    spread = rel(r4["S"], r1["S"])
    print(f"  after {K} its: oracle 1-vs-4-thread spread {spread:.2e}, gpu-oracle {rel(SK, r1['S']):.2e}")
    assert rel(SK, r1["S"]) < max(100 * spread, 1e-9)
    crit = np.array([h["crit"] for h in hK])
    assert np.allclose(crit, r1["crit"], rtol=max(100 * spread, 1e-8), atol=0)
    chi2 = np.array([h["chi2"] for h in hK])
    assert np.allclose(chi2[:1], r1["chi2"][:1], rtol=1e-10, atol=0)
    # convergence in the A-norm against the exact solution
    b, hits = r1["b"], r1["hits"]
    xs, A, idx = _dense_solution(P1, b, hits, npix)
    it = 150
    ro = P1.solve(stop=0.0, maxiter=it)
    Sg, _ = ctx.solve(stop=0.0, maxiter=it)

    def anorm(e):
        return np.sqrt(abs(e[idx] @ A @ e[idx]))

    eo, eg = anorm(ro["S"] - xs) / anorm(xs), anorm(Sg - xs) / anorm(xs)
    print(f"  after {it} its: A-norm error oracle {eo:.2e} gpu {eg:.2e}")
    assert eg <= 3.0 * eo + 1e-10
    assert rel(ctx.nvar(), ro["Mptot"]) < 1e-12
    assert rel(ctx.sumcross(), ro["sumCross"]) < 1e-12


@pytest.mark.parametrize("case", [CASES[0], CASES[3]], ids=lambda c: "n{nside}_d{ndet}_s{nseg}_l{ns}".format(**c))
def test_mixed_precision_operator(case):
    """FP32 operator (TOD, FFT, scatter) against the FP64 oracle: relative error
    of A x at the FP32 level (the FFT's error grows with log2 N)."""
    print("case", case, flush=True)
    cfg, ds = make_case(**case)
    P = oracle_problem(cfg, ds)
    ctx, npix, _ = _gpu(cfg, ds, precision=sanepic_amd.MIXED)
    x = np.random.default_rng(5).standard_normal(3 * npix)
    err = rel(ctx.apply(x), P.apply_A(x))
    print(f"  FP32 operator relative error {err:.2e}")
    assert 1e-9 < err < 2e-6
    assert rel(ctx.rhs(), P.rhs()) < 1e-11          # b is built in FP64
    c_or = P.chi2(x)
    assert abs(ctx.chi2(x) - c_or) / c_or < 1e-10    # chi2 uses the FP64 operator


def test_mixed_precision_solve_reaches_fp64_residual():
    """FP32 operator inside PCG with the FP64 residual replacement every 10
    iterations: the true (FP64) residual still reaches 1e-6, and the map
    agrees with the FP64 solve on the well-determined modes."""
    cfg, ds = make_case(nside=16, ndet=4, nseg=4, ns=2 ** 14, precessions=8.0)
    c64, npix, _ = _gpu(cfg, ds)
    c32, _, _ = _gpu(cfg, ds, precision=sanepic_amd.MIXED)
    S64, h64 = c64.solve(precond=sanepic_amd.PRECOND_BLOCK3, deflate=True, stop=1e-13, maxiter=600)
    S32, h32 = c32.solve(precond=sanepic_amd.PRECOND_BLOCK3, deflate=True, stop=1e-13, maxiter=600)
    t32 = c32.true_residual()
    print(f"  fp64 {len(h64)} its, mixed {len(h32)} its, mixed true residual {t32:.2e}")
    assert t32 <= 1e-6
    e = S32 - S64
    a_e = np.sqrt(abs(e @ c64.apply(e)))
    a_s = np.sqrt(S64 @ c64.apply(S64))
    print(f"  A-norm map difference {a_e / a_s:.2e}")
    assert a_e / a_s < 1e-5


@pytest.mark.parametrize("case", [CASES[0], CASES[3]], ids=lambda c: "n{nside}_d{ndet}_s{nseg}_l{ns}".format(**c))
def test_exact_blocks_are_the_operator_diagonal(case):
    """The exact 3x3 diagonal blocks of A (block-Jacobi's input, built by the
    pair-sum kernel from the noise autocorrelation over all lags) equal the
    diagonal entries of the operator itself, A e_(p,c) at pixel p, for a few
    well-observed pixels; the operator is checked against the oracle above."""
    print("case", case, flush=True)
    cfg, ds = make_case(**case)
    ctx, npix, _ = _gpu(cfg, ds)
    hits = ctx.get_setup()[0]
    blocks = ctx.blocks()
    cand = np.flatnonzero(hits[npix:] >= 8)
    pixels = cand[np.linspace(0, cand.size - 1, 4).astype(int)]
    for p in pixels:
        col = np.zeros((3, 3))
        for c in range(3):
            x = np.zeros(3 * npix)
            x[c * npix + p] = 1.0
            y = ctx.apply(x)
            col[:, c] = y[[p, npix + p, 2 * npix + p]]
        b = blocks[p]
        B = np.array([[b[0], b[1], b[2]], [b[1], b[3], b[4]], [b[2], b[4], b[5]]])
        err = np.abs(B - col).max() / np.abs(col).max()
        print(f"  pixel {p}: |block - A diag| / |A diag| = {err:.2e}", flush=True)
        assert err < 1e-10
    ctx.close()
