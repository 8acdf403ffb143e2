"""Pins the oracle's exact 3x3 diagonal blocks and its block-Jacobi / A-DEF2
PCG (the checker of the GPU's block-Jacobi path, which has no counterpart in
the reference; the reference's diagonal PCG is sanepic.cc:1340-1761).

* blocks: against columns of the oracle's FFT-based operator, A e_(p,c) read
  at pixel p -- a different computation (no pair sums) of the same entries;
* the PCG with the reference's diagonal and no deflation reproduces the
  oracle's restatement of the reference's loop iterate by iterate;
* block-Jacobi and A-DEF2 reach the dense solution of the masked system;
* A-DEF2 keeps Z^T r = 0 (the sum of the I residuals) at every iterate,
  the defining property of the deflated residual;
* the block preconditioner is the inverse of each well-observed block.
"""
import numpy as np

from tests.helpers import make_case, oracle_problem, rel


def _case(**kw):
    base = dict(nside=4, ndet=2, nseg=2, ns=512)
    base.update(kw)
    cfg, ds = make_case(**base)
    return cfg, ds, oracle_problem(cfg, ds, nthreads=2)


def _masked_apply(P, x, hits):
    npix = P.npix
    y = P.apply_A(x)
    m = hits[npix:] < 4
    y[npix:2 * npix][m] = 0.0
    y[2 * npix:][m] = 0.0
    return y


def test_exact_blocks_are_operator_columns():
    cfg, ds, P = _case()
    npix = P.npix
    B = P.exact_blocks()
    worst = 0.0
    for p in range(npix):
        col = np.zeros((3, 3))
        for c in range(3):
            x = np.zeros(3 * npix)
            x[c * npix + p] = 1.0
            col[:, c] = P.apply_A(x)[[p, npix + p, 2 * npix + p]]
        b = B[p]
        M = np.array([[b[0], b[1], b[2]], [b[1], b[3], b[4]], [b[2], b[4], b[5]]])
        worst = max(worst, np.abs(M - col).max() / np.abs(col).max())
    assert worst < 1e-11, worst


def test_exact_blocks_differ_from_truncated_sum():
    """The reference's truncated I sum (|dt| < ns/2, compute_diagPtNPCorr)
    differs from the exact B00 wherever long lags carry weight, so the
    column check above is able to fail."""
    cfg, ds, P = _case(ns=256)
    B = P.exact_blocks()
    P.setup_pass()
    trunc = P.diag_corr()
    assert rel(trunc, B[:, 0]) > 1e-6


def test_diag_pcg_reproduces_reference_loop():
    cfg, ds, P = _case()
    ref = P.solve(stop=0.0, maxiter=8)
    got = P.solve_pcg(kind=0, deflate=False, stop=0.0, maxiter=8)
    assert rel(got["S"], ref["S"]) < 1e-12
    assert np.allclose(got["crit"], ref["crit"], rtol=1e-10, atol=0)


def _dense(P, b, hits):
    npix = P.npix
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


def test_block_pcg_and_deflation_reach_dense_solution():
    cfg, ds, P = _case(nside=4, ndet=2, nseg=2, ns=1024)
    for deflate in (False, True):
        out = P.solve_pcg(kind=1, deflate=deflate, stop=1e-26, maxiter=400)
        xs, A, idx = _dense(P, out["b"], out["hits"])

        def anorm(e):
            return np.sqrt(abs(e[idx] @ A @ e[idx]))

        err = anorm(out["S"] - xs) / anorm(xs)
        print(f"deflate={deflate}: {out['niter']} iterations, A-norm error {err:.2e}")
        assert err < 1e-7


def test_deflated_residual_has_zero_monopole():
    cfg, ds, P = _case(nside=8, ndet=2, nseg=2, ns=2048)
    npix = P.npix
    for k in (1, 3, 7):
        out = P.solve_pcg(kind=1, deflate=True, stop=0.0, maxiter=k)
        r = out["b"] - _masked_apply(P, out["S"], out["hits"])
        s = abs(r[:npix].sum()) / np.abs(r[:npix]).sum()
        print(f"after {k} iterations: |Z^T r| / |r_I|_1 = {s:.2e}")
        assert s < 1e-13
    # without deflation the residual keeps a monopole (small, because the
    # monopole is nearly null for A, but 1e5 times larger than rounding)
    out = P.solve_pcg(kind=1, deflate=False, stop=0.0, maxiter=3)
    r = out["b"] - _masked_apply(P, out["S"], out["hits"])
    assert abs(r[:npix].sum()) / np.abs(r[:npix]).sum() > 1e-9


def test_block_preconditioner_inverts_blocks():
    cfg, ds, P = _case(nside=8, ndet=2, nseg=2, ns=2048)
    npix = P.npix
    out = P.solve_pcg(kind=1, stop=0.0, maxiter=1)
    B, M = P.exact_blocks(), out["M6"]
    hits = out["hits"]
    full = lambda b: np.array([[b[0], b[1], b[2]], [b[1], b[3], b[4]], [b[2], b[4], b[5]]])
    n_checked = 0
    for p in np.flatnonzero(hits[npix:] >= 16)[:200]:
        Bp, Mp = full(B[p]), full(M[p])
        det = np.linalg.det(Bp)
        if det <= 1e-8 * Bp[0, 0] * Bp[1, 1] * Bp[2, 2]:
            continue
        assert np.abs(Mp @ Bp - np.eye(3)).max() < 1e-9
        n_checked += 1
    assert n_checked > 20
    # masked pixels keep 1/B00 on I only
    for p in np.flatnonzero(hits[npix:] < 4):
        assert M[p, 0] == 1.0 / B[p, 0] and not M[p, 1:].any()
