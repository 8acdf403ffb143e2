"""Pins the oracle's GalBP leakage component (-b 1; sanepic.cc:551-575,
2860-2870, 2591-2607, 2565-2566) against dense brute force.

P gets a 4th block of columns: sample t of detector d reads L[p(t)] times
relfGal[d]; A = P^T N^-1 P and b = P^T N^-1 d are formed with explicit
matrices.  The diagonal preconditioner's leakage entry is the I entry times
the first detector's coefficient squared, and the PCG with four components
reaches the dense solution of the masked system.
"""
import numpy as np

from tests.test_oracle_operator import dense, tiny_problem

REL = np.array([0.8, -1.3])


def _dense4(P, dp, psip, ispf):
    Pm, Nm, m = dense(P, dp, psip, ispf)
    npix, ns, nseg = P.npix, P.ns, P.nseg
    det = np.arange(dp.size) // (nseg * ns)
    L = np.zeros((dp.size, npix))
    L[np.arange(dp.size), m] = REL[det]
    return np.hstack([Pm, L]), Nm


def test_leakage_operator_and_rhs_against_dense():
    P, dp, psip, ispf, tod = tiny_problem()
    P.set_leakage(REL)
    assert P.ncomp == 4
    Pm, Nm = _dense4(P, dp, psip, ispf)
    A = Pm.T @ Nm @ Pm
    x = np.random.default_rng(1).standard_normal(4 * P.npix)
    y = P.apply_A(x)
    assert y.size == 4 * P.npix
    assert np.allclose(y, A @ x, rtol=1e-11, atol=1e-11 * np.abs(A @ x).max())
    b = Pm.T @ Nm @ tod
    assert np.allclose(P.rhs(), b, rtol=1e-11, atol=1e-11 * np.abs(b).max())
    # without the 4th block the first three blocks are the plain operator's
    P3, *_ = tiny_problem()
    assert np.allclose(P3.apply_A(x[:3 * P.npix]), (A[:3 * P.npix, :3 * P.npix] @ x[:3 * P.npix]),
                       rtol=1e-11, atol=1e-11 * np.abs(A @ x).max())


def test_leakage_preconditioner_entry():
    P, dp, psip, ispf, tod = tiny_problem(ns=64, nside=1, seed=3)
    P.set_leakage(REL)
    _, hits, mp, sc = P.setup_pass()
    npix = P.npix
    assert mp.size == 4 * npix
    assert np.array_equal(mp[3 * npix:], mp[:npix] * REL[0] * REL[0])
    assert np.array_equal(mp[npix:2 * npix], mp[:npix] / 2.0)


def test_leakage_solve_reaches_dense_solution():
    P, dp, psip, ispf, tod = tiny_problem(ndet=2, nseg=4, ns=64, nside=1, seed=7)
    P.set_leakage(REL)
    Pm, Nm = _dense4(P, dp, psip, ispf)
    out = P.solve(stop=1e-28, maxiter=3000)
    npix = P.npix
    hits = out["hits"]
    keep = np.ones(4 * npix, bool)
    m = hits[npix:] < 4
    keep[npix:2 * npix][m] = False
    keep[2 * npix:3 * npix][m] = False
    idx = np.flatnonzero(keep)
    A = (Pm.T @ Nm @ Pm)[np.ix_(idx, idx)]
    b = (Pm.T @ Nm @ tod)[idx]
    w, V = np.linalg.eigh(0.5 * (A + A.T))
    xs = V @ ((V.T @ b) / w)
    e = out["S"][idx] - xs
    err = np.sqrt(abs(e @ A @ e) / abs(xs @ A @ xs))
    print(f"4-component solve: {out['niter']} iterations, A-norm error {err:.2e}")
    assert err < 1e-6
    assert not out["S"][npix:3 * npix][np.concatenate([m, m])].any()
