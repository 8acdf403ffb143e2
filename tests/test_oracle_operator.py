"""Pins the oracle's operator, right-hand side, preconditioner, chi2 and PCG
(sanepic.cc 509-539, 2390-2762, 1069-1761) against dense brute force.

P is built as an explicit N x 3npix matrix, N^-1 of each (detector, segment)
block as the explicit circulant c2r(SpN)[(t - t') mod ns] (numpy's irfft),
so A = P^T N^-1 P and b = P^T N^-1 d are formed without the oracle's FFT or
scatter code.
"""
import numpy as np
import pytest

import oracle


def spn_row(ispf, ns):
    """SpN_all row for nssp == ns == 2^k, where double(ii)/ns*ns == ii."""
    row = ispf[:ns] / float(ns) / float(ns)
    row = row.copy()
    row[0] = ispf[1] * 1e-6 / float(ns) / float(ns)
    return row


def tiny_problem(ndet=2, nseg=3, ns=32, nside=2, seed=0, npix_used=None):
    rng = np.random.default_rng(seed)
    nsky = 12 * nside * nside
    pool = np.arange(nsky) if npix_used is None else rng.choice(nsky, npix_used, replace=False)
    ntot = ndet * nseg * ns
    dp = rng.choice(pool, ntot)
    psip = rng.uniform(-np.pi, np.pi, ntot)
    k = np.arange(ns)
    kk = np.minimum(k, ns - k).astype(float)
    kk[0] = 1
    ispf = [1.0 / (1 + (0.3 * (d + 1) / (kk / ns)) ** 1.0) * (1 + 0.1 * d) for d in range(ndet)]
    P = oracle.Problem(nside, ndet, nseg, ns, nthreads=2)
    for d in range(ndet):
        P.set_detector(d, 0.0, 0.0, 0.0, ispf[d])
    P.set_pixels(dp, psip)
    tod = rng.standard_normal(ntot)
    P.load_tod(tod)
    return P, dp, psip, ispf, tod


def dense(P, dp, psip, ispf):
    ndet, nseg, ns = P.ndet, P.nseg, P.ns
    _, _, _, ind = P.pixels()
    npix = P.npix
    m = ind[dp]
    N = dp.size
    Pm = np.zeros((N, 3 * npix))
    Pm[np.arange(N), m] = 1.0
    Pm[np.arange(N), npix + m] = np.cos(2 * psip)
    Pm[np.arange(N), 2 * npix + m] = np.sin(2 * psip)
    Nm = np.zeros((N, N))
    t = np.arange(ns)
    for d in range(ndet):
        c = ns * np.fft.irfft(spn_row(np.asarray(ispf[d]), ns)[: ns // 2 + 1], ns)
        C = c[(t[:, None] - t[None, :]) % ns]
        for s in range(nseg):
            o = (d * nseg + s) * ns
            Nm[o:o + ns, o:o + ns] = C
    return Pm, Nm, m


def test_apply_and_rhs_against_dense():
    P, dp, psip, ispf, tod = tiny_problem()
    Pm, Nm, _ = dense(P, dp, psip, ispf)
    A = Pm.T @ Nm @ Pm
    x = np.random.default_rng(1).standard_normal(3 * P.npix)
    assert np.allclose(P.apply_A(x), A @ x, rtol=1e-11, atol=1e-11 * np.abs(A @ x).max())
    b = Pm.T @ Nm @ tod
    assert np.allclose(P.rhs(), b, rtol=1e-11, atol=1e-11 * np.abs(b).max())
    # A is symmetric positive semi-definite
    y = np.random.default_rng(2).standard_normal(3 * P.npix)
    assert abs(x @ P.apply_A(y) - y @ P.apply_A(x)) < 1e-12 * np.linalg.norm(x) * np.linalg.norm(A @ y)
    assert np.linalg.eigvalsh(0.5 * (A + A.T)).min() > -1e-10 * np.abs(A).max()


def test_diag_pair_sum_against_bruteforce():
    P, dp, psip, ispf, tod = tiny_problem(ns=64, nside=1, seed=3)
    _, _, _, ind = P.pixels()
    ns, npix = P.ns, P.npix
    ref = np.zeros(npix)
    for d in range(P.ndet):
        nabs = ns * np.fft.irfft(np.abs(spn_row(np.asarray(ispf[d]), ns))[: ns // 2 + 1], ns)
        for s in range(P.nseg):
            o = (d * P.nseg + s) * ns
            mm = ind[dp[o:o + ns]]
            for i in range(ns):
                for k in range(ns):
                    if mm[i] == mm[k] and abs(i - k) < ns // 2:
                        ref[mm[i]] += nabs[abs(i - k)]
    assert np.allclose(P.diag_corr(), ref, rtol=1e-12)
    ax, hits, mp, sc = P.setup_pass()
    counts = np.bincount(ind[dp], minlength=npix)
    assert np.array_equal(hits, np.concatenate([2 * counts, 2 * counts]))
    assert np.allclose(mp, np.concatenate([2 * ref, ref, ref]), rtol=1e-12)
    m = ind[dp]
    cs = np.bincount(m, weights=np.cos(2 * psip), minlength=npix)
    sn = np.bincount(m, weights=np.sin(2 * psip), minlength=npix)
    assert np.allclose(sc, np.concatenate([cs, sn, np.zeros(npix)]), atol=1e-12)


def test_chi2_closed_form():
    P, dp, psip, ispf, tod = tiny_problem(seed=5)
    Pm, _, _ = dense(P, dp, psip, ispf)
    S = np.random.default_rng(6).standard_normal(3 * P.npix) * 0.1
    ps = Pm @ S
    ns = P.ns
    ref = 0.0
    for d in range(P.ndet):
        w = spn_row(np.asarray(ispf[d]), ns)[: ns // 2 + 1]
        for s in range(P.nseg):
            o = (d * P.nseg + s) * ns
            R = np.fft.rfft(tod[o:o + ns]) - np.fft.rfft(ps[o:o + ns])
            ref += np.sum(2.0 * w * np.abs(R) ** 2)
    assert abs(P.chi2(S) - ref) < 1e-11 * ref


def test_spn_index_truncation_quirk():
    # double(1)/49*49 == 0.9999999999999999 in IEEE double, so the reference's
    # long(double(ii)/nssp*ns) sends bin 1 to iSpf[0] when ns = nssp = 49
    assert 1.0 / 49 * 49 < 1.0
    ispf = np.arange(1.0, 50.0)
    row = oracle.spn_row(ispf, 49)
    low = {1, 2, 4, 8, 16, 27, 32}  # ii with (ii/49)*49 < ii in IEEE double
    for ii in range(1, 49):
        src = ii - 1 if ii in low else ii
        assert row[ii] == ispf[src] / 49.0 / 49.0, ii
    assert row[0] == ispf[1] * 1e-6 / 49.0 / 49.0
    row = oracle.spn_row(np.arange(1.0, 65.0), 64)
    assert np.array_equal(row[1:], np.arange(2.0, 65.0) / 64.0 / 64.0)


def test_pcg_against_dense_solve():
    """Tiny well-covered problem: the reference's PCG reaches the dense
    solution of the masked system (A-norm)."""
    P, dp, psip, ispf, tod = tiny_problem(ndet=2, nseg=4, ns=64, nside=1, seed=9)
    out = P.solve(stop=1e-15, maxiter=2000)
    npix = P.npix
    hits = out["hits"]
    keep = np.ones(3 * npix, bool)
    keep[npix:2 * npix] = hits[npix:] >= 4
    keep[2 * npix:] = hits[npix:] >= 4
    Pm, Nm, _ = dense(P, dp, psip, ispf)
    A = (Pm.T @ Nm @ Pm)[np.ix_(keep, keep)]
    b = (Pm.T @ Nm @ tod)[keep]
    w, V = np.linalg.eigh(0.5 * (A + A.T))
    x = V @ ((V.T @ b) / w)
    assert out["crit"][-1] <= 1e-15
    assert np.linalg.norm(b - A @ out["S"][keep]) / np.linalg.norm(b) < 1e-7
    # the DC bin's 1e-6 weight leaves near-null modes (map monopole, segment
    # offsets); on the well-determined eigen-subspace the solutions agree
    good = w > 1e-4 * w.max()
    c_or, c_x = V.T @ out["S"][keep], V.T @ x
    # there the error is bounded by |r| / lambda ~ 3e-8 / 0.06
    assert np.linalg.norm((c_or - c_x)[good]) / np.linalg.norm(c_x[good]) < 1e-6
    # masked entries are exactly zero, as the reference forces them (1316-1330)
    assert np.all(out["S"][~keep] == 0.0) and np.all(out["Mptot"][~keep] == 0.0)


@pytest.mark.parametrize("nan_at", [0, 5])
def test_tod_nan_takes_previous_sample(nan_at):
    P, dp, psip, ispf, tod = tiny_problem(ndet=1, nseg=2, ns=32, seed=11)
    b0 = P.rhs()
    t2 = tod.copy()
    if nan_at == 0:
        # a leading NaN has no previous sample: replaced value stays NaN
        t2[0] = np.nan
        P.load_tod(t2)
        assert np.isnan(P.rhs()).any()
    else:
        t2[nan_at] = np.nan
        t3 = tod.copy()
        t3[nan_at] = tod[nan_at - 1]
        P.load_tod(t2)
        b2 = P.rhs()
        P.load_tod(t3)
        assert np.array_equal(b2, P.rhs()) and not np.array_equal(b2, b0)
