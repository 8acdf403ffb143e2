"""Pins the oracle's pointingshift (sanepic.cc:3412-3476) and detector-angle
conversion (458-463).

theta/phi: against an independent construction of the detector direction,
d = cos(t) b + sin(t) (-cos(A) e_theta - sin(A) e_phi) with A = phi_ + psi
(the spherical law of cosines the reference's formulas encode), to 1e-12.
Zero offsets: exact identities the reference's arithmetic implies.
psi with non-zero offsets: against a vector construction of the detector's
polarisation direction (the boresight direction rotated about the axis that
tilts the boresight onto the detector, Rodrigues' formula, in the global
frame of the boresight), read out as the reference defines the angle, to
1e-12; the reference's formula drops the second-order term of that rotation,
so the exact rotation stays within the squared offset of it.  HWP (763-765):
psip = -(psi_o + psi_det) + 2 Omega (ii + ns iseg) in closed form.
"""
import math

import numpy as np

import oracle


def detector_direction(x, y, thetap, phip, psip):
    th_ = np.arcsin(np.sqrt(x * x + y * y))
    ph_ = np.arctan2(y, x)
    A = ph_ + psip
    b = np.stack([np.sin(thetap) * np.cos(phip), np.sin(thetap) * np.sin(phip), np.cos(thetap)])
    et = np.stack([np.cos(thetap) * np.cos(phip), np.cos(thetap) * np.sin(phip), -np.sin(thetap)])
    ep = np.stack([-np.sin(phip), np.cos(phip), np.zeros_like(phip)])
    d = np.cos(th_) * b + np.sin(th_) * (-np.cos(A) * et - np.sin(A) * ep)
    return np.arccos(np.clip(d[2], -1, 1)), np.arctan2(d[1], d[0])


def test_offsets_against_rotation():
    rng = np.random.default_rng(2)
    n = 5000
    thp = np.arccos(rng.uniform(-0.98, 0.98, n))
    php = rng.uniform(0, 2 * np.pi, n)
    psp = rng.uniform(-np.pi, np.pi, n)
    for x, y in [(0.01, 0.0), (0.0, -0.02), (0.03, 0.017), (-0.2, 0.1)]:
        th, ph, _ = oracle.pointingshift(x, y, thp, php, psp)
        th2, ph2 = detector_direction(x, y, thp, php, psp)
        assert np.abs(th - th2).max() < 1e-12
        dphi = np.angle(np.exp(1j * (ph - ph2)))
        assert np.abs(dphi * np.sin(th)).max() < 1e-12


def test_zero_offsets_identities():
    rng = np.random.default_rng(4)
    n = 20000
    thp = np.arccos(rng.uniform(-1, 1, n))
    php = rng.uniform(0, 2 * np.pi, n)
    psp = rng.uniform(-np.pi, np.pi, n)
    th, ph, ps = oracle.pointingshift(0.0, 0.0, thp, php, psp)
    assert np.array_equal(ph, php)               # atan2(+-0, y > 0) + phip == phip
    # glibc (Python's math module), not numpy's own SIMD libm
    assert np.array_equal(th, np.array([math.acos(math.cos(t)) for t in thp]))
    d = np.angle(np.exp(1j * (ps - psp)))
    assert np.abs(d).max() < 1e-12
    assert ps.min() >= -np.pi and ps.max() < np.pi + 1e-15


def test_psi_invariants_with_offsets():
    rng = np.random.default_rng(5)
    n = 3000
    thp = np.arccos(rng.uniform(-0.95, 0.95, n))
    php = rng.uniform(0, 2 * np.pi, n)
    psp = rng.uniform(-np.pi, np.pi, n)
    x, y = 0.02, -0.01
    _, _, a = oracle.pointingshift(x, y, thp, php, psp)
    _, _, b = oracle.pointingshift(x, y, thp, php + 1.234, psp)
    assert np.abs(np.angle(np.exp(1j * (a - b)))).max() < 1e-12
    _, _, c = oracle.pointingshift(x, y, thp, php, psp + 2 * np.pi)
    assert np.abs(np.angle(np.exp(1j * (a - c)))).max() < 1e-9
    _, _, s = oracle.pointingshift(1e-9, 1e-9, thp, php, psp)
    assert np.abs(np.angle(np.exp(1j * (s - psp)))).max() < 1e-6


def _frame(theta, phi):
    """Radial, e_theta, e_phi unit vectors at (theta, phi), shape (3, n)."""
    st, ct, sp, cp = np.sin(theta), np.cos(theta), np.sin(phi), np.cos(phi)
    b = np.stack([st * cp, st * sp, ct])
    et = np.stack([ct * cp, ct * sp, -st])
    ep = np.stack([-sp, cp, np.zeros_like(phi)])
    return b, et, ep


def detector_psi(x, y, thetap, phip, psip, exact=False, tilt_sign=1.0):
    """Polarisation angle of a detector at offsets (x, y) from the boresight.

    In the boresight's local frame (pole = boresight, first axis = north,
    second = -e_phi) the boresight's polarisation direction is
    e = (cos psi, sin psi, 0) and the detector sits at the pole tilted by
    t = asin(|(x, y)|) towards azimuth A = atan2(y, x) + psi, i.e. rotated
    about k = (-sin A, cos A, 0).  Rodrigues:
    R e = e cos t + (k x e) sin t + k (k.e)(1 - cos t).  The reference keeps
    the first two terms (exact=False).  The vector goes to the global frame
    through the boresight's (-e_theta, -e_phi, b) and the angle is read as in
    the reference: +-acos of its cosine with e_phi at the detector (sign of
    its z component), plus pi/2, wrapped to [-pi, pi)."""
    t = np.arcsin(np.sqrt(x * x + y * y))
    A = np.arctan2(y, x) + psip
    e = np.stack([np.cos(psip), np.sin(psip), np.zeros_like(psip)])
    k = np.stack([-np.sin(A), np.cos(A), np.zeros_like(A)])
    v = e * np.cos(t) + tilt_sign * np.cross(k, e, axis=0) * np.sin(t)
    if exact:
        v = v + k * (k * e).sum(0) * (1.0 - np.cos(t))
    b, et, ep = _frame(thetap, phip)
    g = v[0] * (-et) + v[1] * (-ep) + v[2] * b
    th_d, ph_d = detector_direction(x, y, thetap, phip, psip)
    _, _, ep_d = _frame(th_d, ph_d)
    c = np.clip((g * ep_d).sum(0) / np.linalg.norm(g, axis=0), -1.0, 1.0)
    pt = np.where(g[2] < 0, -np.arccos(c), np.arccos(c))
    return np.mod(pt + np.pi / 2 + 2 * (2 * np.pi), 2 * np.pi) - np.pi


def test_psi_with_offsets_against_rotation():
    rng = np.random.default_rng(7)
    n = 4000
    thp = np.arccos(rng.uniform(-0.97, 0.97, n))
    php = rng.uniform(0, 2 * np.pi, n)
    psp = rng.uniform(-np.pi, np.pi, n)
    for x, y in [(0.0087, 0.0), (0.0, -0.0087), (0.0061, 0.0061), (-0.03, 0.012), (0.1, -0.05)]:
        _, _, ps = oracle.pointingshift(x, y, thp, php, psp)
        want = detector_psi(x, y, thp, php, psp)
        d = np.angle(np.exp(1j * (ps - want)))
        # 1e-11: both sides read the angle through acos, whose condition
        # number grows where the cosine nears +-1
        assert np.abs(d).max() < 1e-11, (x, y, np.abs(d).max())
        assert np.median(np.abs(d)) < 1e-14
        # the exact rotation differs only at second order in the offset
        ex = detector_psi(x, y, thp, php, psp, exact=True)
        t2 = x * x + y * y
        dex = np.abs(np.angle(np.exp(1j * (ps - ex))))
        assert dex.max() < 2.0 * t2, (x, y, dex.max(), t2)


def test_psi_construction_catches_a_wrong_term():
    """Plausible slips in the tilt term (its sign, which carries the
    reference's third vector component -sin t cos phi_, or the offsets'
    signs) move psi far beyond the bound above, so the check can fail."""
    rng = np.random.default_rng(8)
    n = 500
    thp = np.arccos(rng.uniform(-0.9, 0.9, n))
    php = rng.uniform(0, 2 * np.pi, n)
    psp = rng.uniform(-np.pi, np.pi, n)
    x, y = 0.0087, 0.004
    _, _, ps = oracle.pointingshift(x, y, thp, php, psp)
    for wrong in (detector_psi(x, y, thp, php, psp, tilt_sign=-1.0),
                  detector_psi(-x, -y, thp, php, psp)):
        assert np.abs(np.angle(np.exp(1j * (ps - wrong)))).max() > 1e-6


def test_hwp_closed_form():
    """With -h 1 (763-765) the operator's angle is -(psi_o + psi_det) +
    2 Omega (ii + ns iseg) and the crosslink angle psipFP keeps
    psi_o + psi_det, the angle without the half-wave plate."""
    from tests.helpers import make_case, oracle_problem
    cfg, ds = make_case(nside=8, ndet=2, nseg=3, ns=512, offsets=[(0.3, -0.2), (-0.5, 0.1)])
    omega = 0.0123456789
    P0 = oracle_problem(cfg, ds, hwp=0)
    P1 = oracle_problem(cfg, ds, hwp=1, omega=omega)
    d0, ps0, pf0, _ = P0.pixels()
    d1, ps1, pf1, _ = P1.pixels()
    assert np.array_equal(d0, d1)
    assert np.array_equal(pf1, ps0) and np.array_equal(pf0, ps0)
    ns, nseg = cfg.ns, cfg.nseg
    t = np.tile(np.arange(nseg * ns, dtype=np.float64), cfg.ndet)  # ii + ns iseg
    assert np.array_equal(ps1, -pf1 + 2.0 * omega * t)
