"""Pins the oracle's pointingshift (sanepic.cc:3412-3476) and detector-angle
conversion (458-463).

theta/phi: against an independent construction of the detector direction,
d = cos(t) b + sin(t) (-cos(A) e_theta - sin(A) e_phi) with A = phi_ + psi
(the spherical law of cosines the reference's formulas encode), to 1e-12.
Zero offsets: exact identities the reference's arithmetic implies.
psi with non-zero offsets has no closed-form check here; it is pinned by
invariants (independence from ra, 2 pi periodicity, continuity to the
zero-offset limit): parity partially pinned.
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
