"""Pins the oracle's ang2pix_ring (sanepic.cc:2764-2843) to HEALPix.

Independent checks (Gorski et al. 2005, ApJ 622:759): the nested scheme's
face/x/y construction followed by the nested->ring renumbering, written here
from the paper's geometry; pixel centres from the ring scheme's closed form
mapping back to their own pixel; ring populations; equal areas.
"""
import numpy as np
import pytest

import oracle

JRLL = np.array([2, 2, 2, 2, 3, 3, 3, 3, 4, 4, 4, 4])
JPLL = np.array([1, 3, 5, 7, 0, 2, 4, 6, 1, 3, 5, 7])


def ang2xyf(nside, theta, phi):
    z = np.cos(theta)
    za = abs(z)
    tt = np.mod(phi, 2 * np.pi) / (0.5 * np.pi)
    if za <= 2.0 / 3.0:
        t1 = nside * (0.5 + tt)
        t2 = nside * (z * 0.75)
        jp, jm = int(t1 - t2), int(t1 + t2)
        ifp, ifm = jp // nside, jm // nside
        face = (ifp | 4) if ifp == ifm else (ifp if ifp < ifm else ifm + 8)
        return jm % nside, nside - (jp % nside) - 1, face
    ntt = min(3, int(tt))
    tp = tt - ntt
    tmp = nside * np.sqrt(3.0 * (1.0 - za))
    jp = min(nside - 1, int(tp * tmp))
    jm = min(nside - 1, int((1.0 - tp) * tmp))
    if z >= 0:
        return nside - jm - 1, nside - jp - 1, ntt
    return jp, jm, ntt + 8


def xyf2ring(nside, ix, iy, face):
    nl4 = 4 * nside
    npix = 12 * nside * nside
    jr = JRLL[face] * nside - ix - iy - 1
    if jr < nside:
        nr, nbefore, kshift = jr, 2 * jr * (jr - 1), 0
    elif jr > 3 * nside:
        nr = nl4 - jr
        nbefore, kshift = npix - 2 * (nr + 1) * nr, 0
    else:
        nr = nside
        nbefore = 2 * nside * (nside - 1) + (jr - nside) * nl4
        kshift = (jr - nside) & 1
    jp = (JPLL[face] * nr + ix - iy + 1 + kshift) // 2
    if jp > nl4:
        jp -= nl4
    elif jp < 1:
        jp += nl4
    return nbefore + jp - 1


def ring_centre(nside, ipix):
    npix = 12 * nside * nside
    ncap = 2 * nside * (nside - 1)
    if ipix < ncap:
        iring = (1 + int(np.sqrt(1 + 2 * ipix + 0.5))) >> 1
        while 2 * iring * (iring + 1) <= ipix:
            iring += 1
        while 2 * iring * (iring - 1) > ipix:
            iring -= 1
        iphi = ipix + 1 - 2 * iring * (iring - 1)
        return np.arccos(1.0 - iring * iring / (3.0 * nside * nside)), (iphi - 0.5) * np.pi / (2 * iring)
    if ipix < npix - ncap:
        ip = ipix - ncap
        iring = ip // (4 * nside) + nside
        iphi = ip % (4 * nside) + 1
        fodd = 0.5 * (1 + ((iring + nside) & 1))
        return np.arccos((2 * nside - iring) * 2.0 / (3.0 * nside)), (iphi - fodd) * np.pi / (2 * nside)
    ip = npix - ipix
    iring = (1 + int(np.sqrt(2 * ip - 1 + 0.5))) >> 1
    while 2 * iring * (iring + 1) < ip:
        iring += 1
    while 2 * iring * (iring - 1) >= ip:
        iring -= 1
    iphi = 4 * iring + 1 - (ip - 2 * iring * (iring - 1))
    return np.arccos(-1.0 + iring * iring / (3.0 * nside * nside)), (iphi - 0.5) * np.pi / (2 * iring)


@pytest.mark.parametrize("nside", [1, 2, 4, 8, 16, 64, 512, 8192])
def test_ring_against_nested_construction(nside):
    rng = np.random.default_rng(nside)
    th = np.arccos(rng.uniform(-1, 1, 4000))
    ph = rng.uniform(0, 2 * np.pi, 4000)
    got = oracle.ang2pix_ring(nside, th, ph)
    ref = np.array([xyf2ring(nside, *ang2xyf(nside, t, p)) for t, p in zip(th, ph)])
    assert np.array_equal(got, ref)


@pytest.mark.parametrize("nside", [1, 2, 3, 4, 8, 13, 32])
def test_pixel_centres_round_trip(nside):
    npix = 12 * nside * nside
    cen = [ring_centre(nside, p) for p in range(npix)]
    got = oracle.ang2pix_ring(nside, np.array([c[0] for c in cen]), np.array([c[1] for c in cen]))
    assert np.array_equal(got, np.arange(npix))


def test_ring_populations_and_equal_area():
    nside = 4
    rng = np.random.default_rng(0)
    n = 120000
    th = np.arccos(rng.uniform(-1, 1, n))
    ph = rng.uniform(0, 2 * np.pi, n)
    pix = oracle.ang2pix_ring(nside, th, ph)
    counts = np.bincount(pix, minlength=12 * nside * nside)
    exp = n / counts.size
    chi2 = np.sum((counts - exp) ** 2 / exp)
    assert chi2 < counts.size + 6 * np.sqrt(2 * counts.size)
    # rings: 4i pixels in cap ring i, 4 nside in the belt
    sizes = [4 * i for i in range(1, nside)] + [4 * nside] * (2 * nside + 1) + \
        [4 * i for i in range(nside - 1, 0, -1)]
    assert sum(sizes) == 12 * nside * nside


def test_out_of_range_rejected():
    assert oracle.ang2pix_ring(8, np.array([-1e-9]), 0.0)[0] == -1
    assert oracle.ang2pix_ring(8, np.array([np.pi + 1e-9]), 0.0)[0] == -1
    assert oracle.ang2pix_ring(0, np.array([1.0]), 0.0)[0] == -1
    assert oracle.ang2pix_ring(8193, np.array([1.0]), 0.0)[0] == -1
