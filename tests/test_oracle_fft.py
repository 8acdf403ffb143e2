"""Pins the oracle's FFT to FFTW's published r2c/c2r semantics.

The reference uses FFTW3 (version unpinned; Make.cm2 names 3.3.10), absent
here.  FFTW documents r2c as the unnormalised forward DFT of a real array
(bins 0..n/2) and c2r as the unnormalised inverse of the Hermitian extension
that ignores Im of the DC and (even n) Nyquist bins, so c2r(r2c(x)) = n x.
Checked against numpy.fft (pocketfft, an independent implementation) and a
brute-force O(n^2) DFT.
"""
import numpy as np
import pytest

import oracle

SIZES = [1, 2, 3, 4, 5, 6, 7, 8, 9, 12, 15, 16, 25, 30, 49, 64, 97, 100, 128, 243, 256, 1000,
         1024, 1155, 4096, 3 * 2 ** 10, 5 * 2 ** 9, 2 ** 16]


def dft(x, sign=-1):
    n = x.size
    k = np.arange(n)
    return np.exp(sign * 2j * np.pi * np.outer(k, k) / n) @ x


@pytest.mark.parametrize("n", SIZES)
def test_r2c_against_numpy(n):
    x = np.random.default_rng(n).standard_normal(n)
    X = oracle.r2c(x)
    ref = np.fft.rfft(x)
    assert np.abs(X - ref).max() <= 1e-14 * max(1.0, np.abs(ref).max()) * np.log2(n + 1)


@pytest.mark.parametrize("n", SIZES)
def test_c2r_against_numpy(n):
    rng = np.random.default_rng(1000 + n)
    X = rng.standard_normal(n // 2 + 1) + 1j * rng.standard_normal(n // 2 + 1)
    y = oracle.c2r(X, n)
    ref = n * np.fft.irfft(X, n)  # numpy's irfft also drops Im of DC / Nyquist
    assert np.abs(y - ref).max() <= 1e-14 * max(1.0, np.abs(ref).max()) * np.log2(n + 1)


@pytest.mark.parametrize("n", [1, 2, 3, 5, 8, 12, 17, 31, 64])
def test_against_bruteforce_dft(n):
    rng = np.random.default_rng(7 * n)
    z = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    for sign in (-1, 1):
        assert np.allclose(oracle.c2c(z, sign), dft(z, sign), rtol=0, atol=1e-12)
    x = z.real.copy()
    assert np.allclose(oracle.r2c(x), dft(x)[: n // 2 + 1], rtol=0, atol=1e-12)


@pytest.mark.parametrize("n", [16, 17, 1024, 999])
def test_c2r_ignores_dc_and_nyquist_imag(n):
    rng = np.random.default_rng(3)
    X = rng.standard_normal(n // 2 + 1) + 1j * rng.standard_normal(n // 2 + 1)
    Y = X.copy()
    Y[0] = Y[0].real + 5j
    if n % 2 == 0:
        Y[-1] = Y[-1].real - 3j
    X[0] = X[0].real
    if n % 2 == 0:
        X[-1] = X[-1].real
    assert np.array_equal(oracle.c2r(X, n), oracle.c2r(Y, n))


@pytest.mark.parametrize("n", [8, 15, 1024, 2 ** 19])
def test_roundtrip_is_n_times_identity(n):
    x = np.random.default_rng(n).standard_normal(n)
    assert np.allclose(oracle.c2r(oracle.r2c(x), n), n * x, rtol=0, atol=1e-10 * n)


def test_linearity_and_parseval():
    n = 2 ** 12
    rng = np.random.default_rng(11)
    a, b = rng.standard_normal(n), rng.standard_normal(n)
    assert np.allclose(oracle.r2c(2.0 * a - b), 2.0 * oracle.r2c(a) - oracle.r2c(b), atol=1e-11)
    X = oracle.r2c(a)
    full = np.abs(X[0]) ** 2 + np.abs(X[-1]) ** 2 + 2 * np.sum(np.abs(X[1:-1]) ** 2)
    assert abs(full / n - np.sum(a * a)) < 1e-9 * np.sum(a * a)
