"""ctypes front end of the CPU oracle (test infrastructure only).

The oracle is an FP64 CPU restatement of SANEPIC's map-making path
(src/sanepic.cc of patanch/SANEPIC); see sanepic_oracle.c for the function
map and how each function is pinned.  Only tests/, __graft_entry__.smoke()
and bench.py's cpu_baseline leg import this module.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle_sanepic.so")
_lib = None

_dp = ctypes.POINTER(ctypes.c_double)
_lp = ctypes.POINTER(ctypes.c_long)
_L = ctypes.c_long
_D = ctypes.c_double
_I = ctypes.c_int
_V = ctypes.c_void_p


def build():
    subprocess.run(["make", "-s", "-B", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        sig = {
            "or_r2c": (None, [_L, _dp, _dp]),
            "or_c2r": (None, [_L, _dp, _dp]),
            "or_c2c": (None, [_L, _dp, _I]),
            "or_spn_row": (None, [_dp, _L, _L, _dp]),
            "or_pointingshift": (None, [_D, _D, _L, _dp, _dp, _dp, _dp, _dp, _dp]),
            "or_ang2pix_ring": (_L, [_L, _D, _D]),
            "or_problem_new": (_V, [_L, _L, _L, _L, _I, _I, _D, _I]),
            "or_problem_free": (None, [_V]),
            "or_problem_npix": (_L, [_V]),
            "or_problem_nbad": (_L, [_V]),
            "or_problem_set_detector": (None, [_V, _L, _D, _D, _D, _dp, _L]),
            "or_problem_pixelize": (_I, [_V, _dp, _dp, _dp]),
            "or_problem_set_pixels": (_I, [_V, _lp, _dp]),
            "or_problem_get_pixels": (None, [_V, _lp, _dp, _dp, _lp]),
            "or_problem_load_tod": (_I, [_V, _dp]),
            "or_apply_A": (None, [_V, _dp, _dp]),
            "or_rhs": (None, [_V, _dp]),
            "or_diag_corr": (None, [_V, _dp]),
            "or_chi2": (_D, [_V, _dp]),
            "or_setup_pass": (None, [_V, _dp, _dp]),
            "or_problem_get_setup": (None, [_V, _lp, _dp, _dp]),
            "or_problem_solve": (_I, [_V, _dp, _I, _D, _I, _dp, _dp, _dp, _dp, _dp, _dp]),
            "or_expand_map": (None, [_L, _lp, _dp, _dp]),
            "or_exact_blocks": (None, [_V, _dp]),
            "or_problem_set_dipole": (_I, [_V, _dp]),
            "or_problem_set_leakage": (None, [_V, _dp]),
            "or_problem_ncomp": (_L, [_V]),
            "or_problem_solve_pcg": (_I, [_V, _dp, _I, _D, _I, _I, _I, _D, _dp, _dp, _dp, _dp, _dp]),
            "or_problem_set_gains": (_I, [_V, _dp]),
            "or_recal_sums": (None, [_V, _dp, _dp]),
            "or_recal_chi2": (_D, [_V, _dp, _dp]),
            "or_version": (_I, []),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _d(a):
    return a.ctypes.data_as(_dp)


def _l(a):
    return a.ctypes.data_as(_lp)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def r2c(x):
    x = _f64(x)
    n = x.size
    out = np.empty(2 * (n // 2 + 1))
    lib().or_r2c(n, _d(x), _d(out))
    return out[0::2] + 1j * out[1::2]


def c2r(X, n):
    X = np.ascontiguousarray(X, dtype=np.complex128)
    assert X.size == n // 2 + 1
    buf = np.empty(2 * X.size)
    buf[0::2] = X.real
    buf[1::2] = X.imag
    out = np.empty(n)
    lib().or_c2r(n, _d(buf), _d(out))
    return out


def c2c(x, sign):
    x = np.ascontiguousarray(x, dtype=np.complex128)
    buf = np.empty(2 * x.size)
    buf[0::2] = x.real
    buf[1::2] = x.imag
    lib().or_c2c(x.size, _d(buf), int(sign))
    return buf[0::2] + 1j * buf[1::2]


def spn_row(ispf, ns):
    ispf = _f64(ispf)
    row = np.empty(ns)
    lib().or_spn_row(_d(ispf), ispf.size, ns, _d(row))
    return row


def pointingshift(x, y, thetap, phip, psip):
    thetap, phip, psip = _f64(thetap), _f64(phip), _f64(psip)
    n = thetap.size
    th, ph, ps = np.empty(n), np.empty(n), np.empty(n)
    lib().or_pointingshift(x, y, n, _d(thetap), _d(phip), _d(psip), _d(th), _d(ph), _d(ps))
    return th, ph, ps


def ang2pix_ring(nside, theta, phi):
    f = lib().or_ang2pix_ring
    theta = np.atleast_1d(np.asarray(theta, dtype=np.float64))
    phi = np.broadcast_to(np.asarray(phi, dtype=np.float64), theta.shape)
    return np.array([f(nside, float(t), float(p)) for t, p in zip(theta.ravel(), phi.ravel())],
                    dtype=np.int64).reshape(theta.shape)


def expand_map(nside, indpix, mapr):
    indpix = np.ascontiguousarray(indpix, dtype=np.int64)
    mapr = _f64(mapr)
    full = np.empty(12 * nside * nside)
    lib().or_expand_map(nside, _l(indpix), _d(mapr), _d(full))
    return full


class Problem:
    """Oracle problem: all (detector, segment) blocks of one map-making run."""

    def __init__(self, nside, ndet, nseg, ns, polar=1, hwp=0, omega=0.0, nthreads=1):
        self.nside, self.ndet, self.nseg, self.ns = nside, ndet, nseg, ns
        self.h = lib().or_problem_new(nside, ndet, nseg, ns, polar, hwp, omega, nthreads)
        self._npix = None

    def __del__(self):
        if getattr(self, "h", None):
            lib().or_problem_free(self.h)
            self.h = None

    def set_detector(self, idet, lel, xel, angl, ispf):
        ispf = _f64(ispf)
        lib().or_problem_set_detector(self.h, idet, lel, xel, angl, _d(ispf), ispf.size)

    def pixelize(self, ra, dec, psi):
        ra, dec, psi = _f64(ra), _f64(dec), _f64(psi)
        assert ra.size == self.nseg * self.ns
        lib().or_problem_pixelize(self.h, _d(ra), _d(dec), _d(psi))

    def set_pixels(self, datatopix, psip):
        datatopix = np.ascontiguousarray(datatopix, dtype=np.int64)
        psip = _f64(psip)
        lib().or_problem_set_pixels(self.h, _l(datatopix), _d(psip))

    @property
    def npix(self):
        return lib().or_problem_npix(self.h)

    @property
    def ncomp(self):
        """Map components: 3 (I, Q, U), or 4 with the leakage component."""
        return lib().or_problem_ncomp(self.h)

    def set_leakage(self, relfgal):
        """The GalBP leakage component with per-detector coefficients
        (normalised as the reference does, sanepic.cc:561-570)."""
        relfgal = _f64(relfgal)
        assert relfgal.size == self.ndet
        lib().or_problem_set_leakage(self.h, _d(relfgal))

    @property
    def nbad(self):
        return lib().or_problem_nbad(self.h)

    def pixels(self):
        ntot = self.ndet * self.nseg * self.ns
        dp = np.empty(ntot, dtype=np.int64)
        ps = np.empty(ntot)
        pf = np.empty(ntot)
        ind = np.empty(12 * self.nside ** 2, dtype=np.int64)
        lib().or_problem_get_pixels(self.h, _l(dp), _d(ps), _d(pf), _l(ind))
        return dp, ps, pf, ind

    def load_tod(self, tod):
        tod = _f64(tod)
        assert tod.size == self.ndet * self.nseg * self.ns
        lib().or_problem_load_tod(self.h, _d(tod))

    def set_dipole(self, odip):
        """Orbital dipole [ndet, nseg * ns] out of the data spectra (after load_tod)."""
        odip = _f64(odip)
        assert odip.size == self.ndet * self.nseg * self.ns
        assert lib().or_problem_set_dipole(self.h, _d(odip)) == 0

    def set_gains(self, g):
        """Relative calibration [ndet, nseg] of every (det, seg) block (None: 1)."""
        if g is None:
            assert lib().or_problem_set_gains(self.h, None) == 0
            return
        g = _f64(g)
        assert g.size == self.ndet * self.nseg
        assert lib().or_problem_set_gains(self.h, _d(g)) == 0

    def recal_sums(self, S):
        """[ndet, nseg, 3]: zeta, gamma and the data term of the gain fit of map S."""
        out = np.empty(3 * self.ndet * self.nseg)
        lib().or_recal_sums(self.h, _d(_f64(S)), _d(out))
        return out.reshape(self.ndet, self.nseg, 3)

    def recal_chi2(self, S, g=None):
        """The reference's MPI_compute_chi2 of map S with gains g [ndet, nseg]."""
        return lib().or_recal_chi2(self.h, _d(_f64(S)), None if g is None else _d(_f64(g)))

    def apply_A(self, x):
        x = _f64(x)
        y = np.empty(self.ncomp * self.npix)
        lib().or_apply_A(self.h, _d(x), _d(y))
        return y

    def rhs(self):
        b = np.empty(self.ncomp * self.npix)
        lib().or_rhs(self.h, _d(b))
        return b

    def diag_corr(self):
        m = np.empty(self.npix)
        lib().or_diag_corr(self.h, _d(m))
        return m

    def chi2(self, S):
        S = _f64(S)
        return lib().or_chi2(self.h, _d(S))

    def setup_pass(self, x0=None):
        n3 = self.ncomp * self.npix
        x0 = np.zeros(n3) if x0 is None else _f64(x0)
        ax = np.empty(n3)
        lib().or_setup_pass(self.h, _d(x0), _d(ax))
        hits = np.empty(2 * self.npix, dtype=np.int64)
        mp = np.empty(n3)
        sc = np.empty(n3)
        lib().or_problem_get_setup(self.h, _l(hits), _d(mp), _d(sc))
        return ax, hits, mp, sc

    def solve(self, x0=None, maxiter=2000, stop=1e-15, refresh=10):
        n3 = self.ncomp * self.npix
        S = np.empty(n3)
        Mpt = np.empty(n3)
        b = np.empty(n3)
        crit = np.zeros(maxiter + 1)
        crit2 = np.zeros(maxiter + 1)
        chi2 = np.zeros(maxiter + 1)
        x0p = None if x0 is None else _d(_f64(x0))
        it = lib().or_problem_solve(self.h, x0p, maxiter, stop, refresh, _d(S), _d(Mpt),
                                    _d(crit), _d(crit2), _d(chi2), _d(b))
        if it < 0:
            raise RuntimeError("oracle: Mp has a zero I element")
        hits = np.empty(2 * self.npix, dtype=np.int64)
        mp = np.empty(n3)
        sc = np.empty(n3)
        lib().or_problem_get_setup(self.h, _l(hits), _d(mp), _d(sc))
        return dict(S=S, Mptot=Mpt, b=b, crit=crit[:it], crit2=crit2[:it], chi2=chi2[:it],
                    niter=it, hits=hits, sumCross=sc)

    def exact_blocks(self):
        """[npix, 6] diagonal 3x3 blocks of A (B00 B01 B02 B11 B12 B22), all lags."""
        out = np.empty(6 * self.npix)
        lib().or_exact_blocks(self.h, _d(out))
        return out.reshape(self.npix, 6)

    def solve_pcg(self, kind=1, deflate=False, x0=None, maxiter=2000, stop=1e-15, refresh=10, var00=0.0):
        """The GPU solver's PCG: kind 0 = the reference's diagonal, 1 = 3x3
        block-Jacobi; deflate = A-DEF2 with the I monopole as coarse vector;
        var00 > 0 replaces |b|^2 in crit (the recalibration loop's rule)."""
        n3 = self.ncomp * self.npix
        S = np.empty(n3)
        b = np.empty(n3)
        M6 = np.zeros(6 * self.npix)
        crit = np.zeros(maxiter + 1)
        crit2 = np.zeros(maxiter + 1)
        x0p = None if x0 is None else _d(_f64(x0))
        it = lib().or_problem_solve_pcg(self.h, x0p, maxiter, stop, refresh, int(kind),
                                        1 if deflate else 0, float(var00), _d(S), _d(crit), _d(crit2),
                                        _d(b), _d(M6))
        if it < 0:
            raise RuntimeError("oracle: Mp has a zero I element")
        hits = np.empty(2 * self.npix, dtype=np.int64)
        lib().or_problem_get_setup(self.h, _l(hits), None, None)
        return dict(S=S, b=b, crit=crit[:it], crit2=crit2[:it], niter=it, hits=hits,
                    M6=M6.reshape(self.npix, 6))
