"""sanepic_amd -- MI355X-native CG GLS map-maker (drop-in for SANEPIC's solve).

The compute lives in ``lib/libsanepic_hip.so`` (hand-written HIP for gfx950)
behind the C ABI of ``include/sanepic_hip.h``.  This module is a thin ctypes
front end used by the tests and the benchmark; the drop-in command-line
program is ``bin/sanepic``.  There is no CPU fallback: if the library is
missing, importing ``Context`` raises.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libsanepic_hip.so")
EXE_PATH = os.path.join(HERE, "bin", "sanepic")
CSRC = os.path.join(HERE, "csrc")

FP64, MIXED = 0, 1
PRECOND_DIAG, PRECOND_BLOCK3, PRECOND_LEARNED = 0, 1, 2
LOSS_ANORM, LOSS_BLOCK, LOSS_RESIDUAL = 0, 1, 2
HARVEST_BLOCK3, HARVEST_LEARNED = 0, 1

# The header's function list; tests check the library exports all of them.
ABI_FUNCTIONS = [
    "sanepic_create", "sanepic_destroy", "sanepic_last_error", "sanepic_abi_version",
    "sanepic_set_detector", "sanepic_set_pointing", "sanepic_set_tod", "sanepic_pixelize",
    "sanepic_get_pixels", "sanepic_get_indpix", "sanepic_setup", "sanepic_get_setup",
    "sanepic_get_blocks", "sanepic_get_rhs", "sanepic_apply", "sanepic_chi2", "sanepic_solve_begin",
    "sanepic_solve_step", "sanepic_get_map", "sanepic_get_nvar", "sanepic_get_sumcross",
    "sanepic_true_residual", "sanepic_set_profiling", "sanepic_kernel_time",
    "sanepic_device_bytes",
    "sanepic_run_records",
    "sanepic_fp64_angle_bytes", "sanepic_fp64_sorted_angle_bytes",
    "sanepic_unet_create", "sanepic_unet_destroy", "sanepic_unet_last_error",
    "sanepic_unet_nparams", "sanepic_unet_init", "sanepic_unet_get_params",
    "sanepic_unet_set_params", "sanepic_unet_forward", "sanepic_unet_backward",
    "sanepic_learn_attach", "sanepic_learn_face_index", "sanepic_ring_to_face",
    "sanepic_learn_features", "sanepic_learn_harvest", "sanepic_learn_npairs",
    "sanepic_learn_clear", "sanepic_learn_truncate", "sanepic_learn_get_pair", "sanepic_learn_train",
    "sanepic_learn_eval", "sanepic_learn_grad", "sanepic_learn_stats", "sanepic_host_alloc", "sanepic_host_free", "sanepic_set_pointing_async",
    "sanepic_set_tod_async", "sanepic_autotune", "sanepic_set_fft_split", "sanepic_create_uneven",
    "sanepic_get_seglen", "sanepic_set_pointing_range", "sanepic_set_tod_range",
    "sanepic_set_dipole", "sanepic_set_dipole_range", "sanepic_apply_noalloc", "sanepic_set_leakage",
    "sanepic_set_recalib", "sanepic_set_gains", "sanepic_recal_sums",
]


class Config(ctypes.Structure):
    _fields_ = [("nside", ctypes.c_int64), ("ndet", ctypes.c_int64), ("nseg", ctypes.c_int64),
                ("ns", ctypes.c_int64), ("polar", ctypes.c_int32), ("hwp", ctypes.c_int32),
                ("omega_hwp", ctypes.c_double), ("precision", ctypes.c_int32),
                ("device", ctypes.c_int32)]


class UnevenConfig(ctypes.Structure):
    _fields_ = [("nside", ctypes.c_int64), ("ndet", ctypes.c_int64), ("nsamp", ctypes.c_int64),
                ("seglen", ctypes.POINTER(ctypes.c_int64)), ("polar", ctypes.c_int32),
                ("hwp", ctypes.c_int32), ("omega_hwp", ctypes.c_double),
                ("precision", ctypes.c_int32), ("device", ctypes.c_int32)]


class SolveOpts(ctypes.Structure):
    _fields_ = [("precond", ctypes.c_int32), ("refresh", ctypes.c_int32),
                ("deflate", ctypes.c_int32), ("fallback_growth", ctypes.c_double),
                ("graph", ctypes.c_int32), ("fcg_window", ctypes.c_int32),
                ("learned_iters", ctypes.c_int32), ("refresh_fp32", ctypes.c_int32),
                ("var00", ctypes.c_double)]


class IterInfo(ctypes.Structure):
    _fields_ = [("iter", ctypes.c_int32), ("crit", ctypes.c_double), ("crit2", ctypes.c_double),
                ("chi2", ctypes.c_double), ("alpha", ctypes.c_double), ("beta", ctypes.c_double)]


class HarvestOpts(ctypes.Structure):
    _fields_ = [("seed", ctypes.c_uint64), ("deflate", ctypes.c_int32), ("maxiter", ctypes.c_int32),
                ("keep_crit", ctypes.c_double), ("stop_crit", ctypes.c_double),
                ("capacity", ctypes.c_int64), ("policy", ctypes.c_int32), ("noise_gain", ctypes.c_double)]


class TrainOpts(ctypes.Structure):
    _fields_ = [("steps", ctypes.c_int64), ("lr", ctypes.c_float), ("lr_final", ctypes.c_float),
                ("beta1", ctypes.c_float), ("beta2", ctypes.c_float), ("eps", ctypes.c_float),
                ("weight_decay", ctypes.c_float), ("seed", ctypes.c_uint64),
                ("first", ctypes.c_int64), ("count", ctypes.c_int64), ("loss_kind", ctypes.c_int32),
                ("sched_offset", ctypes.c_int64), ("sched_total", ctypes.c_int64)]


class SanepicError(RuntimeError):
    pass


def build(jobs=8):
    """Compiles the HIP library and the drop-in executable (always rebuilds)."""
    subprocess.run(["make", "-B", f"-j{jobs}", "-C", CSRC], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise SanepicError(f"{LIB_PATH} is missing: build it with __graft_entry__.build()")
        L = ctypes.CDLL(LIB_PATH)
        P = ctypes.c_void_p
        d_p = ctypes.POINTER(ctypes.c_double)
        i64p = ctypes.POINTER(ctypes.c_int64)
        i64 = ctypes.c_int64
        f_p = ctypes.POINTER(ctypes.c_float)
        sig = {
            "sanepic_create": [ctypes.POINTER(Config), ctypes.POINTER(P)],
            "sanepic_destroy": [P],
            "sanepic_abi_version": [],
            "sanepic_set_detector": [P, i64, ctypes.c_double, ctypes.c_double, ctypes.c_double, d_p, i64],
            "sanepic_set_pointing": [P, i64, d_p, d_p, d_p],
            "sanepic_set_tod": [P, i64, i64, d_p],
            "sanepic_pixelize": [P, i64p, i64p],
            "sanepic_get_pixels": [P, i64, i64, i64p],
            "sanepic_get_indpix": [P, i64p],
            "sanepic_setup": [P],
            "sanepic_get_setup": [P, i64p, d_p, d_p],
            "sanepic_get_blocks": [P, d_p],
            "sanepic_get_rhs": [P, d_p],
            "sanepic_apply": [P, d_p, d_p],
            "sanepic_chi2": [P, d_p, d_p],
            "sanepic_solve_begin": [P, ctypes.POINTER(SolveOpts), d_p],
            "sanepic_solve_step": [P, ctypes.POINTER(IterInfo)],
            "sanepic_get_map": [P, d_p],
            "sanepic_get_nvar": [P, d_p],
            "sanepic_get_sumcross": [P, d_p],
            "sanepic_true_residual": [P, d_p],
            "sanepic_set_profiling": [P, ctypes.c_int32],
            "sanepic_kernel_time": [P, ctypes.c_int32, d_p, i64p],
            "sanepic_device_bytes": [P, i64p],
            "sanepic_run_records": [P, i64p, ctypes.POINTER(ctypes.c_int32)],
            "sanepic_fp64_angle_bytes": [P, ctypes.POINTER(ctypes.c_int32)],
            "sanepic_fp64_sorted_angle_bytes": [P, ctypes.POINTER(ctypes.c_int32)],
            "sanepic_unet_create": [i64, ctypes.c_int32, ctypes.POINTER(P)],
            "sanepic_unet_destroy": [P],
            "sanepic_unet_nparams": [P, i64p],
            "sanepic_unet_init": [P, ctypes.c_uint64],
            "sanepic_unet_get_params": [P, f_p],
            "sanepic_unet_set_params": [P, f_p],
            "sanepic_unet_forward": [P, f_p, f_p],
            "sanepic_unet_backward": [P, f_p, f_p, f_p, f_p],
            "sanepic_learn_attach": [P, P],
            "sanepic_learn_face_index": [P, i64p],
            "sanepic_ring_to_face": [i64, i64, i64p, i64p],
            "sanepic_learn_features": [P, d_p, f_p, d_p],
            "sanepic_learn_harvest": [P, ctypes.POINTER(HarvestOpts), i64p,
                                      ctypes.POINTER(ctypes.c_int32)],
            "sanepic_learn_npairs": [P, i64p],
            "sanepic_learn_clear": [P],
            "sanepic_learn_truncate": [P, ctypes.c_int64],
            "sanepic_learn_get_pair": [P, i64, d_p, d_p],
            "sanepic_learn_train": [P, ctypes.POINTER(TrainOpts), f_p],
            "sanepic_learn_eval": [P, i64, ctypes.c_int32, d_p],
            "sanepic_learn_grad": [P, i64, ctypes.c_int32, d_p, f_p],
            "sanepic_learn_stats": [P, i64p, i64p, i64p],
            "sanepic_host_alloc": [ctypes.c_uint64, ctypes.POINTER(P)],
            "sanepic_host_free": [P],
            "sanepic_set_pointing_async": [P, i64, d_p, d_p, d_p],
            "sanepic_set_tod_async": [P, i64, i64, d_p],
            "sanepic_autotune": [P, ctypes.c_char_p, ctypes.POINTER(ctypes.c_int32), d_p],
            "sanepic_set_fft_split": [P, ctypes.c_int32],
            "sanepic_create_uneven": [ctypes.POINTER(UnevenConfig), ctypes.POINTER(P)],
            "sanepic_get_seglen": [P, i64, i64p, i64p],
            "sanepic_set_pointing_range": [P, i64, i64, d_p, d_p, d_p, ctypes.c_int32],
            "sanepic_set_tod_range": [P, i64, i64, i64, d_p, ctypes.c_int32],
            "sanepic_set_dipole": [P, i64, i64, d_p],
            "sanepic_set_dipole_range": [P, i64, i64, i64, d_p],
            "sanepic_apply_noalloc": [P, d_p, d_p],
            "sanepic_set_leakage": [P, d_p],
            "sanepic_set_recalib": [P, ctypes.c_int32],
            "sanepic_set_gains": [P, d_p],
            "sanepic_recal_sums": [P, d_p, d_p],
        }
        for name, args in sig.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = (None if name in ("sanepic_destroy", "sanepic_unet_destroy", "sanepic_host_free")
                          else ctypes.c_int)
        L.sanepic_last_error.argtypes = [P]
        L.sanepic_last_error.restype = ctypes.c_char_p
        L.sanepic_unet_last_error.argtypes = [P]
        L.sanepic_unet_last_error.restype = ctypes.c_char_p
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _ip(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _fp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def ring_to_face(nside, ring):
    """Face position f*S*S + y*S + x of RING pixels (nested faces, host only)."""
    ring = np.ascontiguousarray(ring, dtype=np.int64)
    pos = np.empty_like(ring)
    rc = lib().sanepic_ring_to_face(nside, ring.size, _ip(ring), _ip(pos))
    if rc:
        raise SanepicError("sanepic_ring_to_face: bad nside or pixel")
    return pos


class PinnedInputs:
    """A synth dataset copied into pinned host memory (sanepic_host_alloc),
    the form Context.load_async uploads at the full host-link rate.  Keep it
    alive until Context.setup() has returned."""

    def __init__(self, cfg, ds):
        self._bufs = []
        nsamp = cfg.nsamp
        self.ra = self._alloc(nsamp)
        self.dec = self._alloc(nsamp)
        self.psi = self._alloc(nsamp)
        self.ra[:] = ds["ra"]
        self.dec[:] = ds["dec"]
        self.psi[:] = ds["psi"]
        if hasattr(cfg, "seglen"):  # per-detector lengths: one flat buffer, a view per detector
            off = np.concatenate([[0], np.cumsum([len(t) for t in ds["tod"]])])
            flat = self._alloc(int(off[-1]))
            self.tod = [flat[off[d]:off[d + 1]] for d in range(cfg.ndet)]
            for d in range(cfg.ndet):
                self.tod[d][:] = ds["tod"][d]
        else:
            self.tod = self._alloc(cfg.ndet * nsamp).reshape(cfg.ndet, nsamp)
            self.tod[:] = ds["tod"]
        self.ispf = ds["ispf"]

    def _alloc(self, n):
        p = ctypes.c_void_p()
        if lib().sanepic_host_alloc(n * 8, ctypes.byref(p)):
            raise SanepicError(f"sanepic_host_alloc: {lib().sanepic_last_error(None).decode()}")
        self._bufs.append(p)
        return np.ctypeslib.as_array((ctypes.c_double * n).from_address(p.value))

    def close(self):
        for p in self._bufs:
            lib().sanepic_host_free(p)
        self._bufs = []
        self.ra = self.dec = self.psi = self.tod = None

    def __del__(self):
        self.close()


class UNet:
    """The learned preconditioner's network (wraps sanepic_unet): BF16 U-Net
    on the 12 nested base faces, widths 16-32-64-128, 3x3 convolutions."""

    NFEAT, NOUT = 10, 3

    def __init__(self, nside, device=0, seed=None):
        self.nside = nside
        self.h = None
        h = ctypes.c_void_p()
        rc = lib().sanepic_unet_create(nside, device, ctypes.byref(h))
        if rc:
            raise SanepicError(f"sanepic_unet_create: {lib().sanepic_unet_last_error(None).decode()}")
        self.h = h
        n = ctypes.c_int64()
        lib().sanepic_unet_nparams(self.h, ctypes.byref(n))
        self.nparams = n.value
        if seed is not None:
            self.init(seed)

    def close(self):
        if self.h:
            lib().sanepic_unet_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def _chk(self, rc, what):
        if rc:
            raise SanepicError(f"{what} ({rc}): {lib().sanepic_unet_last_error(self.h).decode()}")

    def init(self, seed):
        self._chk(lib().sanepic_unet_init(self.h, seed), "unet_init")

    def params(self):
        out = np.empty(self.nparams, dtype=np.float32)
        self._chk(lib().sanepic_unet_get_params(self.h, _fp(out)), "unet_get_params")
        return out

    def set_params(self, p):
        p = _f32(p)
        assert p.size == self.nparams
        self._chk(lib().sanepic_unet_set_params(self.h, _fp(p)), "unet_set_params")

    def forward(self, x):
        """x[12][S][S][10] float32 -> y[12][S][S][3]."""
        x = _f32(x)
        S = self.nside
        assert x.size == 12 * S * S * self.NFEAT
        y = np.empty((12, S, S, self.NOUT), dtype=np.float32)
        self._chk(lib().sanepic_unet_forward(self.h, _fp(x), _fp(y)), "unet_forward")
        return y

    def backward(self, x, dy):
        """(y, gradient of sum(y * dy) w.r.t. every parameter)."""
        x, dy = _f32(x), _f32(dy)
        S = self.nside
        y = np.empty((12, S, S, self.NOUT), dtype=np.float32)
        g = np.empty(self.nparams, dtype=np.float32)
        self._chk(lib().sanepic_unet_backward(self.h, _fp(x), _fp(dy), _fp(y), _fp(g)), "unet_backward")
        return y, g


class Context:
    """One map-making run on one GPU (wraps sanepic_ctx)."""

    def __init__(self, nside, ndet, nseg, ns, hwp=0, omega=0.0, precision=FP64, device=0):
        self.nside, self.ndet, self.nseg, self.ns = nside, ndet, nseg, ns
        self.h = None
        cfg = Config(nside, ndet, nseg, ns, 1, hwp, omega, precision, device)
        h = ctypes.c_void_p()
        rc = lib().sanepic_create(ctypes.byref(cfg), ctypes.byref(h))
        if rc:
            raise SanepicError(f"sanepic_create: {lib().sanepic_last_error(None).decode()}")
        self.h = h
        self.npix = None
        self.nc = 3

    @classmethod
    def uneven(cls, nside, seglen, nsamp, hwp=0, omega=0.0, precision=FP64, device=0):
        """A context whose detectors have their own segment lengths (seglen[d])
        on one boresight timeline of nsamp samples."""
        obj = cls.__new__(cls)
        obj.h = None
        sl = np.ascontiguousarray(seglen, dtype=np.int64)
        cfg = UnevenConfig(nside, sl.size, nsamp, _ip(sl), 1, hwp, omega, precision, device)
        h = ctypes.c_void_p()
        rc = lib().sanepic_create_uneven(ctypes.byref(cfg), ctypes.byref(h))
        if rc:
            raise SanepicError(f"sanepic_create_uneven: {lib().sanepic_last_error(None).decode()}")
        obj.h = h
        obj.nside, obj.ndet, obj.nseg, obj.ns, obj.npix = nside, sl.size, None, None, None
        obj.nc = 3
        obj.nsamp = nsamp
        return obj

    def seglen(self, idet):
        """(segment length, number of segments) of detector idet."""
        L, n = ctypes.c_int64(), ctypes.c_int64()
        self._chk(lib().sanepic_get_seglen(self.h, idet, ctypes.byref(L), ctypes.byref(n)), "get_seglen")
        return L.value, n.value

    def set_pointing_range(self, t0, ra, dec, psi, queued=False):
        ra, dec, psi = _f64(ra), _f64(dec), _f64(psi)
        self._keep = getattr(self, "_keep", []) + ([ra, dec, psi] if queued else [])
        self._chk(lib().sanepic_set_pointing_range(self.h, t0, ra.size, _dp(ra), _dp(dec), _dp(psi),
                                                   1 if queued else 0), "set_pointing_range")

    def set_tod_range(self, idet, t0, tod, queued=False):
        tod = _f64(tod)
        self._keep = getattr(self, "_keep", []) + ([tod] if queued else [])
        self._chk(lib().sanepic_set_tod_range(self.h, idet, t0, tod.size, _dp(tod), 1 if queued else 0),
                  "set_tod_range")

    def set_dipole_range(self, idet, t0, odip):
        odip = _f64(odip)
        self._chk(lib().sanepic_set_dipole_range(self.h, idet, t0, odip.size, _dp(odip)), "set_dipole_range")

    def load_uneven(self, cfg, ds, queued=False):
        """Feeds a synth.make_uneven_dataset dataset (dict) or a PinnedInputs of one."""
        get = (lambda k: ds[k]) if isinstance(ds, dict) else (lambda k: getattr(ds, k))
        for k in range(self.ndet):
            lel, xel = cfg.offsets_deg[k]
            self.set_detector(k, lel, xel, cfg.det_angles_deg[k], get("ispf")[k])
        self.set_pointing_range(0, get("ra"), get("dec"), get("psi"), queued)
        tod = get("tod")
        for k in range(self.ndet):
            self.set_tod_range(k, 0, tod[k], queued)

    def close(self):
        if self.h:
            lib().sanepic_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _chk(self, rc, what):
        if rc:
            raise SanepicError(f"{what} ({rc}): {lib().sanepic_last_error(self.h).decode()}")

    def set_detector(self, idet, lel, xel, angle, ispf):
        ispf = _f64(ispf)
        self._chk(lib().sanepic_set_detector(self.h, idet, lel, xel, angle, _dp(ispf), ispf.size),
                  "set_detector")

    def set_pointing(self, iseg, ra, dec, psi):
        ra, dec, psi = _f64(ra), _f64(dec), _f64(psi)
        assert ra.size == dec.size == psi.size == self.ns
        self._chk(lib().sanepic_set_pointing(self.h, iseg, _dp(ra), _dp(dec), _dp(psi)),
                  "set_pointing")

    def set_tod(self, idet, iseg, tod):
        tod = _f64(tod)
        assert self.ns is None or tod.size == self.ns
        self._chk(lib().sanepic_set_tod(self.h, idet, iseg, _dp(tod)), "set_tod")

    def set_leakage(self, relfgal):
        """GalBP leakage component: maps become [I|Q|U|L] (4 npix)."""
        relfgal = _f64(relfgal)
        assert relfgal.size == self.ndet
        self._chk(lib().sanepic_set_leakage(self.h, _dp(relfgal)), "set_leakage")
        self.nc = 4

    def set_recalib(self, on=True):
        """Keep the TOD and dipole resident for the relative calibration loop
        (before setup; one segment length for every detector)."""
        self._chk(lib().sanepic_set_recalib(self.h, 1 if on else 0), "set_recalib")

    def set_gains(self, g):
        """relCalib [ndet, nseg] of every block; rebuilds b (a new solve_begin follows)."""
        g = _f64(g)
        assert g.size == self.ndet * self.nseg
        self._chk(lib().sanepic_set_gains(self.h, _dp(g)), "set_gains")

    def recal_sums(self, x):
        """[ndet, nseg, 3]: zeta, gamma and the data term of the gain fit of map x."""
        x = _f64(x)
        out = np.empty(3 * self.ndet * self.nseg)
        self._chk(lib().sanepic_recal_sums(self.h, _dp(x), _dp(out)), "recal_sums")
        return out.reshape(self.ndet, self.nseg, 3)

    def set_dipole(self, idet, iseg, odip):
        odip = _f64(odip)
        assert odip.size == self.ns
        self._chk(lib().sanepic_set_dipole(self.h, idet, iseg, _dp(odip)), "set_dipole")

    def load(self, cfg, ds):
        """Feeds a sanepic_amd.synth dataset (all detectors and segments)."""
        ns = self.ns
        for k in range(self.ndet):
            lel, xel = cfg.offsets_deg[k]
            self.set_detector(k, lel, xel, cfg.det_angles_deg[k], ds["ispf"][k])
        for s in range(self.nseg):
            sl = slice(s * ns, (s + 1) * ns)
            self.set_pointing(s, ds["ra"][sl], ds["dec"][sl], ds["psi"][sl])
        for k in range(self.ndet):
            for s in range(self.nseg):
                self.set_tod(k, s, ds["tod"][k, s * ns:(s + 1) * ns])
        if ds.get("odip") is not None:
            for k in range(self.ndet):
                for s in range(self.nseg):
                    self.set_dipole(k, s, ds["odip"][k, s * ns:(s + 1) * ns])

    def load_async(self, cfg, pin):
        """Queues the uploads of a PinnedInputs (returns before they finish)."""
        ns = self.ns
        for k in range(self.ndet):
            lel, xel = cfg.offsets_deg[k]
            self.set_detector(k, lel, xel, cfg.det_angles_deg[k], pin.ispf[k])
        for s in range(self.nseg):
            sl = slice(s * ns, (s + 1) * ns)
            self._chk(lib().sanepic_set_pointing_async(self.h, s, _dp(pin.ra[sl]), _dp(pin.dec[sl]),
                                                       _dp(pin.psi[sl])), "set_pointing_async")
        for k in range(self.ndet):
            for s in range(self.nseg):
                self._chk(lib().sanepic_set_tod_async(self.h, k, s, _dp(pin.tod[k, s * ns:(s + 1) * ns])),
                          "set_tod_async")

    def pixelize(self):
        npix = ctypes.c_int64()
        ndef = ctypes.c_int64()
        self._chk(lib().sanepic_pixelize(self.h, ctypes.byref(npix), ctypes.byref(ndef)), "pixelize")
        self.npix = npix.value
        return npix.value, ndef.value

    def pixels(self, idet, iseg):
        n = self.ns if self.ns else self.seglen(idet)[0]
        out = np.empty(n, dtype=np.int64)
        self._chk(lib().sanepic_get_pixels(self.h, idet, iseg, _ip(out)), "get_pixels")
        return out

    def all_pixels(self):
        if self.nseg is None:  # per-detector lengths
            return np.concatenate([self.pixels(d, s) for d in range(self.ndet)
                                   for s in range(self.seglen(d)[1])])
        return np.concatenate([self.pixels(d, s) for d in range(self.ndet) for s in range(self.nseg)])

    def indpix(self):
        out = np.empty(12 * self.nside ** 2, dtype=np.int64)
        self._chk(lib().sanepic_get_indpix(self.h, _ip(out)), "get_indpix")
        return out

    def autotune(self, cache=None):
        """Times every four-step FFT split and keeps the fastest (before setup)."""
        l1 = ctypes.c_int32()
        ms = ctypes.c_double()
        self._chk(lib().sanepic_autotune(self.h, cache.encode() if cache else None, ctypes.byref(l1),
                                         ctypes.byref(ms)), "autotune")
        return l1.value, ms.value

    def set_fft_split(self, log2_n1):
        self._chk(lib().sanepic_set_fft_split(self.h, log2_n1), "set_fft_split")

    def setup(self):
        self._chk(lib().sanepic_setup(self.h), "setup")

    def get_setup(self):
        n = self.npix
        hits = np.empty(2 * n, dtype=np.int64)
        mp = np.empty(self.nc * n)
        sc = np.empty(3 * n)
        self._chk(lib().sanepic_get_setup(self.h, _ip(hits), _dp(mp), _dp(sc)), "get_setup")
        return hits, mp, sc

    def blocks(self):
        out = np.empty(6 * self.npix)
        self._chk(lib().sanepic_get_blocks(self.h, _dp(out)), "get_blocks")
        return out.reshape(self.npix, 6)

    def rhs(self):
        b = np.empty(self.nc * self.npix)
        self._chk(lib().sanepic_get_rhs(self.h, _dp(b)), "get_rhs")
        return b

    def apply(self, x):
        x = _f64(x)
        y = np.empty(self.nc * self.npix)
        self._chk(lib().sanepic_apply(self.h, _dp(x), _dp(y)), "apply")
        return y

    def apply_noalloc(self, x):
        """A x through the entry point that allocates no device memory."""
        x = _f64(x)
        y = np.empty(self.nc * self.npix)
        self._chk(lib().sanepic_apply_noalloc(self.h, _dp(x), _dp(y)), "apply_noalloc")
        return y

    def chi2(self, x):
        x = _f64(x)
        out = ctypes.c_double()
        self._chk(lib().sanepic_chi2(self.h, _dp(x), ctypes.byref(out)), "chi2")
        return out.value

    def solve_begin(self, precond=PRECOND_DIAG, refresh=10, x0=None, deflate=False, fallback_growth=0.0,
                    graph=False, fcg_window=0, learned_iters=0, refresh_fp32=False, var00=0.0):
        opts = SolveOpts(precond, refresh, 1 if deflate else 0, fallback_growth, 1 if graph else 0,
                         fcg_window, learned_iters, 1 if refresh_fp32 else 0, float(var00))
        xp = None if x0 is None else _dp(_f64(x0))
        self._x0_keep = None if x0 is None else _f64(x0)
        self._chk(lib().sanepic_solve_begin(self.h, ctypes.byref(opts), xp), "solve_begin")

    def step(self):
        info = IterInfo()
        self._chk(lib().sanepic_solve_step(self.h, ctypes.byref(info)), "solve_step")
        return dict(iter=info.iter, crit=info.crit, crit2=info.crit2, chi2=info.chi2,
                    alpha=info.alpha, beta=info.beta)

    def solve(self, precond=PRECOND_DIAG, stop=1e-15, maxiter=2000, refresh=10, x0=None,
              deflate=False, fallback_growth=0.0):
        """The reference's loop: while ((iter < maxiter && crit > stop) || iter == 0)."""
        self.solve_begin(precond, refresh, x0, deflate, fallback_growth)
        hist = []
        while True:
            info = self.step()
            hist.append(info)
            it = info["iter"] + 1
            if not ((it < maxiter and info["crit"] > stop) or it == 0):
                break
        return self.map(), hist

    def map(self):
        x = np.empty(self.nc * self.npix)
        self._chk(lib().sanepic_get_map(self.h, _dp(x)), "get_map")
        return x

    def nvar(self):
        x = np.empty(self.nc * self.npix)
        self._chk(lib().sanepic_get_nvar(self.h, _dp(x)), "get_nvar")
        return x

    def sumcross(self):
        x = np.empty(3 * self.npix)
        self._chk(lib().sanepic_get_sumcross(self.h, _dp(x)), "get_sumcross")
        return x

    def true_residual(self):
        out = ctypes.c_double()
        self._chk(lib().sanepic_true_residual(self.h, ctypes.byref(out)), "true_residual")
        return out.value

    def set_profiling(self, on):
        self._chk(lib().sanepic_set_profiling(self.h, 1 if on else 0), "set_profiling")

    def kernel_time(self, which):
        ms = ctypes.c_double()
        n = ctypes.c_int64()
        self._chk(lib().sanepic_kernel_time(self.h, which, ctypes.byref(ms), ctypes.byref(n)),
                  "kernel_time")
        return ms.value, n.value

    def device_bytes(self):
        b = ctypes.c_int64()
        self._chk(lib().sanepic_device_bytes(self.h, ctypes.byref(b)), "device_bytes")
        return b.value

    def fp64_angle_bytes(self):
        """Bytes per sample of the FP64 operator's angle (8: 64-bit turns, 16: pairs)."""
        b = ctypes.c_int32()
        self._chk(lib().sanepic_fp64_angle_bytes(self.h, ctypes.byref(b)), "fp64_angle_bytes")
        return b.value

    def fp64_sorted_angle_bytes(self):
        """Bytes per sorted entry of the angle read by the FP64 P^T over
        samples (8: 64-bit turns, 16: pairs, 0: P^T over run records)."""
        b = ctypes.c_int32()
        self._chk(lib().sanepic_fp64_sorted_angle_bytes(self.h, ctypes.byref(b)),
                  "fp64_sorted_angle_bytes")
        return b.value

    def run_records(self):
        """(number of run records, whether the operator uses them)."""
        n = ctypes.c_int64()
        on = ctypes.c_int32()
        self._chk(lib().sanepic_run_records(self.h, ctypes.byref(n), ctypes.byref(on)), "run_records")
        return n.value, bool(on.value)

    # ------------------------------------------------ learned preconditioner
    def learn_attach(self, unet):
        self._chk(lib().sanepic_learn_attach(self.h, unet.h if unet is not None else None),
                  "learn_attach")
        self._unet = unet

    def face_index(self):
        out = np.empty(self.npix, dtype=np.int64)
        self._chk(lib().sanepic_learn_face_index(self.h, _ip(out)), "learn_face_index")
        return out

    def features(self, r):
        """Network input of residual r (block layout) and its sigma."""
        r = _f64(r)
        un = getattr(self, "_unet", None)
        S = un.nside if un is not None else self.nside  # the network's (the map's or coarser)
        feat = np.empty((12, S, S, UNet.NFEAT), dtype=np.float32)
        sig = ctypes.c_double()
        self._chk(lib().sanepic_learn_features(self.h, _dp(r), _fp(feat), ctypes.byref(sig)),
                  "learn_features")
        return feat, sig.value

    def harvest(self, seed, capacity, deflate=True, maxiter=400, keep_crit=1e-12, stop_crit=1e-17,
                policy=None, noise_gain=1.0):
        """Append training pairs from one solve of a synthetic right-hand side
        (seed); policy HARVEST_BLOCK3 (default) or HARVEST_LEARNED (the attached
        network's own flexible PCG).  Returns (pairs added, iterations)."""
        o = HarvestOpts(seed, 1 if deflate else 0, maxiter, keep_crit, stop_crit, capacity,
                        HARVEST_BLOCK3 if policy is None else policy, noise_gain)
        added = ctypes.c_int64()
        iters = ctypes.c_int32()
        self._chk(lib().sanepic_learn_harvest(self.h, ctypes.byref(o), ctypes.byref(added),
                                              ctypes.byref(iters)), "learn_harvest")
        return added.value, iters.value

    def npairs(self):
        n = ctypes.c_int64()
        self._chk(lib().sanepic_learn_npairs(self.h, ctypes.byref(n)), "learn_npairs")
        return n.value

    def clear_pairs(self):
        self._chk(lib().sanepic_learn_clear(self.h), "learn_clear")

    def truncate_pairs(self, n):
        self._chk(lib().sanepic_learn_truncate(self.h, n), "learn_truncate")

    def pair(self, k):
        r = np.empty(3 * self.npix)
        e = np.empty(3 * self.npix)
        self._chk(lib().sanepic_learn_get_pair(self.h, k, _dp(r), _dp(e)), "learn_get_pair")
        return r, e

    def train(self, steps, lr=1e-3, lr_final=1e-5, seed=0, first=0, count=0, beta1=0.9,
              beta2=0.999, eps=1e-8, weight_decay=0.0, loss=LOSS_ANORM, sched_offset=0, sched_total=0):
        o = TrainOpts(steps, lr, lr_final, beta1, beta2, eps, weight_decay, seed, first, count, loss,
                      sched_offset, sched_total)
        hist = np.empty(steps, dtype=np.float32)
        self._chk(lib().sanepic_learn_train(self.h, ctypes.byref(o), _fp(hist)), "learn_train")
        return hist

    def eval_pair(self, k, loss=LOSS_ANORM):
        out = ctypes.c_double()
        self._chk(lib().sanepic_learn_eval(self.h, k, loss, ctypes.byref(out)), "learn_eval")
        return out.value

    def grad_pair(self, k, loss=LOSS_RESIDUAL):
        """(loss, gradient w.r.t. the attached network's parameters) of pair k."""
        n = ctypes.c_int64()
        self._chk(lib().sanepic_unet_nparams(self._unet.h, ctypes.byref(n)), "unet_nparams")
        g = np.empty(n.value, dtype=np.float32)
        out = ctypes.c_double()
        self._chk(lib().sanepic_learn_grad(self.h, k, loss, ctypes.byref(out), _fp(g)), "learn_grad")
        return out.value, g

    def learn_stats(self):
        a, b, c = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
        self._chk(lib().sanepic_learn_stats(self.h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c)),
                  "learn_stats")
        return dict(learned=a.value, fallback=b.value, rejected=c.value)
