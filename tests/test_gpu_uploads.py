"""Queued uploads (pinned buffers, per-block asynchronous copies on several
upload streams, the set-up's TOD-free half queued before the wait for the
TOD) against synchronous ones: the same kernels on the same data, so b, the
operator and chi2 must equal the synchronous path's bit for bit, NaN samples
included."""
import numpy as np
import pytest

import sanepic_amd
from tests.helpers import make_case

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_queued_uploads_are_bit_identical(prec):
    P = sanepic_amd.FP64 if prec == "fp64" else sanepic_amd.MIXED
    # N = 2^16 (four-step path), 5 segments (the last pair holds one segment)
    cfg, ds = make_case(nside=32, ndet=3, nseg=5, ns=2 ** 16, fknee=0.2)
    ns = cfg.ns
    tod = ds["tod"]
    tod[1, 2 * ns + 100:2 * ns + 104] = np.nan   # filled from the sample before (784-790)
    tod[2, 4 * ns + 7] = np.nan
    print(f"queued vs synchronous uploads, {prec}: 3 det x 5 x 2^16, NaN samples in two segments", flush=True)
    a = sanepic_amd.Context(cfg.nside, cfg.ndet, cfg.nseg, cfg.ns, precision=P)
    a.load(cfg, ds)
    a.pixelize()
    a.setup()
    pin = sanepic_amd.PinnedInputs(cfg, ds)
    b = sanepic_amd.Context(cfg.nside, cfg.ndet, cfg.nseg, cfg.ns, precision=P)
    b.load_async(cfg, pin)
    b.pixelize()
    b.setup()
    pin.close()
    try:
        assert np.isfinite(a.rhs()).all()
        assert np.array_equal(a.rhs(), b.rhs())
        x = np.random.default_rng(3).standard_normal(3 * a.npix)
        assert np.array_equal(a.apply(x), b.apply(x))
        assert a.chi2(x) == b.chi2(x)
    finally:
        a.close()
        b.close()
