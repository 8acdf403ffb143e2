"""Segment lengths 3 * 2^l and 5 * 2^l (the odd radix rides in the four-step
columns) against the CPU oracle, whose FFT handles radices 2, 3, 4 and 5
(pinned against numpy in test_oracle_fft.py)."""
import numpy as np
import pytest

import sanepic_amd
from tests.helpers import make_case, oracle_problem, rel

pytestmark = pytest.mark.gpu

CASES = [dict(nside=32, ndet=2, nseg=3, ns=3 * 2 ** 13), dict(nside=32, ndet=1, nseg=3, ns=5 * 2 ** 13),
         dict(nside=32, ndet=2, nseg=2, ns=3 * 2 ** 15)]


@pytest.mark.parametrize("case", CASES, ids=lambda c: "ns{ns}_d{ndet}".format(**c))
def test_mixed_radix_operator_matches_oracle(case):
    print("case", case, flush=True)
    cfg, ds = make_case(**case)
    P = oracle_problem(cfg, ds)
    x = np.random.default_rng(11).standard_normal(3 * P.npix)
    ax_ref, b_ref = P.apply_A(x), P.rhs()
    for prec, tol in ((sanepic_amd.FP64, 1e-11), (sanepic_amd.MIXED, 2e-6)):
        ctx = sanepic_amd.Context(cfg.nside, cfg.ndet, cfg.nseg, cfg.ns, precision=prec)
        ctx.load(cfg, ds)
        npix, _ = ctx.pixelize()
        assert npix == P.npix
        ctx.setup()
        e_op = rel(ctx.apply(x), ax_ref)
        e_b = rel(ctx.rhs(), b_ref)
        print(f"  precision {prec}: A x {e_op:.2e}, b {e_b:.2e}", flush=True)
        assert e_op < tol and e_b < 1e-11
        if prec == sanepic_amd.FP64:
            c_or = P.chi2(x)
            assert abs(ctx.chi2(x) - c_or) / c_or < 1e-10
        ctx.close()


def test_mixed_radix_every_split():
    case = CASES[2]  # 3 * 2^15: N1 = 2^9 .. 2^11 keeps the columns 3 * 16 .. 3 * 64
    cfg, ds = make_case(**case)
    P = oracle_problem(cfg, ds)
    x = np.random.default_rng(12).standard_normal(3 * P.npix)
    ref = P.apply_A(x)
    for split in (9, 10, 11):
        ctx = sanepic_amd.Context(cfg.nside, cfg.ndet, cfg.nseg, cfg.ns)
        ctx.load(cfg, ds)
        ctx.pixelize()
        ctx.set_fft_split(split)
        ctx.setup()
        assert rel(ctx.apply(x), ref) < 1e-11
        ctx.close()


def test_bad_lengths_rejected():
    for ns in (7 * 2 ** 13, 3 * 2 ** 12, 15 * 2 ** 13, 1000):
        with pytest.raises(sanepic_amd.SanepicError):
            sanepic_amd.Context(32, 1, 2, ns)
