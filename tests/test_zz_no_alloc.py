"""sanepic_apply_noalloc (no device allocation; staging in the library's
static per-GPU area) against sanepic_apply and the oracle, on two calls in
a row, the second on a context with fewer seen pixels (the area keeps the
first call's contents)."""
import numpy as np
import pytest

import sanepic_amd
from tests.helpers import make_case, oracle_problem, rel

pytestmark = pytest.mark.gpu


def _ctx(cfg, ds, precision):
    ctx = sanepic_amd.Context(cfg.nside, cfg.ndet, cfg.nseg, cfg.ns, precision=precision)
    ctx.load(cfg, ds)
    ctx.pixelize()
    ctx.setup()
    return ctx


@pytest.mark.parametrize("precision", [sanepic_amd.FP64, sanepic_amd.MIXED])
def test_apply_noalloc_matches_apply(precision):
    cases = [make_case(nside=32, ndet=2, nseg=2, ns=2 ** 14), make_case(nside=8, ndet=1, nseg=3, ns=1024)]
    ctxs = [_ctx(cfg, ds, precision) for cfg, ds in cases]
    assert ctxs[1].npix < ctxs[0].npix
    rng = np.random.default_rng(21)
    xs = [rng.standard_normal(3 * c.npix) for c in ctxs]
    want = [c.apply(x) for c, x in zip(ctxs, xs)]
    got = [c.apply_noalloc(x) for c, x in zip(ctxs, xs)]  # two calls in a row, the second smaller
    for (cfg, ds), c, x, w, g in zip(cases, ctxs, xs, want, got):
        print(f"  npix {c.npix}: noalloc vs apply identical {np.array_equal(g, w)}", flush=True)
        assert np.array_equal(g, w)
        e = rel(g, oracle_problem(cfg, ds).apply_A(x))
        assert e < (1e-11 if precision == sanepic_amd.FP64 else 2e-6), e
    for c in ctxs:
        c.close()
