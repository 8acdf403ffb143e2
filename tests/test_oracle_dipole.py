"""Pins the oracle's orbital-dipole subtraction (-o; sanepic.cc:700-720 read
and r2c, 1055-1067 fdatas0 -= fodipole).

Closed form: a TOD that carries the dipole, with the dipole subtracted, gives
the right-hand side and chi2 of the same TOD without it (to rounding), while
leaving it in changes b by far more; NaN samples are filled in the TOD before
the dipole comes off (the dipole itself is not filled)."""
import numpy as np

from tests.helpers import make_case, oracle_problem, rel


def test_dipole_subtraction_recovers_clean_data():
    kw = dict(nside=8, ndet=2, nseg=3, ns=1024)
    cfg_d, ds_d = make_case(dipole=True, **kw)
    cfg_0, ds_0 = make_case(dipole=False, **kw)
    assert np.abs(ds_d["odip"]).max() > 1e-4
    assert np.allclose(ds_d["tod"] - ds_d["odip"], ds_0["tod"], rtol=0, atol=1e-15)
    Pd = oracle_problem(cfg_d, ds_d)
    P0 = oracle_problem(cfg_0, ds_0)
    assert rel(Pd.rhs(), P0.rhs()) < 1e-11
    x = np.random.default_rng(3).standard_normal(3 * P0.npix) * 1e-4
    assert abs(Pd.chi2(x) - P0.chi2(x)) / P0.chi2(x) < 1e-10
    # without the subtraction the dipole dominates b's change
    ds_keep = dict(ds_d, odip=None)
    Pk = oracle_problem(cfg_d, ds_keep)
    assert rel(Pk.rhs(), P0.rhs()) > 1e-3


def test_nan_samples_are_filled_before_the_dipole_comes_off():
    cfg, ds = make_case(nside=8, ndet=1, nseg=2, ns=512, dipole=True)
    tod = ds["tod"].copy()
    tod[0, [5, 6, 700]] = np.nan
    od = np.array(ds["odip"])
    Pd = oracle_problem(cfg, dict(ds, tod=tod))
    filled = tod.copy()
    for t in (5, 6, 700):
        filled[0, t] = filled[0, t - 1]
    Pm = oracle_problem(cfg, dict(ds, tod=filled - od, odip=None))
    assert rel(Pd.rhs(), Pm.rhs()) < 1e-12
