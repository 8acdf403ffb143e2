"""The C-ABI library loads and exports every entry point include/sanepic_hip.h
declares; argument validation runs before any device call."""
import ctypes
import os
import re

import pytest

import sanepic_amd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    txt = open(os.path.join(ROOT, "include", "sanepic_hip.h")).read()
    return sorted(set(re.findall(r"\b(sanepic_[a-z0-9_]+)\s*\(", txt)))


def test_header_matches_binding_list():
    assert header_functions() == sorted(sanepic_amd.ABI_FUNCTIONS)


def test_library_exports_all_symbols():
    L = ctypes.CDLL(sanepic_amd.LIB_PATH)
    for name in header_functions():
        assert hasattr(L, name), name
    assert sanepic_amd.lib().sanepic_abi_version() == 1


@pytest.mark.parametrize("bad", [
    dict(nside=0), dict(nside=8193), dict(ndet=0), dict(nseg=0), dict(ns=1000), dict(ns=8),
    dict(polar=0), dict(precision=7),
])
def test_create_rejects_bad_config(bad):
    cfg = dict(nside=8, ndet=1, nseg=2, ns=1024, polar=1, hwp=0, omega_hwp=0.0, precision=0, device=0)
    cfg.update(bad)
    c = sanepic_amd.Config(**cfg)
    h = ctypes.c_void_p()
    rc = sanepic_amd.lib().sanepic_create(ctypes.byref(c), ctypes.byref(h))
    assert rc == 1
    assert h.value is None
    assert len(sanepic_amd.lib().sanepic_last_error(None)) > 0
