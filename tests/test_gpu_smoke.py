"""The smoke entry point (one small solve on device 0 checked
against the CPU oracle) must pass as committed."""
import pytest

pytestmark = pytest.mark.gpu


def test_graft_entry_smoke():
    import __graft_entry__
    __graft_entry__.smoke()
