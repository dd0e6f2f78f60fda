"""The entry point's smoke run (one forward + render + backward through the
public API on cuda:0, checked against the oracle) as a test."""
import pytest

pytestmark = pytest.mark.gpu


def test_graft_entry_smoke(device):
    import __graft_entry__
    __graft_entry__.smoke()
