"""The build's smoke check (one ADER step and average down of the HIP path
on a two-level vortex, against the oracle to 1e-12) as a GPU test."""
import os
import sys

import pytest

pytestmark = pytest.mark.gpu


def test_smoke_entry_point():
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    sys.path.insert(0, root)
    import __graft_entry__
    __graft_entry__.smoke()
