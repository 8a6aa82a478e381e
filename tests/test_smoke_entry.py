"""The round-trip check run by __graft_entry__.smoke(): cholesky, the two
triangular solves and the projected inverse of a 24 x 24 grid on cuda:0
against the oracle."""
import os
import sys

import pytest

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))


@pytest.mark.gpu
def test_smoke_entry_point():
    import __graft_entry__ as g
    g.smoke()
