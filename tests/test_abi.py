"""The C-ABI library loads and exports every symbol include/chompack_amd.h
declares (no device calls)."""
import re
from pathlib import Path

from chompack_amd import _native


def test_exports_match_header():
    hdr = (Path(__file__).parents[1] / "include" / "chompack_amd.h").read_text()
    names = set(re.findall(r"\b(chp_[a-z_0-9]+)\s*\(", hdr))
    assert "chp_cholesky" in names and "chp_plan_create" in names
    lib = _native.lib()
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing


def test_version():
    assert _native.lib().chp_version() >= 1
