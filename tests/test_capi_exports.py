"""The C-ABI library loads and exports every symbol include/channel_flow.h
declares (no GPU calls)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols():
    txt = open(os.path.join(ROOT, "include", "channel_flow.h")).read()
    return sorted(set(re.findall(r"\b(cf_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_all_symbols():
    from channel_flow import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.fail("libcfhip.so missing; run channel-flow_amd/build.sh")
    lib = ctypes.CDLL(_lib.LIB_PATH)
    syms = header_symbols()
    assert len(syms) >= 25
    missing = [s for s in syms if not hasattr(lib, s)]
    assert not missing, missing
    # the ctypes signature table covers the header too
    assert set(syms) <= set(_lib.SIGNATURES), set(syms) - set(_lib.SIGNATURES)


def test_missing_library_fails_loudly(tmp_path):
    from channel_flow import _lib
    with pytest.raises(RuntimeError):
        _lib.load(str(tmp_path / "nope.so"))
