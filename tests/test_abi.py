"""The C ABI library exports every entry point include/mirage_hip.h declares
(loaded only; no device calls)."""
import os
import re

import pytest

from mirage_amd import _build, native

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include",
                      "mirage_hip.h")


def declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(mh_[a-z_]+)\s*\(", text)))


def test_header_lists_match():
    assert set(declared()) == set(native.EXPORTS)


@pytest.mark.skipif(not os.path.exists(_build.hip_lib_path()), reason="library not built")
def test_library_exports_every_declared_symbol():
    lib = native.load()
    missing = [s for s in declared() if not hasattr(lib, s)]
    assert not missing, missing
