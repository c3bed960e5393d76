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


def test_python_constants_match_header():
    """Every status, flag and result constant the wrapper defines has the
    value of the header's #define of the same name."""
    import re
    from mirage_amd import native
    defs = {m.group(1): int(m.group(2), 0) for m in
            re.finditer(r"#define (MH_[A-Z0-9_]+) \(?(-?(?:0x[0-9a-fA-F]+|\d+))u?\)?", open(HEADER).read())}
    names = [n for n in dir(native) if n.startswith("MH_") and isinstance(getattr(native, n), int)]
    assert len(names) >= 12
    for n in names:
        assert n in defs and defs[n] == getattr(native, n), n
