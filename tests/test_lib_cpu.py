"""CPU-side checks of the HIP C-ABI library: it loads, exports every symbol the
header declares, and the Python binding covers the header (no GPU calls)."""
import re
from pathlib import Path

from pyquake3d_amd import _lib

HEADER = Path(__file__).resolve().parents[1] / "include" / "pyquake3d_amd.h"


def header_functions():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"\b([A-Za-z_][A-Za-z0-9_]*)\s*\(", text)
    skip = {"if", "for", "while", "return", "sizeof", "defined"}
    return [n for n in dict.fromkeys(names) if n not in skip]


def test_header_declares_reference_symbols():
    names = header_functions()
    assert "TDstressFS_C" in names
    assert "TDstressEachSourceAtReceiver_C" in names


def test_library_exports_every_declared_symbol(pq3d_lib):
    missing = [n for n in header_functions() if not hasattr(pq3d_lib, n)]
    assert not missing, f"declared but not exported: {missing}"


def test_binding_covers_header():
    names = set(header_functions())
    assert names == set(_lib.SIGNATURES), (names ^ set(_lib.SIGNATURES))


def test_library_is_gfx950_code_object(pq3d_lib):
    data = _lib.LIB_PATH.read_bytes()
    assert b"gfx950" in data


def test_noalloc_matvec_size_limit_is_checked_before_any_device_call(pq3d_lib):
    """A product needing more than the static scratch (8 Mi doubles) is
    rejected with a clear message before the library touches a device (so
    the check runs here, without a GPU)."""
    import pytest
    args = [None, 1, None, None, 0, None, None, 1, None, None, 4, 16, None, None,
            (8 << 20) // 16, 1, None, None, None, 0, None]
    with pytest.raises(RuntimeError, match="holds at most 8388608"):
        _lib.call("pq3d_hmat_matvec_noalloc", *args)
    args[14], args[15] = -1, 0
    with pytest.raises(RuntimeError, match="must be >= 0"):
        _lib.call("pq3d_hmat_matvec_noalloc", *args)


def test_controller_record_widths(pq3d_lib):
    assert _lib.controller_widths() == {"run": 64, "lstate": 16, "sample": 8}
