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
