"""The C-ABI library loads and exports every entry point include/tsph.h declares.

CPU only: no compute call is made (there may be no GPU here)."""
import ctypes
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
LIB = ROOT / "tensorsph_amd" / "libtsph.so"


def declared():
    text = (ROOT / "include" / "tsph.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(tsph_[a-z0-9_]+)\s*\(", text)))


@pytest.mark.skipif(not LIB.exists(), reason="libtsph.so not built (run __graft_entry__.build())")
def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(str(LIB))
    names = declared()
    assert len(names) >= 10
    missing = [name for name in names if not hasattr(lib, name)]
    assert not missing, missing
    lib.tsph_abi_version.restype = ctypes.c_int
    assert lib.tsph_abi_version() == 1


def test_header_declares_the_core_entry_points():
    names = declared()
    for core in ("tsph_pairs_search", "tsph_pairs_export", "tsph_wcsph_rhs",
                 "tsph_nlist_from_pairs", "tsph_last_error"):
        assert core in names
