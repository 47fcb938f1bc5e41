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


@pytest.mark.skipif(not LIB.exists(), reason="libtsph.so not built (run __graft_entry__.build())")
def test_ctypes_structs_match_the_compiled_layout():
    """Every parameter struct the Python binding restates has the compiled
    struct's size, and each of its fields the compiled offset and size."""
    from tensorsph_amd import _lib
    mirror = {"tsph_wcsph_params": _lib.WcsphParams, "tsph_threed_params": _lib.ThreedParams,
              "tsph_dp_params": _lib.DeltaPlusParams, "tsph_batch_caps": _lib.BatchCaps,
              "tsph_dp_case": _lib.DpCase, "tsph_sim_params": _lib.SimParams,
              "tsph_shift_params": _lib.ShiftParams, "tsph_tvc_params": _lib.TvcParams,
              "tsph_gauge_spec": _lib.GaugeSpec}
    lib = ctypes.CDLL(str(LIB))
    fn = lib.tsph_abi_field
    fn.restype = ctypes.c_int
    cp, i64 = ctypes.c_char_p, ctypes.c_int64
    seen = {}
    k = 0
    while True:
        sn, fname = cp(), cp()
        off, size, ssize = i64(), i64(), i64()
        if fn(k, ctypes.byref(sn), ctypes.byref(fname), ctypes.byref(off), ctypes.byref(size),
              ctypes.byref(ssize)) != 0:
            break
        k += 1
        s, f = sn.value.decode(), fname.value.decode()
        cls = mirror[s]
        assert ctypes.sizeof(cls) == ssize.value, s
        desc = getattr(cls, f)
        assert (desc.offset, desc.size) == (off.value, size.value), (s, f)
        seen.setdefault(s, set()).add(f)
    assert set(seen) == set(mirror)
    for s, cls in mirror.items():
        assert seen[s] == {name for name, *_ in cls._fields_}, s
