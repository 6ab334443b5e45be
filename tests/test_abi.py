"""The C-ABI library loads and exports every symbol include/hc.h declares
(no compute calls: those need a GPU)."""
import ctypes
import os
import re

import pytest

from hochunk3d_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    txt = open(os.path.join(ROOT, "include", "hc.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char \*)\s*(hc_\w+)\s*\(", txt, re.M)))


def test_header_matches_binding():
    assert sorted(_lib.EXPORTS) == _declared()


@pytest.mark.skipif(not os.path.exists(_lib.library_path()), reason="library not built")
def test_library_exports_every_symbol():
    lib = ctypes.CDLL(_lib.library_path())
    for name in _declared():
        assert hasattr(lib, name), name
    lib.hc_abi_version.restype = ctypes.c_int
    assert lib.hc_abi_version() == 1


def test_trace_record_size_matches_header():
    # 13 doubles, 6 int64, 1 uint64, 5 int32 -> 180 bytes padded to 184
    assert _lib.TRACE_DTYPE.itemsize == 184
