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


FLANG = "/opt/rocm/lib/llvm/bin/flang"


@pytest.mark.skipif(not (os.path.exists(FLANG) and os.path.exists(_lib.library_path())),
                    reason="flang or the library absent")
def test_fortran_binding_matches_c_layout(tmp_path):
    """hc_iface.f90's bind(C) types have the C struct sizes, and a Fortran
    program links against the library and calls it."""
    import subprocess
    src = os.path.join(ROOT, "hochunk3d_amd", "fortran")
    libdir = os.path.dirname(_lib.library_path())
    exe = tmp_path / "hc_abi_check"
    r = subprocess.run([FLANG, "-J", str(tmp_path), os.path.join(src, "hc_iface.f90"),
                        os.path.join(src, "hc_abi_check.f90"), "-o", str(exe),
                        f"-L{libdir}", "-lhochunk3d_amd", f"-Wl,-rpath,{libdir}"],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60).stdout.split()
    assert out[out.index("abi") + 1] == "1"
    sizes = [int(x) for x in out[out.index("sizes") + 1:out.index("sizes") + 5]]
    assert sizes == [ctypes.sizeof(_lib.HcDims), ctypes.sizeof(_lib.HcSources),
                     ctypes.sizeof(_lib.HcTallies), ctypes.sizeof(_lib.HcDensityModel)]
