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


@pytest.mark.skipif(not os.path.exists(_lib.library_path()), reason="library not built")
def test_argument_errors_are_reported_before_any_device_call():
    """Entry points reject bad arguments with HC_E_ARG before touching a
    device, so these run without a GPU."""
    import numpy as np
    lib = _lib.load()
    E_ARG = -1
    assert lib.hc_create(None, 0, None) == E_ARG
    assert lib.hc_set_option(None, b"regen_lanes", 8) == E_ARG
    assert lib.hc_run_iteration(None, 10, 0, 1, 2, 1, 0, None, None) == E_ARG
    assert lib.hc_build_density(None, 0, None, None, None) == E_ARG
    # emissivity: more than 64 quadrature nodes
    a = np.ones(8)
    p = _lib._ptr(a)
    out = np.zeros(64)
    assert lib.hc_build_emissivity(0, 8, p, p, p, 2, p, 4, p, 65, p, p, _lib._ptr(out),
                                   _lib._ptr(out)) == E_ARG
    # clump points: n^(levels+1) above 2^31, an even PCG increment, too many levels
    w = np.linspace(0.0, 1.0, 5)
    pw = _lib._ptr(w)
    cnt = np.zeros(125, np.uint32)
    pc = cnt.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))
    sc = np.full(8, 0.5)
    ok = np.array([1, 2, 3, 5], np.uint64)
    even = np.array([1, 2, 4, 5], np.uint64)

    def frac(n, levels, pcg):
        return lib.hc_fractal_counts(0, 5, 5, 5, pw, pw, pw, pw, n, levels, _lib._ptr(sc), 1.0,
                                     pcg.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)), pc)
    assert frac(64, 5, ok) == E_ARG          # 64^6 = 2^36 points
    assert frac(4, 9, ok) == E_ARG           # levels <= 8
    assert frac(4, 2, even) == E_ARG


def test_filter_and_diffusion_entry_points_reject_a_null_context():
    lib = _lib.load()
    E_ARG = -1
    assert lib.hc_set_filters(None, 0, 0, None, None, None) == E_ARG
    assert lib.hc_get_filter_images(None, None, None, None, None) == E_ARG
    assert lib.hc_set_diffusion(None, None, None, None, None) == E_ARG
