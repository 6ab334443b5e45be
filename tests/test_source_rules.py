"""Rules of the device code that a compiler does not check.

The transport kernel reads its launch parameters from the kernel-argument
segment (kernel_params(), spelled c_P).  A device function emitted as a call
does not receive that segment's address -- the builtin yields a null pointer
there -- so a function marked noinline must not read the parameters; it
takes them as arguments instead.
"""
import os
import re

CSRC = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                    "hochunk3d_amd", "csrc")


def _bodies_of_noinline(text: str):
    for m in re.finditer(r"noinline\)\)[^;{]*?([A-Za-z_][A-Za-z_0-9]*)\s*\(", text):
        name = m.group(1)
        i = text.index("{", m.end())
        depth, j = 0, i
        while True:
            if text[j] == "{":
                depth += 1
            elif text[j] == "}":
                depth -= 1
                if depth == 0:
                    break
            j += 1
        yield name, text[i:j + 1]


def test_noinline_functions_do_not_read_launch_parameters():
    found = []
    for fn in sorted(os.listdir(CSRC)):
        if not fn.endswith((".hip", ".h")):
            continue
        text = open(os.path.join(CSRC, fn)).read()
        for name, body in _bodies_of_noinline(text):
            found.append(name)
            assert "c_P" not in body and "kernel_params" not in body, (fn, name)
    assert "axis_phi_cell_call" in found


def test_rule_checker_catches_a_planted_read():
    planted = ("__device__ __attribute__((noinline)) int f(int x) {\n"
               "  if (x) { return c_P.npg; }\n  return 0;\n}\n")
    (name, body), = list(_bodies_of_noinline(planted))
    assert name == "f" and "c_P" in body


LLVM = "/opt/rocm/llvm/bin"


def _device_functions(lib_path: str, tmp):
    """(kernels, other functions) of the library's gfx950 code objects:
    the .hip_fatbin section split into its offload bundles, each bundle's
    device ELF unbundled, and the FUNC symbols of its symbol table (a kernel
    has a matching '<name>.kd' descriptor symbol).  Reads symbol tables only."""
    import subprocess
    fat = tmp / "fatbin.bin"
    subprocess.run(["objcopy", f"--dump-section=.hip_fatbin={fat}", lib_path, str(tmp / "x.so")],
                   check=True, capture_output=True)
    data = fat.read_bytes()
    magic = b"__CLANG_OFFLOAD_BUNDLE__"
    starts = [i for i in range(len(data)) if data.startswith(magic, i)]
    kernels, others = set(), set()
    for n, (a, b) in enumerate(zip(starts, starts[1:] + [len(data)])):
        part = tmp / f"bundle{n}.bin"
        part.write_bytes(data[a:b])
        dev = tmp / f"dev{n}.o"
        subprocess.run([f"{LLVM}/clang-offload-bundler", "--type=o", "--unbundle",
                        f"--input={part}", "--targets=hipv4-amdgcn-amd-amdhsa--gfx950",
                        f"--output={dev}"], check=True, capture_output=True)
        out = subprocess.run([f"{LLVM}/llvm-readelf", "--symbols", str(dev)], check=True,
                             capture_output=True, text=True).stdout
        funcs, objs = set(), set()
        for line in out.splitlines():
            f = line.split()
            if len(f) >= 8 and f[3] == "FUNC":
                funcs.add(f[7])
            elif len(f) >= 8 and f[3] == "OBJECT":
                objs.add(f[7])
        for fn in funcs:
            (kernels if fn + ".kd" in objs else others).add(fn)
    return kernels, others


def test_device_code_has_no_outlined_helpers(tmp_path):
    """Every device function of the build is inlined into its kernel except
    the polar-axis phi cell (which takes its inputs as arguments): a helper
    or event-loop lambda the compiler emitted as a call would read a null
    kernel-argument pointer through kernel_params()."""
    import pytest

    from hochunk3d_amd import _lib
    lib = _lib.library_path()
    if not (os.path.exists(lib) and os.path.exists(f"{LLVM}/llvm-readelf")):
        pytest.skip("library or LLVM tools absent")
    kernels, others = _device_functions(lib, tmp_path)
    assert any("transport" in k for k in kernels)
    assert any("densities" in k for k in kernels)
    assert {o for o in others if "axis_phi_cell_call" not in o} == set(), others
