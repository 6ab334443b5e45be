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
