"""Writes the synthetic dust table in the reference's 7-column format with 17
significant digits, for the golden-vector driver."""
import os
import sys

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "..", ".."))
from hochunk3d_amd.dust import synthetic_dust_table  # noqa: E402

tab = synthetic_dust_table()
with open(sys.argv[1], "w") as f:
    for row in tab:
        f.write(" ".join(f"{v:.17e}" for v in row) + "\n")
