"""Turns the output of the reference-leaf driver (oracle/_ref/golden.txt, made
by `make -C oracle/ref` from Hochunk3d's own sources in the build container)
into the committed fixture tests/golden/ref_leaves.npz.

Each section of golden.txt is one routine of the reference: its inputs (or the
formula-defined tables it read) and its outputs, with 17 significant digits.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "..", "..", "oracle", "_ref", "golden.txt")
DST = os.path.join(HERE, "ref_leaves.npz")


def parse(path):
    sec = {}
    with open(path) as f:
        for line in f:
            parts = line.split()
            if not parts:
                continue
            try:
                vals = [float(x) for x in parts[1:]]
            except ValueError:  # the reference's own progress messages
                continue
            sec.setdefault(parts[0], []).append(vals)
    return {k: np.array(v, dtype=np.float64) for k, v in sec.items()}


def main():
    src = sys.argv[1] if len(sys.argv) > 1 else SRC
    data = parse(src)
    np.savez_compressed(DST, **data)
    for k, v in data.items():
        print(f"{k:12s} {v.shape}")


if __name__ == "__main__":
    main()
