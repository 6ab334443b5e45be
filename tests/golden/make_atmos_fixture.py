"""Writes tests/golden/kt04000g+4.0z-2.0.npz: the rows (wavelength um,
F_nu) of the reference's T = 4000 K model atmosphere kt04000g+4.0z-2.0.ascii,
the one its config-1-like models name, as data for the atmosphere tests (the
reference tree does not travel to the GPU box).  Run in the build container:

    python tests/golden/make_atmos_fixture.py [/root/reference]
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from hochunk3d_amd.atmos import read_atmosphere  # noqa: E402

ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
name = "kt04000g+4.0z-2.0"
wave, flux = read_atmosphere(os.path.join(ref, name + ".ascii"))
np.savez_compressed(os.path.join(HERE, name + ".npz"), wave=wave, flux=flux)
print(f"{name}: {wave.size} rows, {wave[0]}..{wave[-1]} um")
