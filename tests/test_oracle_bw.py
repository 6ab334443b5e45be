"""The oracle's Bjorkman-Wood re-emission (ilucy = 0) on the CPU.

dustfreq and dusttemp themselves are pinned bit for bit against the
reference's compiled dust.f90/dusttemp.f90 (tests/test_oracle_golden.py).
Here the composition in the pass (emit_common.f90:57-70): each absorption
updates its cell-species temperature to dusttemp of the absorption count so
far, so after the pass every absorbing cell-species sits at the fixed point
of dusttemp for its final count, cells without absorptions keep their
temperature, and a count off by one (a plausible mistake) would leave the
cells off that fixed point."""
import ctypes as C

import numpy as np
import pytest

from hochunk3d_amd import build_model, config1
from oracle.pyoracle import Oracle


@pytest.fixture(scope="module")
def bw_pass():
    m = build_model(config1(nr=24, ntheta=24, nphi=12, nximg=21, nfreq=50))
    m.tdust[...] = 0.1 / 11605.0
    o = Oracle(m, bw=True)
    t0 = o.tdust.copy()
    n = 2000
    r = o.run(n, ipeel=0, trace=True)
    return m, o, t0, r, n


def test_temperatures_follow_absorption_counts(bw_pass):
    m, o, t0, r, n = bw_pass
    T = o.tdust
    nabs = r["nabs"]
    assert nabs.sum() == r["trace"]["nabs"].sum() > 0
    assert np.array_equal(T[nabs == 0], t0[nabs == 0])
    assert np.all(T[nabs > 0] > t0[nabs > 0])
    lib, d = o.lib, C.byref(o.dust)
    idx = np.argwhere(nabs > 0)
    rng = np.random.default_rng(3)
    pick = idx[rng.choice(len(idx), size=min(600, len(idx)), replace=False)]
    off = 0
    for i in map(tuple, pick):
        na, t, m4, sp = nabs[i], T[i], m.massarr[i], i[3]
        te = lib.hco_dusttemp(d, na, t, m.ltot, float(n), m4, sp)
        assert abs(te - t) <= 2e-5 * t, (i, na, t, te)
        # one absorption fewer is off the fixed point where counts are small
        if na <= 20:
            tm = lib.hco_dusttemp(d, na - 1.0 if na > 1 else 0.5, t, m.ltot, float(n), m4, sp)
            off += abs(tm - t) > 2e-5 * t
    assert off > 0


def test_lucy_mode_leaves_temperatures(bw_pass):
    m, _, _, _, _ = bw_pass
    o = Oracle(m)
    t0 = o.tdust.copy()
    o.run(200, ipeel=0)
    assert np.array_equal(o.tdust, t0)
