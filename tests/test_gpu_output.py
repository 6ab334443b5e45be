"""The output files written from a GPU pass against the same files written
from the oracle's pass on the same packets: every file of output() plus the
image cube.  The tallies agree to 1e-9 (atomics sum in another order), so
the printed five-digit values agree except where a value sits within that
of a rounding tie; every line must parse to the same numbers within the
last printed digit, and all but a handful must be identical bytes."""
import os
from types import SimpleNamespace

import numpy as np
import pytest

from hochunk3d_amd import build_model, config1
from hochunk3d_amd.output import read_unf_records, write_run_outputs
from hochunk3d_amd.testing import synthetic_temperature

pytestmark = pytest.mark.gpu


def test_output_files_of_a_gpu_pass_match_the_oracle(tmp_path):
    from hochunk3d_amd._lib import Transport
    from oracle.pyoracle import Oracle
    m = build_model(config1(nr=32, ntheta=33, nphi=8, nximg=21, nfreq=50, npeel=3))
    m.tdust[...] = synthetic_temperature(m)
    n = 3000
    g = Transport(m, device=0)
    try:
        rg = g.run(n, ipeel=1)
        t2, _ = g.temperature(n)
    finally:
        g.close()
    rc = SimpleNamespace(**Oracle(m).run(n, ipeel=1))
    dg, dc = tmp_path / "gpu", tmp_path / "oracle"
    pg = write_run_outputs(str(dg), m, rg, n, m.tdust, t2)
    pc = write_run_outputs(str(dc), m, rc, n, m.tdust, t2)
    names = sorted(os.path.basename(p) for p in pg)
    assert names == sorted(os.path.basename(p) for p in pc)
    nlines = ndiff = 0
    for name in names:
        a, b = dg / name, dc / name
        if name.endswith(".unf"):
            ra, rb = read_unf_records(str(a)), read_unf_records(str(b))
            assert [len(x) for x in ra] == [len(x) for x in rb], name
            assert ra[:4] == rb[:4], name
            dt = "<f4" if name == "imcube.unf" else "<f8"
            for xa, xb in zip(ra[1 if name == "imcube.unf" else 4:], rb[1 if name == "imcube.unf" else 4:]):
                va, vb = np.frombuffer(xa, dt), np.frombuffer(xb, dt)
                np.testing.assert_allclose(va, vb, rtol=1e-6, atol=1e-30, err_msg=name)
            continue
        la, lb = a.read_text().splitlines(), b.read_text().splitlines()
        assert len(la) == len(lb), name
        for x, y in zip(la, lb):
            nlines += 1
            if x == y:
                continue
            ndiff += 1
            fx = np.array([float(v) for v in x.split()])
            fy = np.array([float(v) for v in y.split()])
            # one unit in the fifth significant digit
            np.testing.assert_allclose(fx, fy, rtol=2e-4, atol=1e-30, err_msg=f"{name}: {x!r} {y!r}")
    print(f"{len(names)} files, {nlines} lines, {ndiff} differ in the last digit")
    assert ndiff <= max(3, nlines // 2000)
