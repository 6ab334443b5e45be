"""Bjorkman-Wood re-emission (ilucy = 0) on the GPU.

- Serial mode (one lane runs the packets in their order, the reference's
  loop): every packet's trace, the absorption counts and the temperatures the
  pass leaves match the oracle's Bjorkman-Wood pass bit for bit.
- Parallel mode: packets in flight take their absorption counts in the order
  of the atomics, so histories differ from the serial order; with many more
  packets than lanes in flight the temperatures agree with the Lucy
  temperatures of the same model (the reference's own consistency check of
  the two methods, the tmidplane columns).
"""
import numpy as np
import pytest

from hochunk3d_amd import build_model, config1
from hochunk3d_amd.testing import compare_traces

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def model():
    m = build_model(config1(nr=24, ntheta=24, nphi=12, nximg=21, nfreq=50))
    m.tdust[...] = 0.1 / 11605.0
    return m


def _mass_weighted_rel(m, a, b, nabs, nmin):
    sel = (nabs >= nmin) & (b > 0)
    w = (m.densarr * m.vcell[..., None])[sel]
    rel = np.abs(a[sel] / b[sel] - 1.0)
    return float(np.sum(w * rel) / np.sum(w)), int(sel.sum())


def test_serial_bw_pass_matches_oracle(model):
    from oracle.pyoracle import Oracle

    from hochunk3d_amd._lib import Transport
    n = 600
    g = Transport(model)
    try:
        g.set_bw(True)
        g.set_option("serial", 1)
        rg = g.run(n, ipeel=0, trace=True)
        tg = g.get_temperature()
    finally:
        g.close()
    o = Oracle(model, bw=True)
    rc = o.run(n, ipeel=0, trace=True)
    compare_traces(rg.trace, rc["trace"])
    np.testing.assert_array_equal(rg.nabs, rc["nabs"])
    act = model.densarr > 0.0
    np.testing.assert_array_equal(tg[act], o.tdust[act])
    assert rc["nabs"].sum() > 1000


def test_parallel_bw_temperatures_match_lucy(model):
    """The parallel pass's temperatures against Lucy iterations of the same
    model (the reference's consistency check of the two methods).  Packets in
    flight see each other's counts, so a pass whose packets are all in flight
    at once runs hot (+5 % on this model at 4e4 packets); with many more
    packets than lanes the pass tends to the equilibrium (+0.8 % at 4e5,
    -0.1 % at 4e6 in the measurement that set these bounds)."""
    from hochunk3d_amd._lib import Transport
    n = 4_000_000
    g = Transport(model)
    try:
        g.set_bw(True)
        rb = g.run(n, ipeel=0)
        tb = g.get_temperature()
        g.set_bw(False)
        g.set_temperature(model.tdust)
        nl = 400_000
        for i in range(8):
            g.run_scalars(nl, seg0=(i + 2) * n)
            g.temperature(nl, apply=True)
        tl = g.get_temperature()
    finally:
        g.close()
    s = rb.scalars
    assert s["n_err_nowall"] == 0 and s["n_warn"] == 0 and s["nabs_ev"] > 1e8
    # cell-species with at least 1000 absorptions, whose temperature the
    # count fixes to ~1/(4 sqrt(1000)) = 0.8 %
    dusty = (model.densarr > 0.0) & (rb.nabs >= 1000) & (tl > 0.0)
    w = (model.densarr * model.vcell[..., None])[dusty]
    rel = tb[dusty] / tl[dusty] - 1.0
    mean = float(np.sum(w * rel) / np.sum(w))
    mabs = float(np.sum(w * np.abs(rel)) / np.sum(w))
    print(f"Bjorkman-Wood (parallel, {n} packets) vs Lucy: mass-weighted mean {mean:+.4f}, "
          f"mean |rel| {mabs:.4f} over {int(dusty.sum())} cell-species")
    assert abs(mean) < 0.02 and mabs < 0.04, (mean, mabs)
