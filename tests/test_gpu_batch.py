"""Parameter-grid batches on the GPU.

- Models of the parameter grid (uneven nr/ntheta/nphi, down to one phi
  cell) trace the oracle's cell sequences packet for packet.
- Models run side by side on their own contexts and streams give the
  results of running them one after the other: the same event counts and
  the same temperatures (up to the order of the floating-point tally sums).
"""
import numpy as np
import pytest

from hochunk3d_amd.batch import parameter_grid, run_batch
from hochunk3d_amd.model import build_model
from hochunk3d_amd.testing import (assert_temperatures_close, compare_traces,
                                   synthetic_temperature)

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("index", [5, 10])   # 160x196x2 and 196x176x1 cells
def test_grid_model_traces_match_oracle(index):
    from oracle.pyoracle import Oracle

    from hochunk3d_amd._lib import Transport
    gm = parameter_grid(12, max_cells=200_000)[index]
    assert gm.cfg.nphi <= 2
    m = build_model(gm.cfg)
    m.tdust[...] = synthetic_temperature(m)
    n = 400
    g = Transport(m)
    try:
        r = g.run(n, trace=True)
    finally:
        g.close()
    c = Oracle(m).run(n, trace=True)
    compare_traces(r.trace, c["trace"])


def test_concurrent_batch_matches_sequential():
    models = parameter_grid(4, seed=11, packet_scale=2e-3, max_cells=150_000)
    seq = run_batch(models, iterations=2, streams=1)
    par = run_batch(models, iterations=2, streams=4)
    assert seq.packets == par.packets == 2 * sum(gm.packets for gm in models)
    for a, b in zip(seq.per_model, par.per_model):
        for i, (sa, sb) in enumerate(zip(a["scalars"], b["scalars"])):
            # the first pass starts from the same temperatures: the same
            # histories; later passes from temperatures equal up to the sum
            # order of the tallies, where a few packets in a million may differ
            for k in ("ncross", "nscat_ev", "nabs_ev", "n_err_nowall", "n_err_dust"):
                if i == 0:
                    assert sa[k] == sb[k], k
                else:
                    assert abs(sa[k] - sb[k]) <= 1e-5 * max(sa[k], 1.0), k
            np.testing.assert_allclose(sb["flux"], sa["flux"], rtol=1e-12 if i == 0 else 1e-5)
        assert_temperatures_close(b["tdust"], a["tdust"], False)
    # the same batch with every model's iterations replayed from a graph
    grf = run_batch(models, iterations=2, streams=4, graph=True)
    for a, b in zip(seq.per_model, grf.per_model):
        # the second pass starts from temperatures that differ in the last
        # bits (tally sum order): a few packets in a million may differ
        for k in ("ncross", "nscat_ev", "nabs_ev"):
            assert abs(a["scalars"][-1][k] - b["scalars"][-1][k]) <= 1e-5 * a["scalars"][-1][k], k
        assert_temperatures_close(b["tdust"], a["tdust"], False)


def test_fused_batch_matches_models_alone():
    """hc_run_batch: the passes of FP64 and FP32 models in one launch per
    kernel variant give each model's own results: the first iteration, from
    the same temperatures, has the same histories (event counts, fluxes and
    SED up to the tally sum order) as the model alone; the second starts
    from temperatures equal up to that order."""
    from hochunk3d_amd._lib import HC_FP32, HC_FP64, Transport, run_batch_lucy
    models = parameter_grid(5, seed=11, packet_scale=2e-3, max_cells=150_000)
    built = [build_model(gm.cfg) for gm in models]
    precs = [HC_FP64, HC_FP32, HC_FP64, HC_FP32, HC_FP64]
    nps = [gm.packets for gm in models]
    alone = []
    for m, p, n in zip(built, precs, nps):
        t = Transport(m, precision=p)
        try:
            s, _, sed = t.run_sed(n, seg0=0)
            t.temperature(n, apply=True)
            s2, _ = t.run_scalars(n, seg0=n)
            t.temperature(n, apply=True)
            alone.append((s, sed, s2, t.get_temperature()))
        finally:
            t.close()
    trs = [Transport(m, precision=p) for m, p in zip(built, precs)]
    try:
        res1, ms1 = run_batch_lucy(trs, 1, nps, sed=True)
        res2, ms2 = run_batch_lucy(trs, 1, nps, seg0=nps)
        temps = [t.get_temperature() for t in trs]
    finally:
        for t in trs:
            t.close()
    assert ms1 > 0 and ms2 > 0
    for (s, sed, s2, temp), (f, _, fsed), (f2, _, _), ft in zip(alone, res1, res2, temps):
        for k in ("ncross", "nscat_ev", "nabs_ev", "tot", "n_err_nowall", "n_err_dust", "n_warn",
                  "max_packet_work"):
            assert f[k] == s[k], k
        np.testing.assert_allclose(f["flux"], s["flux"], rtol=1e-12)
        np.testing.assert_allclose(fsed, sed, rtol=1e-9, atol=1e-12 * np.abs(sed).max())
        for k in ("ncross", "nscat_ev", "nabs_ev"):
            assert abs(f2[k] - s2[k]) <= 1e-5 * max(s2[k], 1.0), k
        assert_temperatures_close(ft, temp, False)


def test_fused_batch_graph_replays():
    """Two iterations replayed from the batch's graph equal two batch calls of
    one iteration each (the segment bases advance on the device)."""
    from hochunk3d_amd._lib import Transport, run_batch_lucy
    models = parameter_grid(3, seed=5, packet_scale=1e-3, max_cells=100_000)
    built = [build_model(gm.cfg) for gm in models]
    nps = [gm.packets for gm in models]
    out = []
    for split in (False, True):
        trs = [Transport(m) for m in built]
        try:
            if split:
                run_batch_lucy(trs, 1, nps)
                res, _ = run_batch_lucy(trs, 1, nps, seg0=nps)
            else:
                res, _ = run_batch_lucy(trs, 2, nps, oversubscribe=2)
            out.append((res, [t.get_temperature() for t in trs]))
        finally:
            for t in trs:
                t.close()
    (ra, ta), (rb, tb) = out
    for a, b, x, y in zip(ra, rb, ta, tb):
        for k in ("ncross", "nscat_ev", "nabs_ev"):
            assert abs(a[0][k] - b[0][k]) <= 1e-5 * max(b[0][k], 1.0), k
        assert_temperatures_close(x, y, False)
