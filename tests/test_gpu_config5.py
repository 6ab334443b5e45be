"""The large hydro-like 3-D model (config 5) on the GPU.

* Reduced size (30 x 31 x 16 cells, densities through densarr.unf, the
  log-normal multiplier on disks and envelope): Lucy and peeled passes
  packet by packet against the oracle; the FP32 image cube (the reference's
  real*4 image arrays, hc_dims.imcube = 2) against the oracle's FP64 cube
  to FP32 summation rounding.
* Full size (400 x 300 x 256 cells, 20 observers, 500^2 pixels, 200
  frequencies: a 44 GB FP32 IQUV cube resident in HBM), FP32 path: every
  packet of a Lucy pass leaves with its energy or is counted (escaped +
  killed = packets, no lost or dropped packet), and in a peeled pass whose
  image covers the grid each observer's cube holds exactly the energy of
  its peeled SED, in I, Q, U and V."""
import numpy as np
import pytest

from hochunk3d_amd import build_model, config5
from hochunk3d_amd.testing import compare_traces, rel_close, synthetic_temperature

pytestmark = pytest.mark.gpu


def test_reduced_config5_through_densarr_unf_matches_oracle(tmp_path):
    from hochunk3d_amd._lib import Transport
    from hochunk3d_amd.hydro import write_densarr_unf
    from oracle.pyoracle import Oracle
    cfg = dict(nr=30, ntheta=31, nphi=16, nximg=41, nfreq=50, npeel=5)
    src = build_model(config5(**cfg))
    path = tmp_path / "densarr.unf"
    write_densarr_unf(str(path), src.rarr, src.thetarr, src.phiarr, src.densarr[..., :2])
    m = build_model(config5(**cfg, densinput=str(path),
                            lognormal={"sigma": 1.5, "slope": -11.0 / 3.0, "seed": 11,
                                       "species": (2,)}))
    m.tdust[...] = synthetic_temperature(m)
    print("reduced config5: lucy 5000, peeled 400 packets", flush=True)
    g = Transport(m, cube_fp32=True)
    try:
        rg = g.run(5000, ipeel=0, trace=True)
        rp = g.run(400, ipeel=1, trace=True, seg0=90_000)
    finally:
        g.close()
    c = Oracle(m)
    rc = c.run(5000, ipeel=0, trace=True)
    compare_traces(rg.trace, rc["trace"])
    rel_close(rg.dtauabs, rc["dtauabs"], rtol=1e-9, atol=1e-300)
    rel_close(rg.sed, rc["sed"], rtol=1e-9, atol=1e-12)
    cp = c.run(400, ipeel=1, trace=True, seg0=90_000)
    compare_traces(rp.trace, cp["trace"])
    rel_close(rp.peel, cp["peel"], rtol=1e-9, atol=1e-12)
    assert rp.img_i.dtype == np.float32
    for k in ("img_i", "img_q", "img_u", "img_v"):
        a, b = getattr(rp, k).astype(np.float64), cp[k]
        scale = np.abs(b).max()
        if k == "img_v" and scale == 0.0:   # no circular polarisation from these events
            assert np.all(a == 0.0)
            continue
        assert scale > 0, k
        np.testing.assert_allclose(a, b, rtol=2e-5, atol=2e-6 * scale, err_msg=k)


def test_full_config5_energy_bookkeeping():
    import time

    from hochunk3d_amd._lib import HC_FP32, Transport
    t0 = time.time()
    m = build_model(config5(rmaxi_au=5000.0), device=0)
    m.tdust[...] = synthetic_temperature(m)
    print(f"config5 {m.nrg - 1}x{m.ntg - 1}x{m.npg - 1} cells built in {time.time() - t0:.1f}s",
          flush=True)
    g = Transport(m, precision=HC_FP32, cube_fp32=True)
    try:
        n = 2_000_000
        s, ms = g.run_scalars(n)
        print(f"lucy pass {n} packets: {ms:.0f} ms, longest {s['max_packet_work']:.0f}",
              flush=True)
        assert s["n_err_nowall"] == 0 and s["n_err_dust"] == 0 and s["n_warn"] == 0
        # every packet escapes with weight 1 (FP32 Stokes I) or is killed
        assert abs(s["sumall"] + s["abflux"] - n) <= 1e-5 * n
        npl = 20_000
        r = g.run(npl, ipeel=1, seg0=n)
        print(f"peeled pass {npl} packets: {r.kernel_ms:.0f} ms", flush=True)
    finally:
        g.close()
    sc = r.scalars
    assert sc["n_err_nowall"] == 0 and sc["n_warn"] == 0
    assert r.img_i.shape == (20, 500, 500, 200, 8) and r.img_i.dtype == np.float32
    big = r.peel.shape[3] - 1                       # the largest aperture
    # per observer, in one pass over each (F-order) cube
    tots = [r.img_i[..., 0].sum(axis=(1, 2, 3), dtype=np.float64)] + [
        getattr(r, name).sum(axis=(1, 2, 3), dtype=np.float64)
        for name in ("img_q", "img_u", "img_v")]
    ref = r.peel[0, :, :, big, 0].sum(axis=0)
    assert np.all(ref > 0)
    for j in range(4):
        sed = r.peel[j, :, :, big, 0].sum(axis=0)
        np.testing.assert_allclose(tots[j], sed, rtol=0, atol=1e-4 * ref.max())
