"""The diffusion approximation (DIFFUSION=YES) on the GPU.

- Serial mode (one lane runs the packets in their order, the reference's
  loop): every packet's trace, the tallies and the temperatures the pass
  leaves in and behind the layer match the oracle bit for bit, for the
  gridset layer of a massive disk (theta surfaces, above and below the
  midplane), for a radial layer with accretion packets born inside it, and
  for a peeled pass (rays peeled from the layer's surfaces, on a wall).
- Parallel mode: packets in flight heat the surface cells one step per hit
  (compare-and-swap), so each surface cell's T^4 is its start plus an
  integer number of steps 4 pi T*^4 / (nstar dA); the layer counts no
  absorption; the device temperature update restarts the layer at 0.1 K.
"""
import numpy as np
import pytest

from hochunk3d_amd import build_model, config1
from hochunk3d_amd.testing import compare_traces, rel_close, synthetic_temperature

pytestmark = pytest.mark.gpu

PI = 3.1415926535897932384626433
TSUB = 1600.0 / 11605.0


@pytest.fixture(scope="module")
def thick():
    m = build_model(config1(diffusion=True, massd=0.1, nphi=12, nximg=21, nfreq=50))
    m.tdust[...] = synthetic_temperature(m)
    return m


def _radial(accretion):
    m = build_model(config1(massd=0.1, nr=24, ntheta=32, nphi=6, nximg=21, nfreq=50,
                            mdot_disk=1e-6 if accretion else 0.0))
    m.tdust[...] = 300.0 / 11605.0
    d = np.zeros((m.nrg, m.ntg, m.npg), np.int8, order="F")
    dd = np.zeros_like(d)
    d[3:9, 13:19, :-1] = 1
    dd[3:12, 13:19, :-1] = -1
    m.diffus, m.diffdir = d, dd
    return m


def _serial(model, n, ipeel=0, seg0=0):
    from hochunk3d_amd._lib import Transport
    from oracle.pyoracle import Oracle
    g = Transport(model)
    try:
        g.set_option("serial", 1)
        rg = g.run(n, ipeel=ipeel, trace=True, seg0=seg0)
        tg = g.get_temperature()
    finally:
        g.close()
    o = Oracle(model)
    rc = o.run(n, ipeel=ipeel, trace=True, seg0=seg0)
    return rg, tg, rc, o.tdust


def _check(model, rg, tg, rc, tc):
    compare_traces(rg.trace, rc["trace"])
    np.testing.assert_array_equal(rg.nabs, rc["nabs"])
    np.testing.assert_array_equal(rg.dbig, rc["dbig"])
    rel_close(rg.dtauabs, rc["dtauabs"], rtol=1e-9, atol=1e-300)
    rel_close(rg.sed, rc["sed"], rtol=1e-9, atol=1e-12)
    for k in ("n_err_nowall", "n_err_dust"):
        assert rg.scalars[k] == rc["scalars"][k] == 0
    act = model.densarr > 0.0
    act[..., :2] = True                   # the layer writes both disk species everywhere
    np.testing.assert_array_equal(tg[act], tc[act])
    lay = model.diffus == 1
    assert rc["dbig"][lay].sum() > 20     # packets reached the layer


def test_serial_theta_layer_matches_oracle(thick):
    rg, tg, rc, tc = _serial(thick, 400)
    _check(thick, rg, tg, rc, tc)


def test_serial_radial_layer_with_accretion_matches_oracle():
    m = _radial(True)
    assert m.accfrac > 0.05
    rg, tg, rc, tc = _serial(m, 1000, seg0=11)
    _check(m, rg, tg, rc, tc)
    lay = m.diffus == 1
    assert rc["nabs"][lay].sum() > 0      # accretion packets born in the layer


def test_serial_peeled_pass_with_layer_matches_oracle():
    m = _radial(False)
    rg, tg, rc, tc = _serial(m, 150, ipeel=1, seg0=5)
    _check(m, rg, tg, rc, tc)
    rel_close(rg.peel, rc["peel"], rtol=1e-9, atol=1e-12)


def test_parallel_layer_surface_steps_and_reset(thick):
    from hochunk3d_amd._lib import HC_FP32, HC_FP64, Transport
    m = thick
    m0 = m.tdust.copy()
    for prec in (HC_FP64, HC_FP32):
        g = Transport(m, precision=prec)
        try:
            n = 200_000
            r = g.run(n, ipeel=0)
            t = g.get_temperature()
            s = r.scalars
            assert s["n_err_nowall"] == 0 and s["n_err_dust"] == 0 and s["n_warn"] == 0
            lay = m.diffus == 1
            assert r.nabs[lay].sum() == 0.0 and r.dbig[lay].sum() > 100
            # surface cells of the theta layer: T^4 = T0^4 + k c, k an integer
            ns = n                                    # no accretion: every packet stellar
            checked = 0
            for ir in range(1, m.nrg):
                band = np.nonzero(m.diffus[ir - 1, :, 0])[0] + 1
                if band.size == 0:
                    continue
                for surf, wall in ((band[0], band[0]), (band[-1], band[-1] + 1)):
                    dA = PI * np.sin(m.thetarr[wall - 1]) * (m.rarr[ir] ** 2 - m.rarr[ir - 1] ** 2)
                    c = 4.0 * PI * m.tstar_ev ** 4 / (ns * dA)
                    for ip in range(1, m.npg):
                        for id_ in (0, 1):
                            tt = t[ir - 1, surf - 1, ip - 1, id_]
                            t0 = m0[ir - 1, surf - 1, ip - 1, id_]
                            if tt == t0 or tt >= TSUB:
                                continue
                            k = (tt ** 4 - t0 ** 4) / c
                            assert abs(k - round(k)) < 1e-6 * max(1.0, k), (ir, surf, ip, k)
                            checked += 1
            assert checked > 10
            # the device update restarts the layer at 0.1 K (initarr.f90:78-84)
            g.temperature(n, apply=True)
            t2 = g.get_temperature()
            assert np.all(t2[..., :2][np.repeat(lay[..., None], 2, axis=3)] == 0.1 / 11605.0)
        finally:
            g.close()


def test_diffusion_with_bjorkman_wood_is_refused(thick):
    from hochunk3d_amd._lib import HcError, Transport
    g = Transport(thick)
    try:
        g.set_bw(True)
        with pytest.raises(HcError):
            g.run(100)
        g.set_bw(False)
        assert g.run(100).scalars["n_err_dust"] == 0
    finally:
        g.close()
