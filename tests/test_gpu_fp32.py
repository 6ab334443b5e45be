"""The FP32 path (FP32 packet directions, Stokes vectors, frequencies and
event arithmetic with native elementary functions; FP64 positions, wall roots
and accumulators) against the plain FP64 path on the same packets.

Targets: SED error <= 1 % and Lucy temperature error <= 0.5 % against FP64.
The two runs draw the same random segments, so their difference is the
precision effect plus the Monte Carlo noise of histories that diverge."""
import numpy as np
import pytest

from hochunk3d_amd import build_model, config1
from hochunk3d_amd.testing import synthetic_temperature

pytestmark = pytest.mark.gpu

N = 2_000_000


@pytest.fixture(scope="module")
def runs():
    from hochunk3d_amd._lib import HC_FP32, HC_FP64, Transport
    m = build_model(config1(nfreq=50))
    m.tdust[...] = synthetic_temperature(m)
    out = {}
    for prec in (HC_FP64, HC_FP32):
        g = Transport(m, precision=prec)
        r = g.run(N, ipeel=0)
        t2, _ = g.temperature(N)
        g.close()
        out[prec] = (r, t2)
    return m, out[HC_FP64], out[HC_FP32]


def test_fp32_sed_within_one_percent(runs):
    m, (r64, _), (r32, _) = runs
    # SED summed over frequency per inclination bin (slot 0: all photons)
    s64 = r64.sed[0, :, :, 0, 0, 0].sum(axis=0)
    s32 = r32.sed[0, :, :, 0, 0, 0].sum(axis=0)
    err = np.abs(s32 - s64) / s64
    print("SED per inclination, rel err:", np.round(err, 5))
    assert err.max() <= 0.01
    # the spectrum itself, luminosity-weighted
    f64 = r64.sed[0, :, :, 0, 0, 0].sum(axis=1)
    f32 = r32.sed[0, :, :, 0, 0, 0].sum(axis=1)
    werr = np.abs(f32 - f64).sum() / f64.sum()
    print("luminosity-weighted spectral difference", werr)
    assert werr <= 0.01


def test_fp32_temperature_within_half_percent(runs):
    m, (r64, t64), (r32, t32) = runs
    well = r64.nabs >= 2000        # cells with a well-sampled absorbed energy
    assert well.sum() > 100
    rel = np.abs(t32[well] - t64[well]) / t64[well]
    print("T rel err: median", np.median(rel), "p95", np.percentile(rel, 95), "max", rel.max())
    assert np.median(rel) <= 0.005
    assert np.percentile(rel, 95) <= 0.01


def test_fp32_energy_bookkeeping(runs):
    m, (r64, _), (r32, _) = runs
    for k in ("sumall", "flux"):
        assert abs(r32.scalars[k] - r64.scalars[k]) <= 0.002 * r64.scalars[k], k
    # no packet is lost or dropped at the work bound on either path
    for r in (r64, r32):
        assert r.scalars["n_err_nowall"] == 0
        assert r.scalars["n_err_dust"] == 0
        assert r.scalars["n_warn"] == 0


def test_fp32_launch_on_a_phi_wall_walks():
    """Packet 1116954 of the 2e6-packet pass draws a stellar longitude of
    exactly half a turn in FP32.  Its launch point used to land exactly on
    the phi = pi wall of the stellar surface (the native FP32 sine of pi is
    0), where the walk stepped between the two phi cells without moving until
    the work bound dropped it.  The launch point is FP64 now: the packet
    walks out like its FP64 twin, and its path is the FP64 twin's up to the
    FP32 event arithmetic."""
    from hochunk3d_amd._lib import HC_FP32, HC_FP64, Transport
    m = build_model(config1(nfreq=50))
    m.tdust[...] = synthetic_temperature(m)
    out = {}
    for prec in (HC_FP32, HC_FP64):
        g = Transport(m, precision=prec)
        r = g.run(1, ipeel=0, seg0=1116954, trace=True)
        g.close()
        out[prec] = r
    t32, t64 = out[HC_FP32].trace[0], out[HC_FP64].trace[0]
    print("fp32", t32, "\nfp64", t64)
    assert out[HC_FP32].scalars["n_warn"] == 0
    assert t32["status"] in (0, 1) and t64["status"] in (0, 1)
    assert t32["path"] > 1.0 and t32["ncross"] < 100_000


def test_fp32_peeled_images_and_sed_within_one_percent():
    """The final peeled pass on the FP32 path against FP64 on the same
    packets: per observer, the image flux (Stokes I summed over pixels and
    frequencies), the polarised fraction of the image sums, and the peeled
    SED summed over frequency."""
    from hochunk3d_amd._lib import HC_FP32, HC_FP64, Transport
    m = build_model(config1(nfreq=50, nximg=41))
    m.tdust[...] = synthetic_temperature(m)
    out = {}
    for prec in (HC_FP64, HC_FP32):
        g = Transport(m, precision=prec)
        out[prec] = g.run(200_000, ipeel=1)
        g.close()
    a, b = out[HC_FP64], out[HC_FP32]
    ia = a.img_i[..., 0].sum(axis=(1, 2, 3))     # (npeel,) all photons
    ib = b.img_i[..., 0].sum(axis=(1, 2, 3))
    err = np.abs(ib - ia) / ia
    print("image flux per observer, rel err:", np.round(err, 5))
    assert err.max() <= 0.01
    qa = np.hypot(a.img_q.sum(axis=(1, 2, 3)), a.img_u.sum(axis=(1, 2, 3))) / ia
    qb = np.hypot(b.img_q.sum(axis=(1, 2, 3)), b.img_u.sum(axis=(1, 2, 3))) / ib
    print("polarised fraction per observer:", np.round(qa, 5), np.round(qb, 5))
    assert np.abs(qb - qa).max() <= 0.002
    pa = a.peel[0, :, :, 0, 0].sum(axis=0)        # peeled SED, I, per observer
    pb = b.peel[0, :, :, 0, 0].sum(axis=0)
    assert (np.abs(pb - pa) / pa).max() <= 0.01
