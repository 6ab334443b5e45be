"""The oracle's stellar launch against what the geometry fixes: the surface
point and hot-spot selection (initpspot.f90:3-120) and the emission direction
(initp_common.f90:1-94, limb = 0).  These sources need the missing tts_mod,
so no golden vectors exist for them.  A packet leaves the surface point with
normal n = (sinb cos lp, sinb sin lp, cosb):

* its direction u is a unit vector with u . n = sqrt(xi1), xi1 the packet's
  first draw (isotropic intensity: the emitted flux goes as mu, F(mu) = mu^2);
* its azimuth psi about n, measured from the direction of increasing polar
  angle, has cos psi = cos(2 pi xi2), xi2 the second draw, and one
  handedness for every packet; psi is uniform.

A dropped term or a flipped sign in initp_transform breaks one of these.
The surface point is uniform on the star, or within the spot caps for spot
packets; a stellar packet that starts in a cap is a spot packet."""
import ctypes as C

import numpy as np
import pytest
from scipy import stats

from oracle import pyoracle as po


@pytest.fixture(scope="module", params=["libm", "dm"])
def lib(request):
    return po.load(request.param)


@pytest.mark.parametrize("cosb,lp", [(0.6, 0.4), (-0.35, 4.0), (0.02, 2.9), (-0.97, 6.0)])
def test_emission_direction(lib, cosb, lp):
    sinb = np.sqrt(1.0 - cosb * cosb)
    n = np.array([sinb * np.cos(lp), sinb * np.sin(lp), cosb])
    # tangent frame at the surface point
    e1 = np.array([cosb * np.cos(lp), cosb * np.sin(lp), -sinb])
    e2 = np.array([-np.sin(lp), np.cos(lp), 0.0])
    r = po.Rng()
    lib.hco_rng_packet(C.byref(r), 12345678, 87654321, 1, 0)
    m = 40_000
    mu = np.empty(m)
    xi = np.empty(m)
    xi2 = np.empty(m)
    psi = np.empty(m)
    ct, st, ph = C.c_double(), C.c_double(), C.c_double()
    for k in range(m):
        peek = po.Rng(r.s1, r.s2)
        xi[k] = lib.hco_ran(C.byref(peek))
        xi2[k] = lib.hco_ran(C.byref(peek))
        lib.hco_initp_direction(C.byref(r), cosb, sinb, lp, C.byref(ct), C.byref(st),
                                C.byref(ph))
        u = np.array([st.value * np.cos(ph.value), st.value * np.sin(ph.value), ct.value])
        assert abs(u @ u - 1.0) < 1e-13
        mu[k] = u @ n
        psi[k] = np.arctan2(u @ e2, u @ e1)
    np.testing.assert_allclose(mu, np.sqrt(xi), atol=1e-9)
    # the local azimuth: cos exact, one handedness (away from the poles of n)
    ok = mu < 0.999
    np.testing.assert_allclose(np.cos(psi[ok]), np.cos(2 * np.pi * xi2[ok]), atol=1e-6)
    s2 = np.sin(2 * np.pi * xi2[ok])
    big = np.abs(s2) > 1e-3
    hand = np.sign(np.sin(psi[ok][big])) * np.sign(s2[big])
    assert np.all(hand == hand[0])
    assert stats.kstest(mu, lambda x: np.clip(x, 0, 1) ** 2).pvalue > 1e-4
    h, _ = np.histogram(psi, bins=24, range=(-np.pi, np.pi))
    assert stats.chisquare(h).pvalue > 1e-4


D2R = np.pi / 180.0


def _spots(th1, ph1, th2, ph2, rad, nspot, fspot=0.06):
    s = po.Setup()
    s.ispot, s.nspot, s.fspot, s.csprad = 1, nspot, fspot, float(np.cos(rad * D2R))
    for k, (th, ph) in ((1, (th1, ph1)), (2, (th2, ph2))):
        t, p = th * D2R, ph * D2R
        setattr(s, f"cspot{k}", float(np.cos(t)))
        setattr(s, f"sspot{k}", float(np.sin(t)))
        setattr(s, f"scspot{k}", float(np.sin(t) * np.cos(p)))
        setattr(s, f"ssspot{k}", float(np.sin(t) * np.sin(p)))
    return s


def _points(lib, s, hotspot, m, seg=0):
    r = po.Rng()
    lib.hco_rng_packet(C.byref(r), 12345678, 87654321, 1, seg)
    cb, sb, lp = C.c_double(), C.c_double(), C.c_double()
    out = np.empty((m, 4))
    for k in range(m):
        f = lib.hco_star_point(C.byref(s), C.byref(r), hotspot, C.byref(cb), C.byref(sb),
                               C.byref(lp))
        out[k] = (cb.value, sb.value, lp.value, f)
    n = np.stack([out[:, 1] * np.cos(out[:, 2]), out[:, 1] * np.sin(out[:, 2]), out[:, 0]], 1)
    return n, out[:, 3]


def _centre(s, k):
    return np.array([getattr(s, f"scspot{k}"), getattr(s, f"ssspot{k}"), getattr(s, f"cspot{k}")])


def test_stellar_points_are_uniform_on_the_star(lib):
    s = po.Setup()                                   # no spot
    n, flag = _points(lib, s, 0, 40_000)
    assert np.all(flag == 0)
    assert stats.kstest(n[:, 2], "uniform", args=(-1, 2)).pvalue > 1e-4
    lp = np.arctan2(n[:, 1], n[:, 0]) % (2 * np.pi)
    assert stats.kstest(lp, "uniform", args=(0, 2 * np.pi)).pvalue > 1e-4


@pytest.mark.parametrize("nspot", [1, 2])
def test_hot_spot_points_fill_the_caps_uniformly(lib, nspot):
    """initpspot.f90: a spot packet starts uniformly within a cap of angular
    radius acos(csprad) about a spot centre (either one when nspot = 2)."""
    s = _spots(30, 70, 140, 200, 20, nspot)
    n, flag = _points(lib, s, 1, 30_000, seg=5)
    assert np.all(flag == 1)
    d1, d2 = n @ _centre(s, 1), n @ _centre(s, 2)
    inside1 = d1 >= s.csprad
    inside = inside1 | (d2 >= s.csprad) if nspot == 2 else inside1
    assert np.all(inside)
    d = d1[inside1] if nspot == 1 else np.concatenate([d1[inside1], d2[~inside1]])
    # uniform on a cap: the cosine to the centre is uniform on [csprad, 1]
    assert stats.kstest(d, "uniform", args=(s.csprad, 1 - s.csprad)).pvalue > 1e-4
    if nspot == 2:
        assert abs(inside1.mean() - 0.5) < 5 * 0.5 / np.sqrt(len(n))


def test_stellar_packets_that_land_in_a_spot(lib):
    """A stellar packet becomes a spot packet when it starts inside a cap.
    The reference tests the caps by longitude difference from 0 (spot 1) and
    from pi (spot 2) (initpspot.f90), which is the cap test for spots centred
    at those longitudes."""
    s = _spots(30, 0, 140, 180, 20, 2)
    n, flag = _points(lib, s, 0, 40_000, seg=9)
    d = np.maximum(n @ _centre(s, 1), n @ _centre(s, 2))
    edge = np.abs(d - s.csprad) < 1e-12
    np.testing.assert_array_equal(flag[~edge] == 1, d[~edge] > s.csprad)
    assert 0.01 < flag.mean() < 0.2
