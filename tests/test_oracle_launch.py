"""The oracle's stellar emission direction (initp_common.f90:1-94, limb = 0)
against what the geometry fixes.  initp_common.f90 needs the missing tts_mod,
so no golden vectors exist for it.  A packet leaves the surface point with
normal n = (sinb cos lp, sinb sin lp, cosb):

* its direction u is a unit vector with u . n = sqrt(xi1), xi1 the packet's
  first draw (isotropic intensity: the emitted flux goes as mu, F(mu) = mu^2);
* its azimuth psi about n, measured from the direction of increasing polar
  angle, has cos psi = cos(2 pi xi2), xi2 the second draw, and one
  handedness for every packet; psi is uniform.

A dropped term or a flipped sign in initp_transform breaks one of these."""
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
