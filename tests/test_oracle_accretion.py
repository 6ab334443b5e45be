"""The oracle's disk-accretion emission point (initpacc.f90:27-52) against
the distributions its sampling steps fix in closed form:
  - the radius: 1/r is drawn with density ~ 1 - sqrt(1/r) and kept inside
    (rddustmin, rmaxdmax), so p(r) ~ (1 - r^-1/2) / r^2 there;
  - the azimuth (rantheta, random.f90:28-47): uniform on [0, 2 pi);
  - the height: z / sigma(r) is standard normal (gasdev, random.f90:140-159),
    sigma(r) = sum over the disk species of fmass z1 r^b.
A wrong exponent, a dropped sqrt or a swapped sin/cos fails one of them."""
import ctypes as C

import numpy as np
import pytest
from scipy import stats

from hochunk3d_amd import build_model, config1

N = 20000


@pytest.fixture(scope="module")
def samples():
    from oracle.pyoracle import Oracle, Rng
    m = build_model(config1(mdot_disk=3e-6, nr=24, ntheta=24, nphi=12))
    o = Oracle(m)
    rng = Rng()
    o.lib.hco_rng_packet(C.byref(rng), 12345678, 87654321, 1, 7)
    out = np.zeros((N, 3))
    xyz = (C.c_double * 3)()
    for i in range(N):
        o.lib.hco_acc_position(C.byref(o.setup), C.byref(rng), xyz)
        out[i] = xyz[:]
    return m, out


def test_radius_distribution(samples):
    m, xyz = samples
    r = np.hypot(xyz[:, 0], xyz[:, 1])
    a, b = m.rddustmin, m.rmaxdmax
    assert r.min() > a and r.max() < b

    def prim(x):  # integral of (1 - x^-1/2) x^-2
        return -1.0 / x + (2.0 / 3.0) * x ** -1.5

    cdf = lambda x: (prim(np.clip(x, a, b)) - prim(a)) / (prim(b) - prim(a))  # noqa: E731
    assert stats.kstest(r, cdf).pvalue > 1e-3


def test_azimuth_uniform(samples):
    _, xyz = samples
    phi = np.arctan2(xyz[:, 1], xyz[:, 0]) % (2 * np.pi)
    assert stats.kstest(phi / (2 * np.pi), "uniform").pvalue > 1e-3
    # the sign of sin(phi) follows y: a swapped sin/cos would skew the quadrants
    q = np.histogram(phi, bins=4, range=(0, 2 * np.pi))[0]
    assert np.all(np.abs(q - N / 4) < 5 * np.sqrt(N / 4))


def test_height_gaussian_in_scale_heights(samples):
    m, xyz = samples
    r = np.hypot(xyz[:, 0], xyz[:, 1])
    sig = sum(m.fmass8[i] * m.z1_8[i] * r ** m.bflare8[i] for i in (0, 1, 4, 5))
    z = xyz[:, 2] / sig
    assert stats.kstest(z, "norm").pvalue > 1e-3
