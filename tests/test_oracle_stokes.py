"""The oracle's `stokes` (stokes.f90:1-392) and `stokespeel`
(stokespeel.f90:1-382) against what the mathematics fixes.

Neither builds in this tree (they call WRIMSG, whose source is absent), so no
golden vectors exist for them.  Their phase matrix, dustmat, and the peel
geometry, findangle3, are pinned bit-exactly against the reference
(test_oracle_golden.py); these checks pin how the two use them:

* the new direction makes the sampled angle with the old one: after one
  scattering of unpolarised light the degree of polarisation is |P12/P11| at
  mu = u_old . u_new, with I = 1 and V = 0;
* unpolarised light: mu follows P11 (the Cornette-Shanks function, mean
  3g(g^2+4)/(5(g^2+2))) and the scattering plane's azimuth about the old
  direction is uniform;
* linearly polarised light: the azimuth distribution is modulated by
  P12/P11: E[cos 2psi] = (1/2) int P12 / int P11 for Q = 1, and the same
  modulation moves to sin 2psi for U = 1;
* stokespeel's weights are the phase-matrix rows stokes samples from, and
  keep every physical Stokes vector physical.

Each check fails for a flipped sign in the Q or U coupling, in the new
direction or in the azimuth update, or for sin and cos exchanged in the U
coupling (tried on a copy of the oracle).

Both builds of the oracle are checked: the C library's elementary functions
and the deterministic set the GPU path is compared against.
"""
import ctypes as C

import numpy as np
import pytest
from scipy import integrate, stats

from oracle import pyoracle as po

N = 60_000


@pytest.fixture(scope="module", params=["libm", "dm"])
def lib(request):
    return po.load(request.param)


def _opac(g, pl):
    o = po.Opac()
    o.g[0], o.pl[0] = g, pl
    o.g2[0] = g * g
    o.g2p1[0] = o.g2[0] + 1.0
    o.onemg2[0] = 1.0 - o.g2[0]
    o.twog[0] = 2.0 * g
    return o


def _p(lib, o, mu):
    lib.hco_dustmat.argtypes = [C.POINTER(po.Opac), C.c_int, C.c_double, C.c_double,
                                C.POINTER(C.c_double), C.POINTER(C.c_double),
                                C.POINTER(C.c_double)]
    a, b, c = C.c_double(), C.c_double(), C.c_double()
    lib.hco_dustmat(C.byref(o), 0, mu, mu * mu, C.byref(a), C.byref(b), C.byref(c))
    return a.value, b.value


def _scatter(lib, o, iquv, n=N, cost=0.3, phi=1.1, seg=0):
    r = po.Rng()
    lib.hco_rng_packet(C.byref(r), 12345678, 87654321, 1, seg)
    sint = np.sqrt(1.0 - cost * cost)
    out = np.zeros((n, 8))
    s = po.Scat()
    for k in range(n):
        s.cost, s.sint, s.phi = cost, sint, phi
        s.cosp, s.sinp = np.cos(phi), np.sin(phi)
        s.I, s.Q, s.U, s.V = iquv
        lib.hco_stokes(C.byref(s), C.byref(o), 0, C.byref(r))
        out[k] = (s.cost, s.sint, s.phi, s.I, s.Q, s.U, s.V, s.cosp)
    uo = np.array([sint * np.cos(phi), sint * np.sin(phi), cost])
    e1 = np.array([cost * np.cos(phi), cost * np.sin(phi), -sint])
    e2 = np.array([-np.sin(phi), np.cos(phi), 0.0])
    un = np.stack([out[:, 1] * np.cos(out[:, 2]), out[:, 1] * np.sin(out[:, 2]), out[:, 0]], 1)
    return dict(mu=un @ uo, psi=np.arctan2(un @ e2, un @ e1), I=out[:, 3], Q=out[:, 4],
                U=out[:, 5], V=out[:, 6], cost=out[:, 0], sint=out[:, 1], phi=out[:, 2],
                cosp=out[:, 7])


def _moments(lib, o):
    p11 = integrate.quad(lambda m: _p(lib, o, m)[0], -1, 1, epsabs=1e-12)[0]
    p12 = integrate.quad(lambda m: _p(lib, o, m)[1], -1, 1, epsabs=1e-12)[0]
    mu1 = integrate.quad(lambda m: m * _p(lib, o, m)[0], -1, 1, epsabs=1e-12)[0]
    return p11, p12, mu1


@pytest.mark.parametrize("g,pl", [(0.5, 0.5), (-0.3, 0.8), (0.0, 1.0)])
def test_unpolarised_light(lib, g, pl):
    o = _opac(g, pl)
    s = _scatter(lib, o, (1.0, 0.0, 0.0, 0.0))
    # unit directions, cos/sin of phi kept with phi
    np.testing.assert_allclose(s["cost"] ** 2 + s["sint"] ** 2, 1.0, atol=1e-14)
    np.testing.assert_allclose(s["cosp"], np.cos(s["phi"]), atol=1e-15)
    # I stays 1, V stays 0, and the degree of polarisation is |P12/P11| at
    # the geometric scattering angle
    np.testing.assert_allclose(s["I"], 1.0, atol=1e-14)
    assert np.all(s["V"] == 0.0)
    mu = np.clip(s["mu"], -1.0, 1.0)
    ratio = np.array([abs(b / a) for a, b in (_p(lib, o, m) for m in mu[:5000])])
    pol = np.hypot(s["Q"], s["U"])[:5000]
    np.testing.assert_allclose(pol, ratio, atol=1e-9)
    # mu follows P11: the Cornette-Shanks mean cosine, and a KS test on the
    # CDF of P11
    p11, _, mu1 = _moments(lib, o)
    assert p11 == pytest.approx(2.0, rel=1e-9)            # normalised over 4 pi
    assert mu1 / p11 == pytest.approx(3 * g * (g * g + 4) / (5 * (g * g + 2)), abs=1e-9)
    se = mu.std() / np.sqrt(N)
    assert abs(mu.mean() - mu1 / p11) < 5 * se
    grid = np.linspace(-1.0, 1.0, 4001)
    pdf = np.array([_p(lib, o, m)[0] for m in grid])
    cdf = np.concatenate([[0.0], np.cumsum(0.5 * (pdf[1:] + pdf[:-1]) * np.diff(grid))])
    cdf /= cdf[-1]
    ks = stats.kstest(mu, lambda x: np.interp(x, grid, cdf))
    assert ks.pvalue > 1e-4, ks
    # the scattering plane's azimuth is uniform
    psi = s["psi"]
    h, _ = np.histogram(psi, bins=24, range=(-np.pi, np.pi))
    assert stats.chisquare(h).pvalue > 1e-4
    assert abs(np.cos(2 * psi).mean()) < 5 / np.sqrt(2 * N)
    assert abs(np.sin(2 * psi).mean()) < 5 / np.sqrt(2 * N)


@pytest.mark.parametrize("g,pl", [(0.2, 0.9), (0.6, 0.5)])
def test_linearly_polarised_light(lib, g, pl):
    """The azimuth of the scattering plane follows P11 + P12 cos 2(psi - psi0),
    psi0 = 0 for Q = 1 and 45 degrees (in either sense) for U = 1."""
    o = _opac(g, pl)
    p11, p12, _ = _moments(lib, o)
    c = 0.5 * p12 / p11
    assert c < 0.0                      # P12 < 0: the plane prefers psi = 90 degrees
    sq = _scatter(lib, o, (1.0, 1.0, 0.0, 0.0), seg=1)
    tol = 5 / np.sqrt(2 * N)
    assert abs(np.cos(2 * sq["psi"]).mean() - c) < tol
    assert abs(np.sin(2 * sq["psi"]).mean()) < tol
    su = _scatter(lib, o, (1.0, 0.0, 1.0, 0.0), seg=2)
    assert abs(abs(np.sin(2 * su["psi"]).mean()) - abs(c)) < tol
    assert abs(np.cos(2 * su["psi"]).mean()) < tol
    # the modulation is not noise
    assert abs(c) > 4 * tol


def _peel_geometry(lib, rng, n):
    """Random incoming and observer directions with the scattering cosine
    and rotation angle findangle3 (pinned against the reference) gives."""
    lib.hco_findangle3.argtypes = [C.c_double] * 10 + [C.POINTER(C.c_double)] * 4
    out = []
    for _ in range(n):
        c1, c2 = rng.uniform(-0.98, 0.98, 2)
        f1, f2 = rng.uniform(0.0, 2 * np.pi, 2)
        s1, s2 = np.sqrt(1 - c1 * c1), np.sqrt(1 - c2 * c2)
        o = [C.c_double() for _ in range(4)]
        lib.hco_findangle3(c1, s1, f1, np.cos(f1), np.sin(f1), c2, s2, f2, np.cos(f2),
                           np.sin(f2), *[C.byref(x) for x in o])
        out.append((c1, s1, c2, s2, o[0].value, o[3].value))
    return out


def _peel(lib, o, geo, iquv):
    c1, s1, c2, s2, bmu, ri1 = geo
    v = [C.c_double(x) for x in iquv]
    ce, se = C.c_double(c2), C.c_double(s2)
    lib.hco_stokespeel(*[C.byref(x) for x in v], C.byref(ce), C.byref(se), ri1, bmu, c1, s1,
                       C.byref(o), 0)
    return np.array([x.value for x in v])


@pytest.mark.parametrize("g,pl", [(0.5, 0.5), (-0.2, 0.9)])
def test_stokespeel_weights(lib, g, pl):
    """stokespeel.f90: the un-normalised phase-matrix weight of a peel-off
    ray.  Unpolarised in: I = P11, degree |P12|, V = 0.  Q = 1 in:
    I = P11 + P12 cos 2 i1; U = 1 in: I = P11 - P12 sin 2 i1 (the weights
    stokes samples from).  Any physical Stokes vector stays physical."""
    o = _opac(g, pl)
    rng = np.random.default_rng(9)
    for geo in _peel_geometry(lib, rng, 400):
        bmu, ri1 = geo[4], geo[5]
        p1, p2 = _p(lib, o, bmu)
        w = _peel(lib, o, geo, (1.0, 0.0, 0.0, 0.0))
        assert w[0] == p1 and w[3] == 0.0
        assert np.hypot(w[1], w[2]) == pytest.approx(abs(p2), rel=1e-9, abs=1e-14)
        assert _peel(lib, o, geo, (1.0, 1.0, 0.0, 0.0))[0] == pytest.approx(
            p1 + p2 * np.cos(2 * ri1), rel=1e-12, abs=1e-14)
        assert _peel(lib, o, geo, (1.0, 0.0, 1.0, 0.0))[0] == pytest.approx(
            p1 - p2 * np.sin(2 * ri1), rel=1e-12, abs=1e-14)
        q, u, v = rng.normal(size=3)
        pol = rng.uniform(0.0, 1.0) / np.sqrt(q * q + u * u + v * v)
        w = _peel(lib, o, geo, (1.0, q * pol, u * pol, v * pol))
        assert np.sqrt(w[1] ** 2 + w[2] ** 2 + w[3] ** 2) <= w[0] * (1 + 1e-12)
