"""The stellar spectrum table (atmosinit, starbarb.f90:3-118), its sampler
atmosfreq (starbarb.f90:120-156) and the Eddington limb darkening of the
stellar launch (samptabl.f90, initp_common.f90:1-29, newdisktherm.F90:349-355)
pinned by what the mathematics fixes.  starbarb.f90 needs the missing
tts_mod and samptabl.f90 the missing tabl_mod, so neither builds from the
reference tree; the fixture is the reference's own T = 4000 K model
atmosphere (tests/golden/kt04000g+4.0z-2.0.npz, made by
tests/golden/make_atmos_fixture.py).

* atmosinit: the normalised cumulative flux equals an independent
  trapezoid quadrature (scipy) of F_nu over 1/lambda; with CPLANCKST the
  table follows the analytic Planck integral;
* atmosfreq: it maps the table's nodes onto themselves, is monotone, fills
  each bracket with the probability the table gives it, and within a bracket
  follows the log-log interpolated cumulative flux;
* a pass through an empty grid emits the table's spectrum;
* limb darkening: the emission cosine has the cumulative probability
  (mu^3 + mu^2)/2 (intensity ~ 1 + 3/2 mu), the direct peel weight
  12/5 (mu + mu^2) has mean 1 and mean square 2.976, and the flux of the
  limb-darkened disk inside projected radius a is
  (a^2 + 1 - (1 - a^2)^(3/2)) / 2."""
import ctypes as C
import os

import numpy as np
import pytest
from scipy import integrate, stats

from hochunk3d_amd import build_model, config1
from hochunk3d_amd.atmos import atmosinit, bnu, limb_table
from oracle import pyoracle as po
from oracle.pyoracle import Oracle

HERE = os.path.dirname(os.path.abspath(__file__))
FIX = np.load(os.path.join(HERE, "golden", "kt04000g+4.0z-2.0.npz"))
WAVE, FLUX = FIX["wave"], FIX["flux"]
TSTAR_EV = 4000.0 / 11605.0
ORACLE = "libm"


def test_table_is_the_trapezoid_cumulative_flux():
    a = atmosinit(WAVE, FLUX)
    winv = 1.0 / WAVE[::-1]
    cum = integrate.cumulative_trapezoid(FLUX[::-1], winv, initial=0.0)
    np.testing.assert_allclose(a.hnuint, cum / cum[-1], rtol=1e-12, atol=1e-15)
    np.testing.assert_allclose(a.atmosnu, 1.2398 * winv, rtol=1e-15)
    assert a.hnuint[0] == 0.0 and a.hnuint[-1] == 1.0
    np.testing.assert_allclose(10.0 ** a.lognu, a.atmosnu, rtol=1e-14)
    np.testing.assert_allclose(10.0 ** a.loghnu[1:], a.hnuint[1:], rtol=1e-14)


def test_cplanckst_table_follows_the_planck_integral():
    a = atmosinit(WAVE, FLUX, TSTAR_EV, iplanckst=True)
    x = a.atmosnu / TSTAR_EV
    f = lambda y: y ** 3 / np.expm1(y) if y > 0 else 0.0  # noqa: E731
    exact = np.array([integrate.quad(f, 0.0, v, limit=200)[0] for v in x]) / (np.pi ** 4 / 15)
    # the trapezoid on the file's wavelength grid is within 2e-3 of the
    # Planck integral
    assert np.abs(a.hnuint - exact).max() < 2e-3
    # and bnu is the normalised Planck function (integrates to 1 in nu/T)
    assert abs(integrate.quad(lambda y: float(bnu(y, 1.0)), 0, 60, limit=400)[0] - 1.0) < 1e-6


def _setup(atm):
    s = po.Setup()
    keep = [np.ascontiguousarray(getattr(atm, k)) for k in ("atmosnu", "hnuint", "lognu",
                                                           "loghnu")]
    s.nhnu = atm.nhnu
    s.atmosnu, s.hnuint, s.lognu, s.loghnu = (k.ctypes.data_as(C.POINTER(C.c_double))
                                              for k in keep)
    return s, keep


def _freq(lib, s, xi):
    return np.array([lib.hco_atmosfreq_xi(C.byref(s), float(v)) for v in xi])


@pytest.mark.parametrize("variant", ["libm", "dm"])
def test_atmosfreq_maps_nodes_to_nodes_and_is_monotone(variant):
    lib = po.load(variant)
    a = atmosinit(WAVE, FLUX)
    s, keep = _setup(a)
    # nodes of strictly increasing cumulative flux (above the first bin)
    j = np.nonzero(np.diff(a.hnuint) > 0)[0] + 1
    j = j[j >= 2]
    got = _freq(lib, s, a.hnuint[j][a.hnuint[j] < 1.0])
    np.testing.assert_allclose(got, a.atmosnu[j][a.hnuint[j] < 1.0], rtol=1e-12)
    xi = np.sort(np.random.default_rng(3).uniform(0, 1, 20000))
    nu = _freq(lib, s, xi)
    assert np.all(np.diff(nu) >= 0.0)
    assert nu.min() >= a.atmosnu[0] and nu.max() <= a.atmosnu[-1]


def test_atmosfreq_fills_brackets_with_their_probability_and_log_law():
    lib = po.load(ORACLE)
    a = atmosinit(WAVE, FLUX)
    s, keep = _setup(a)
    n = 200_000
    nu = _freq(lib, s, np.random.default_rng(11).uniform(0, 1, n))
    # brackets grouped into 40 equal-probability groups
    edges_p = np.linspace(0, 1, 41)
    idx = np.searchsorted(a.hnuint, edges_p)
    edges_nu = a.atmosnu[np.clip(idx, 0, a.nhnu - 1)]
    edges_nu[0], edges_nu[-1] = 0.0, np.inf
    p = np.diff(a.hnuint[np.clip(idx, 0, a.nhnu - 1)])
    p[-1] += 1.0 - p.sum()
    obs = np.histogram(nu, bins=edges_nu)[0]
    keep_ = n * p > 5
    chi2 = ((obs[keep_] - n * p[keep_]) ** 2 / (n * p[keep_])).sum()
    assert stats.chi2.sf(chi2, keep_.sum() - 1) > 1e-4, chi2
    # within the most probable bracket: the cumulative flux interpolated in
    # log nu - log hnuint (np.interp), KS
    k = int(np.argmax(np.diff(a.hnuint)))
    inb = nu[(nu > a.atmosnu[k]) & (nu <= a.atmosnu[k + 1])]
    cdf = lambda v: (10.0 ** np.interp(np.log10(v), a.lognu[k:k + 2], a.loghnu[k:k + 2])  # noqa
                     - a.hnuint[k]) / (a.hnuint[k + 1] - a.hnuint[k])
    assert inb.size > 500
    assert stats.kstest(inb, cdf).pvalue > 1e-4


def _transparent(**over):
    m = build_model(config1(nr=6, ntheta=6, nphi=4, nfreq=60, npeel=3, nap=2, **over))
    m.densarr[...] = 0.0
    m.aperture2[:] = [0.5 ** 2, 1e12]
    return m


def test_pass_emits_the_atmosphere_spectrum():
    m = _transparent(atmosphere=(WAVE, FLUX))
    o = Oracle(m, ORACLE)
    n = 60_000
    r = o.run(n, ipeel=0)
    cnt = r["sed"][8][:, :, :, 1, 0].sum(axis=(1, 2))
    a = m.atmosphere
    edges = m.numin * np.exp(np.arange(m.cfg.nfreq + 1) * m.lnurat / m.cfg.nfreq)
    F = 10.0 ** np.interp(np.log10(np.clip(edges, a.atmosnu[1], a.atmosnu[-1])), a.lognu,
                          a.loghnu)
    F[edges <= a.atmosnu[1]] = 0.0
    p = np.diff(F)
    p[-1] += 1.0 - F[-1]
    keep = n * p > 5
    chi2 = ((cnt[keep] - n * p[keep]) ** 2 / (n * p[keep])).sum()
    assert stats.chi2.sf(chi2, keep.sum() - 1) > 1e-4, chi2
    assert r["scalars"]["sumall"] == n


def test_limb_darkening_cosine_law():
    lib = po.load(ORACLE)
    xmu, prob = limb_table(1000)
    s = po.Setup()
    s.ntab = xmu.size
    s.xmu = xmu.ctypes.data_as(C.POINTER(C.c_double))
    s.prob = prob.ctypes.data_as(C.POINTER(C.c_double))
    s.delmu = 1.0 / xmu.size
    # nodes map onto nodes
    got = np.array([lib.hco_darkening(C.byref(s), float(v)) for v in prob[2:-1]])
    np.testing.assert_allclose(got, xmu[2:-1], rtol=1e-12)
    # halfway in probability between two nodes is halfway in mu: the table
    # is interpolated linearly within each bin
    mid = np.array([lib.hco_darkening(C.byref(s), float(v))
                    for v in 0.5 * (prob[2:-1] + prob[3:])])
    np.testing.assert_allclose(mid, 0.5 * (xmu[2:-1] + xmu[3:]), rtol=1e-12)
    xi = np.random.default_rng(5).uniform(0, 1, 100_000)
    mu = np.array([lib.hco_darkening(C.byref(s), float(v)) for v in xi])
    assert mu.min() > 0.0 and mu.max() < 1.0
    assert stats.kstest(mu, lambda x: 0.5 * (x ** 3 + x ** 2)).pvalue > 1e-4
    # the isotropic law F = mu^2 fails it
    assert stats.kstest(mu, lambda x: x ** 2).pvalue < 1e-6


def test_limb_darkened_disk_flux_and_peel_weight():
    m = _transparent(limb=1)
    o = Oracle(m, ORACLE)
    n = 60_000
    r = o.run(n, ipeel=1, images=False)
    a = 0.5
    frac = r["sed"][8][:, :, :, 0, 0].sum() / n
    assert abs(frac - (a * a + 1 - (1 - a * a) ** 1.5) / 2) < 0.01, frac
    for k in range(m.cfg.npeel):
        w = r["peel"][0][:, k, 1, 0].sum() / n
        w2 = r["peel"][4][:, k, 1, 0].sum() / n
        c = r["peel"][8][:, k, 1, 0].sum() / n
        # E[W^2] = 2.88 (1/3 + 1/2 + 1/5); Var W^2 = E W^4 - (E W^2)^2 < 26
        assert abs(w - 1.0) < 4 * np.sqrt((2.976 - 1.0) / n), w
        assert abs(w2 - 2.976) < 4 * np.sqrt(26.0 / n), w2
        assert abs(c - 0.5) < 4 * np.sqrt(0.25 / n), c
