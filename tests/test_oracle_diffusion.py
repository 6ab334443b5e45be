"""The diffusion approximation (gridset.f90:666-789, emit_common.f90:80-387)
in the oracle and the host mask, pinned by what geometry and the definitions
fix (isotrp, chiR and ierfc are bit-exact against the reference's routines in
test_oracle_golden.py):

* isotrp's directions have density 2 mu in mu = u.s (a Lambertian surface):
  the cosine's CDF is mu^2, its mean 2/3;
* ierfc inverts erfc to A&S 26.2.23's accuracy (4.5e-4 in the normal tail's
  argument, 3.2e-4 in erfc's);
* the host mask of a massive disk is a band about the midplane, mirror
  symmetric, the same in every phi cell, directed up above and down below;
* a packet leaving through a radial surface lands on the layer's inner wall
  on its own radial line, heading inwards, in the cell below; one leaving
  through a theta surface moves along its meridian to cos(theta) = 1.00001
  cos of the surface cone, heading away from the layer, in the cell beyond;
* the surface cell's T^4 gains 4 pi T*^4 / (nstar dA) with dA the ring area
  of the surface (2 pi r^2 (cos t1 - cos t2) radial, pi sin t (r2^2 - r1^2)
  in theta), for both disk species, capped at 1600 K;
* behind the surface, with no absorbed counts the cells take the surface's
  T^4; with counts n_j and a constant Rosseland ratio chi they take
  T^4 = T_s^4 + 3 pi T*^4 / (nstar dA) kappa_V R* chi
  sum_{j<i} (sum_{k>=j} n_k) rho_j dr_j (the column's cumulative sums).

Each holds against a planted mistake of test_oracle_plants.py."""
import ctypes as C

import numpy as np
import pytest
from scipy import special, stats

from hochunk3d_amd import build_model, config1
from hochunk3d_amd.diffusion import ierfc
from hochunk3d_amd.model import PI, RSOL
from oracle import pyoracle as po

ORACLE = "libm"
TSUB = 1600.0 / 11605.0


def _rng(lib, gen=9, seg=0):
    r = po.Rng()
    lib.hco_rng_packet(C.byref(r), 12345678, 87654321, gen, seg)
    return r


def test_isotrp_is_lambertian_about_the_normal():
    lib = po.load(ORACLE)
    r = _rng(lib)
    g = np.random.default_rng(4)
    u = np.zeros(3)
    mus = []
    for _ in range(20000):
        s = g.normal(size=3)
        s /= np.linalg.norm(s)
        lib.hco_isotrp(C.byref(r), s[0], s[1], s[2], u.ctypes.data_as(C.POINTER(C.c_double)))
        assert abs(np.linalg.norm(u) - 1.0) < 1e-14
        mus.append(float(np.dot(u, s)))
    mus = np.array(mus)
    assert mus.min() >= 0.0
    assert stats.kstest(mus, lambda x: np.clip(x, 0, 1) ** 2).pvalue > 1e-3
    assert abs(mus.mean() - 2.0 / 3.0) < 4 * np.sqrt(1.0 / 18.0 / mus.size)


def test_ierfc_inverts_erfc():
    xs = np.linspace(0.05, 3.0, 300)
    got = np.array([ierfc(float(special.erfc(x))) for x in xs])
    assert np.abs(got - xs).max() < 3.2e-4
    assert np.all(np.diff(got) > 0)


def thick_model():
    """A massive disk (0.1 Msun) whose inner midplane holds a layer."""
    return build_model(config1(diffusion=True, massd=0.1, nphi=6, nfreq=40, nximg=21))


def radial_model():
    m = build_model(config1(massd=0.1, nr=24, ntheta=32, nphi=6, nfreq=40, nximg=21))
    m.tdust[...] = 300.0 / 11605.0
    return _setup_radial(m)


@pytest.fixture(scope="module")
def thick():
    return thick_model()


def test_mask_is_a_midplane_band(thick):
    m = thick
    d, dd = m.diffus, m.diffdir
    assert d.sum() > 0
    assert np.all(d[..., :-1] == d[..., :1])              # the same in every phi cell
    assert np.all(d[:, :, -1] == 0) and np.all(d[-1] == 0) and np.all(d[:, -1] == 0)
    assert np.all(d[:2] == 0)                              # nr0 = 3: cells 1, 2 excluded
    nt = m.ntg - 1
    for ir in range(m.nrg - 1):
        band = np.nonzero(d[ir, :nt, 0])[0]
        if band.size == 0:
            continue
        assert np.all(np.diff(band) == 1)                  # contiguous
        assert np.array_equal(d[ir, :nt, 0], d[ir, :nt, 0][::-1])   # mirror symmetric
        for it in band:
            assert dd[ir, it, 0] == (-2 if m.thetarr[it] < PI / 2 else 2)
    rows = np.nonzero(d[:, :, 0].any(axis=1))[0]
    assert np.all(np.diff(rows) == 1) and rows[0] == 2     # from the first allowed cell out


def _setup_radial(m):
    """A synthetic radial layer: cells ir 6..9 (1-based) of theta cells
    it 12..21 in every phi cell, direction -1 there and in ir 10..12."""
    d = np.zeros((m.nrg, m.ntg, m.npg), np.int8, order="F")
    dd = np.zeros_like(d)
    d[5:9, 11:21, :-1] = 1
    dd[5:12, 11:21, :-1] = -1
    m.diffus, m.diffdir = d, dd
    return m


@pytest.fixture()
def radial():
    return radial_model()


def _kapd(m):
    return np.array([m.dust.kappav[i] * m.rstar * RSOL * 1e-3 for i in range(8)])


def test_radial_surface_move_direction_and_temperatures(radial):
    m = radial
    o = po.Oracle(m, ORACLE)
    lib = o.lib
    rng = _rng(lib)
    nstar = 1.0e5
    it, ip = 15, 3
    t0 = o.tdust.copy()
    th = 0.5 * (m.thetarr[it - 1] + m.thetarr[it])
    ph = 0.5 * (m.phiarr[ip - 1] + m.phiarr[ip])
    r = 0.5 * (m.rarr[7] + m.rarr[8])                     # inside cell ir = 8
    p = r * np.array([np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), np.cos(th)])
    id2, x, y, z, ux, uy, uz, nub, cell, onw = o.emit_diffusion(p, (8, it, ip), _kapd(m), nstar,
                                                              rng)
    assert id2 in (1, 2) and onw and cell == (5, it, ip)
    q = np.array([x, y, z])
    np.testing.assert_allclose(q, p * m.rarr[5] / np.linalg.norm(p), rtol=1e-15, atol=0)
    assert np.dot([ux, uy, uz], q) < 0.0                  # inwards
    dA = 2.0 * PI * m.rarr[5] ** 2 * (np.cos(m.thetarr[it - 1]) - np.cos(m.thetarr[it]))
    inc = 4.0 * PI * m.tstar_ev ** 4 / (nstar * dA)
    for id_ in (0, 1):
        ts = o.tdust[5, it - 1, ip - 1, id_]
        want = min(TSUB, (t0[5, it - 1, ip - 1, id_] ** 4 + inc) ** 0.25)
        assert abs(ts - want) <= 1e-14 * want
        # nothing absorbed behind: every cell of the -1 column takes T_s
        for ir in range(7, 13):
            got = o.tdust[ir - 1, it - 1, ip - 1, id_]
            assert abs(got - max(t0[ir - 1, it - 1, ip - 1, id_], want)) <= 1e-14 * want
        assert o.tdust[12, it - 1, ip - 1, id_] == t0[12, it - 1, ip - 1, id_]
    # other theta cells untouched
    assert np.array_equal(o.tdust[:, it, ...], t0[:, it, ...])
    assert nub > 0


def test_column_sums_with_absorbed_counts(radial):
    m = radial
    m.dust.kappar = np.full((m.dust.lTint.size, 8), 0.37, order="F")   # chiR = 0.37
    o = po.Oracle(m, ORACLE)
    rng = _rng(o.lib, seg=3)
    nstar = 2.0e4
    it, ip = 13, 2
    nabs = np.zeros((m.nrg, m.ntg, m.npg, 8), order="F")
    cnt = {6: 3.0, 7: 0.5, 8: 2.25, 9: 1.0, 10: 0.75, 11: 4.0, 12: 0.125}
    for ir, n in cnt.items():
        nabs[ir - 1, it - 1, ip - 1, 0] = n
        nabs[ir - 1, it - 1, ip - 1, 1] = 2.0 * n
    t0 = o.tdust.copy()
    th = 0.5 * (m.thetarr[it - 1] + m.thetarr[it])
    ph = 0.5 * (m.phiarr[ip - 1] + m.phiarr[ip])
    r = 0.5 * (m.rarr[6] + m.rarr[7])                     # cell 7
    p = r * np.array([np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), np.cos(th)])
    o.emit_diffusion(p, (7, it, ip), _kapd(m), nstar, rng, nabs=nabs)
    dA = 2.0 * PI * m.rarr[5] ** 2 * (np.cos(m.thetarr[it - 1]) - np.cos(m.thetarr[it]))
    A = 3.0 * PI * m.tstar_ev ** 4 / (nstar * dA)
    irs = np.arange(6, 13)                                 # the -1 column from the surface
    for id_, f in ((0, 1.0), (1, 2.0)):
        n = np.array([cnt[i] * f for i in irs])
        S = np.cumsum(n[::-1])[::-1]                       # S_j = sum_{k >= j} n_k
        rho = m.densarr[irs - 1, it - 1, ip - 1, id_]
        w = m.rarr[irs] - m.rarr[irs - 1]
        T4s = t0[5, it - 1, ip - 1, id_] ** 4 + 4.0 * PI * m.tstar_ev ** 4 / (nstar * dA)
        terms = A * m.dust.kappav[id_] * m.rstar * RSOL * 0.37 * S * rho * w
        for k, ir in enumerate(irs[1:]):
            want = min(TSUB, (T4s + terms[:k + 1].sum()) ** 0.25)
            want = max(want, t0[ir - 1, it - 1, ip - 1, id_])
            got = o.tdust[ir - 1, it - 1, ip - 1, id_]
            assert abs(got - want) <= 1e-12 * want, (id_, ir, got, want)


def test_theta_surfaces_up_and_down(thick):
    m = thick
    m.tdust[...] = 200.0 / 11605.0
    d = m.diffus
    ir, ip = 4, 2                                          # 1-based
    band = np.nonzero(d[ir - 1, :, ip - 1])[0] + 1         # layer cells, 1-based
    assert band.size >= 4
    nstar = 5.0e4
    for it0, up in ((band[1], True), (band[-2], False)):
        o = po.Oracle(m, ORACLE)
        rng = _rng(o.lib, seg=int(it0))
        t0 = o.tdust.copy()
        th = 0.5 * (m.thetarr[it0 - 1] + m.thetarr[it0])
        ph = 0.5 * (m.phiarr[ip - 1] + m.phiarr[ip])
        r = 0.5 * (m.rarr[ir - 1] + m.rarr[ir])
        p = r * np.array([np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), np.cos(th)])
        id2, x, y, z, ux, uy, uz, nub, cell, onw = o.emit_diffusion(p, (ir, it0, ip), _kapd(m),
                                                                  nstar, rng)
        assert id2 in (1, 2) and not onw
        wall = band[0] if up else band[-1] + 1             # the surface cone (1-based wall)
        cw = np.cos(m.thetarr[wall - 1])
        q = np.array([x, y, z])
        assert abs(z / r - 1.00001 * cw) < 1e-14
        assert abs(np.linalg.norm(q) - r) < 1e-12 * r      # on its sphere
        assert abs(np.arctan2(y, x) - np.arctan2(p[1], p[0])) < 1e-14   # on its meridian
        # the surface's normal away from the layer: -theta-hat above, +theta-hat below
        sth, cth = np.sqrt(x * x + y * y) / r, z / r
        that = np.array([cth * x / (r * sth), cth * y / (r * sth), -sth])
        nrm = -that if up else that
        assert np.dot([ux, uy, uz], nrm) > 0.0
        assert cell == (ir, band[0] - 1 if up else band[-1] + 1, ip)
        surf = band[0] if up else band[-1]
        dA = PI * np.sin(m.thetarr[wall - 1]) * (m.rarr[ir] ** 2 - m.rarr[ir - 1] ** 2)
        inc = 4.0 * PI * m.tstar_ev ** 4 / (nstar * dA)
        for id_ in (0, 1):
            want = min(TSUB, (t0[ir - 1, surf - 1, ip - 1, id_] ** 4 + inc) ** 0.25)
            got = o.tdust[ir - 1, surf - 1, ip - 1, id_]
            assert abs(got - want) <= 1e-14 * want
            # the cells behind, to the midplane, take T_s (no counts)
            side = [i for i in band if (i > surf and m.thetarr[i - 1] < PI / 2)] if up else \
                   [i for i in band if (i < surf and m.thetarr[i - 1] >= PI / 2)]
            for i in side:
                g2 = o.tdust[ir - 1, i - 1, ip - 1, id_]
                assert abs(g2 - max(t0[ir - 1, i - 1, ip - 1, id_], want)) <= 1e-14 * want
        assert nub > 0


def test_pass_with_a_layer_counts_no_absorption_there(thick):
    m = thick
    from hochunk3d_amd.testing import synthetic_temperature
    m.tdust[...] = synthetic_temperature(m)
    o = po.Oracle(m, "dm")
    r = o.run(3000, ipeel=0, trace=True)
    s = r["scalars"]
    assert s["n_err_nowall"] == 0 and s["n_err_dust"] == 0
    lay = m.diffus == 1
    assert r["dbig"][lay].sum() > 0                       # packets were absorbed there
    assert r["nabs"][lay].sum() == 0.0                    # and not counted
    mask = np.repeat(lay[..., None], 8, axis=3)
    assert np.any(o.tdust[mask] != m.tdust[mask])         # and heated the layer
