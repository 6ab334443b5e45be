"""The CPU oracle pinned against Hochunk3d's own leaf routines.

tests/golden/ref_leaves.npz holds input/output vectors that
oracle/ref/golden_driver.f90 wrote after compiling the reference's sources
(ecuyer_cote_mod, random, radface, phiface, closest_wall, nr.f locate, dust.f90
opacset/setemiss/KapP/chiR, lucydustfreq, dusttemp, select_dust, dustmat,
findangle3, mathsubs ierfc, isotrp) in the build container.  Everything but
setemiss (adaptive Simpson in the reference, Gauss-Legendre here) must match
bit for bit.
"""
import ctypes as C
import os

import numpy as np
import pytest

from oracle import pyoracle as po

HERE = os.path.dirname(os.path.abspath(__file__))
G = np.load(os.path.join(HERE, "golden", "ref_leaves.npz"))
SEEDS = (12345678, 87654321)


@pytest.fixture(scope="module")
def lib():
    return po.load("libm")


def _rng(lib, gen, seg=0):
    r = po.Rng()
    lib.hco_rng_packet(C.byref(r), SEEDS[0], SEEDS[1], gen, seg)
    return r


def test_rng_streams_and_segments(lib):
    for row in G["rng"]:
        gen, seg = int(row[0]), int(row[1])
        r = _rng(lib, gen, seg)
        got = [lib.hco_ran(C.byref(r)) for _ in range(6)]
        assert got == list(row[2:]), (gen, seg)


def test_bbfreq(lib):
    r = _rng(lib, 1)
    for T, nu in G["bbfreq"]:
        assert lib.hco_bbfreq(C.byref(r), T) == nu


def test_wall_roots(lib):
    for row in G["walls"]:
        onw = int(row[0])
        ux, uy, uz, x, y, z, r2a, r2b, a1, b1, a2, b2 = row[1:13]
        in1, ip1, ip2, face = (int(v) for v in row[13:17])
        tn1, tp1, tp2, t = row[17:21]
        w = po.Walls()
        lib.hco_reset_t(C.byref(w))
        lib.hco_radface(C.byref(w), ux, uy, uz, x, y, z, r2a, r2b)
        lib.hco_phiface(C.byref(w), a1, b1, a2, b2, ux, uy, x, y)
        assert (w.tn1, w.tp1, w.tp2) == (tn1, tp1, tp2)
        assert (w.in1, w.ip1, w.ip2) == (in1, ip1, ip2)
        if face >= 0:
            tt, ff = C.c_double(), C.c_int()
            assert lib.hco_find_next_wall(C.byref(w), onw, C.byref(tt), C.byref(ff)) == 1
            assert (tt.value, ff.value) == (t, face)


def test_locate(lib):
    xx = np.array([i * i * 0.25 for i in range(1, 13)])
    for x, j in G["locate"]:
        assert lib.hco_locate(xx.ctypes.data_as(C.POINTER(C.c_double)), 12, x) == int(j)


@pytest.fixture(scope="module")
def dust():
    from hochunk3d_amd.dust import dustinit, synthetic_dust_table
    return dustinit(synthetic_dust_table())


def test_dustinit_arrays(dust):
    nl, kv = G["kappav"][0]
    assert int(nl) == dust.nlambda[0] and kv == dust.kappav[0]
    d = G["dustarr"]
    i = d[:, 0].astype(int) - 1
    np.testing.assert_array_equal(dust.nudust[i, 0], d[:, 1])
    np.testing.assert_array_equal(dust.kappad[i, 0], d[:, 2])
    np.testing.assert_array_equal(dust.adust[i, 0], d[:, 3])
    np.testing.assert_array_equal(dust.gdust[i, 0], d[:, 4])


def _dust_struct(dust, kapBint=None, kappap=None, kappar=None):
    keep = {}

    def p(name, arr):
        a = np.asfortranarray(arr, dtype=np.float64)
        keep[name] = a
        return a.ctypes.data_as(C.POINTER(C.c_double))

    d = po.Dust()
    d.nlammax = dust.nudust.shape[0]
    for i in range(8):
        d.nlambda[i] = int(dust.nlambda[i])
        d.kappav[i] = float(dust.kappav[i])
    d.nudust, d.kappad = p("nud", dust.nudust), p("kd", dust.kappad)
    d.adust, d.gdust, d.pldust = p("ad", dust.adust), p("gd", dust.gdust), p("pd", dust.pldust)
    d.nT, d.nnu = dust.lTint.size, dust.nuint.size
    d.lTint, d.nuint = p("lT", dust.lTint), p("nu", dust.nuint)
    d.kapBint = p("kb", dust.kapBint if kapBint is None else kapBint)
    d.kappap = p("kp", dust.kappap if kappap is None else kappap)
    if kappar is not None:
        d.kappar = p("kr", kappar)
    return d, keep


def test_opacset(lib, dust):
    d, keep = _dust_struct(dust)
    for nu, kap, alb, g, pl in G["opacset"]:
        o = po.Opac()
        lib.hco_opacset(C.byref(d), nu, C.byref(o))
        assert (o.kappa[0], o.albedo[0], o.g[0], o.pl[0]) == (kap, alb, g, pl), nu


def test_setemiss_tables(dust):
    """Quadrature tables: 1e-5 relative is the reference's own QSIMP tolerance."""
    kb = G["kapBint"]
    got = dust.kapBint[kb[:, 0].astype(int) - 1, kb[:, 1].astype(int) - 1, 0]
    np.testing.assert_allclose(got, kb[:, 2], rtol=2e-5, atol=1e-12)
    kp = G["kappap"]
    np.testing.assert_array_equal(dust.lTint, kp[:, 2])
    np.testing.assert_allclose(dust.kappap[:, 0], kp[:, 1], rtol=2e-5)


def test_setemiss_bw_table(dust):
    """The Bjorkman-Wood table kapdBint (kappa_abs dB/dT) of the same setemiss
    call, to the reference's QSIMP tolerance."""
    from hochunk3d_amd.dust import kapdB_table
    kdb = kapdB_table(dust)
    ref = G["kapdBint"]
    got = kdb[ref[:, 0].astype(int) - 1, ref[:, 1].astype(int) - 1, 0]
    np.testing.assert_allclose(got, ref[:, 2], rtol=2e-5, atol=1e-12)


def test_kapp_on_reference_table(lib, dust):
    kp = np.asfortranarray(np.tile(G["kappap"][:, 1:2], (1, 8)))
    d, keep = _dust_struct(dust, kappap=kp)
    for T, val in G["kapp"]:
        assert lib.hco_kapp(C.byref(d), T, 0) == val


def _formula_kapB(nT, nnu):
    j = (np.arange(nnu, dtype=float) / 1024.0)[None, :]
    w = (np.arange(1, nT + 1, dtype=float) / 1000.0)[:, None]
    one = j ** 2 * w + j * (1.0 - w)
    return np.asfortranarray(np.repeat(one[:, :, None], 8, axis=2))


def test_lucydustfreq_on_formula_table(lib, dust):
    d, keep = _dust_struct(dust, kapBint=_formula_kapB(dust.lTint.size, dust.nuint.size))
    r = _rng(lib, 2)
    for T, nu in G["lucy"]:
        assert lib.hco_lucydustfreq(C.byref(d), C.byref(r), T, 0) == nu, T


def test_dustfreq_on_formula_table(lib, dust):
    """dustfreq (Bjorkman-Wood re-emission) on the formula table in
    kapdBint, below and above the 1600 K sublimation temperature (where it
    draws bbfreq): the reference's draws and frequencies bit for bit."""
    d, keep = _dust_struct(dust)
    kb = _formula_kapB(dust.lTint.size, dust.nuint.size)
    keep["kdb"] = kb
    d.kapdBint = kb.ctypes.data_as(C.POINTER(C.c_double))
    r = _rng(lib, 7)
    hot = 0
    for T, nu in G["dustfreq"]:
        assert lib.hco_dustfreq(C.byref(d), C.byref(r), T, 0, 1600.0) == nu, T
        hot += T >= 1600.0 / 11605.0
    assert 1 <= hot < len(G["dustfreq"])


def test_dusttemp_on_formula_table(lib, dust):
    i = np.arange(1, 1001, dtype=float)
    kp = np.asfortranarray(np.tile((1e-4 * i + 1e-6 * (i * i))[:, None], (1, 8)))
    d, keep = _dust_struct(dust, kappap=kp)
    for m4, na, T0, Td in G["dusttemp"]:
        assert lib.hco_dusttemp(C.byref(d), na, T0, 3.839e33, 100000.0, m4, 0) == Td


def test_select_dust(lib):
    kapd = np.array([1.0 + 0.5 * i for i in range(1, 9)])
    dens = np.array([0.0 if i % 3 == 0 else 1.0 / i for i in range(1, 9)])
    r = _rng(lib, 4)
    kp = kapd.ctypes.data_as(C.POINTER(C.c_double))
    dp = dens.ctypes.data_as(C.POINTER(C.c_double))
    lib.hco_select_dust.argtypes = [C.POINTER(C.c_double), C.POINTER(C.c_double),
                                    C.POINTER(po.Rng)]
    for (idust2,) in G["select"]:
        assert lib.hco_select_dust(kp, dp, C.byref(r)) == int(idust2)


def test_dustmat(lib):
    lib.hco_dustmat.argtypes = [C.POINTER(po.Opac), C.c_int, C.c_double, C.c_double,
                                C.POINTER(C.c_double), C.POINTER(C.c_double),
                                C.POINTER(C.c_double)]
    for g, pl, cost, p1, p2, p3 in G["dustmat"]:
        o = po.Opac()
        o.g[0], o.pl[0] = g, pl
        o.g2[0] = g * g
        o.g2p1[0] = o.g2[0] + 1.0
        o.onemg2[0] = 1.0 - o.g2[0]
        o.twog[0] = 2.0 * g
        a, b, c = C.c_double(), C.c_double(), C.c_double()
        lib.hco_dustmat(C.byref(o), 0, cost, cost * cost, C.byref(a), C.byref(b), C.byref(c))
        assert (a.value, b.value, c.value) == (p1, p2, p3)


def test_findangle3(lib):
    for row in G["findangle3"]:
        out = [C.c_double() for _ in range(4)]
        lib.hco_findangle3(*[float(v) for v in row[:10]], *[C.byref(o) for o in out])
        assert [o.value for o in out] == list(row[10:14])


def test_setemiss_rosseland_table(dust):
    """kappar (the Rosseland mean of setemiss, the diffusion layer's chiR)
    to the reference's QSIMP tolerance."""
    from hochunk3d_amd.dust import ensure_kappar
    ref = G["kappar"]
    got = ensure_kappar(dust)[ref[:, 0].astype(int) - 1, 0]
    np.testing.assert_allclose(got, ref[:, 1], rtol=2e-5)


def test_chir_on_reference_table(lib, dust):
    """chiR (dust.f90:83-111) on the reference's own kappar, inside the table
    and beyond both ends: bit for bit, the oracle's and the host's."""
    from hochunk3d_amd.diffusion import chir
    kr = np.asfortranarray(np.tile(G["kappar"][:, 1:2], (1, 8)))
    d, keep = _dust_struct(dust, kappar=kr)
    for T, val in G["chir"]:
        assert lib.hco_chir(C.byref(d), T, 0) == val, T
        assert chir(kr, dust.lTint, T, 0) == val, T


def test_ierfc(lib):
    from hochunk3d_amd.diffusion import ierfc
    for x, val in G["ierfc"]:
        assert ierfc(x) == val, x


def test_isotrp(lib):
    """isotrp (isotrp.f90) on generator 8: the reference's draws and
    directions bit for bit."""
    r = _rng(lib, 8)
    u = np.zeros(3)
    for sux, suy, suz, ux, uy, uz in G["isotrp"]:
        lib.hco_isotrp(C.byref(r), sux, suy, suz, u.ctypes.data_as(C.POINTER(C.c_double)))
        assert (u[0], u[1], u[2]) == (ux, uy, uz)
