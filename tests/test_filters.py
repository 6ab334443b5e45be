"""Filter curves (filt_nu.f90) and the oracle's filter-convolved peel images
(peeloff_3d.f90:179-211), pinned by what the definitions fix:

* a curve flat in wavelength normalises to 1 / (its frequency width), and a
  curve linear in frequency to (nu - nu0) / ((nu1 - nu0)^2 / 2): the
  trapezoidal rule is exact on both, and an integral over wavelength instead
  of frequency, or with the nodes' order reversed, gives other values;
* the reference's own filter files read to filt_mod.f90's line counts give
  curves whose frequencies increase strictly and whose response-weighted
  mean wavelengths fall in the named bands (2MASS J, H, K; IRAC 3.6-8.0;
  PACS 70-160, SPIRE 250-500);
* a ray peeled from a known point of an empty grid lands in the pixel of its
  sky projection in the filter whose node range holds its frequency, with
  weight the response at the last node below that frequency (not the node
  above), Stokes Q, U, V weighted like I, and nothing outside the node
  ranges (strict bounds at both ends).

The writer's text layout is checked against the edit descriptors of
output.F90:380-411."""
import os

import numpy as np
import pytest

from hochunk3d_amd import filters as F
from hochunk3d_amd.output import diskflux, write_filter_images
from oracle.pyoracle import Oracle
from test_oracle_closed_forms import _kap, cell_of, uniform_model

ORACLE = "libm"
REF = "/root/reference"


def test_flat_curve_normalises_to_one_over_frequency_width():
    w = np.linspace(2.0, 1.0, 11)                     # decreasing, as in the files
    f = F.filt([("A", w, np.full(11, 3.0))])
    nu = F.FACT / w
    assert np.all(np.diff(f.filtwave[:11, 0]) > 0)
    np.testing.assert_allclose(f.filtwave[:11, 0], nu, rtol=1e-15)
    want = 1.0 / (F.FACT / 1.0 - F.FACT / 2.0)
    np.testing.assert_allclose(f.filtphi[:11, 0], want, rtol=1e-13)
    assert f.nwfilt[0] == 12 and f.filtphi[11:, 0].sum() == 0.0


def test_curve_linear_in_frequency_normalises_to_its_exact_integral():
    w = np.array([3.0, 2.5, 2.2, 1.7, 1.5, 1.1, 1.0])   # uneven in frequency
    nu = F.FACT / w
    phi = nu - nu[0]
    f = F.filt([("B", w, phi)])
    want = (nu - nu[0]) / (0.5 * (nu[-1] - nu[0]) ** 2)
    np.testing.assert_allclose(f.filtphi[:7, 0], want, rtol=1e-12, atol=1e-30)


@pytest.mark.skipif(not os.path.isdir(REF), reason="reference filter files absent")
def test_reference_filter_files_read_to_their_line_counts():
    f = F.reference_filters(REF)
    assert len(f.names) >= 10
    # band centres in microns (response-weighted mean wavelength)
    bands = {"2J": (1.1, 1.4), "2H": (1.5, 1.8), "2K": (2.0, 2.4), "I1": (3.2, 4.0),
             "I2": (4.0, 5.0), "I3": (5.2, 6.3), "I4": (7.0, 9.0), "H1": (60, 85),
             "H2": (85, 120), "H3": (130, 190), "H4": (220, 280), "H5": (310, 390),
             "H6": (440, 560)}
    for k, n in enumerate(f.names):
        nf = int(f.nwfilt[k]) - 1
        assert f.nwfilt[k] == F.NWFILT[F.FILTNAMES.index(n)]
        nu = f.filtwave[:nf, k]
        ph = f.filtphi[:nf, k]
        assert np.all(np.diff(nu) > 0), n
        assert abs(np.trapezoid(ph, nu) - 1.0) < 1e-12, n
        if n in bands:
            lam = F.FACT / nu
            mean = np.trapezoid(ph * lam, nu) / np.trapezoid(ph, nu)
            lo, hi = bands[n]
            assert lo < mean < hi, (n, mean)
    # the 2K file has one row more than filt_mod counts; filt reads 75
    if "2K" in f.names:
        k = f.names.index("2K")
        w, _ = F.read_filter_file(os.path.join(REF, "2K.txt"))
        assert len(w) == 76 and f.nwfilt[k] - 1 == 75


def test_filter_weight_is_the_last_node_below():
    f = F.synthetic_filters()
    rng = np.random.default_rng(3)
    for k in range(f.nfilt):
        nf = int(f.nwfilt[k]) - 1
        fw, fp = f.filtwave[:nf, k], f.filtphi[:nf, k]
        for i in rng.integers(0, nf - 1, 12):
            nu = np.sqrt(fw[i] * fw[i + 1])
            assert F.filter_weight(f, k, nu) == fp[i]
        assert F.filter_weight(f, k, fw[0]) is None        # strict at both ends
        assert F.filter_weight(f, k, fw[nf - 1]) is None
        assert F.filter_weight(f, k, fw[3]) == fp[2]        # on a node: the one below


def _observer(m, k):
    ob = np.array([m.sinte[k] * m.cospe[k], m.sinte[k] * m.sinpe[k], m.coste[k]])
    east = np.cross([0.0, 0.0, 1.0], ob)
    east /= np.linalg.norm(east)
    return ob, east, np.cross(ob, east)


def _filter_model():
    return uniform_model(0.0, nr=10, ntheta=8, nphi=6, npeel=3, nximg=31, nfreq=40)


def test_filter_image_pixel_weight_and_stokes():
    m = _filter_model()
    o = Oracle(m, ORACLE)
    f = F.synthetic_filters()
    o.set_filters(f)
    nx, rmaxi = m.cfg.nximg, m.rmaxi
    kd, _ = _kap(0.0)
    rng = np.random.default_rng(11)
    stokes = (0.8, 0.3, -0.2, 0.1)
    for k in range(3):
        ob, east, north = _observer(m, k)
        for j in range(f.nfilt):
            nf = int(f.nwfilt[j]) - 1
            fw, fp = f.filtwave[:nf, j], f.filtphi[:nf, j]
            i = int(rng.integers(1, nf - 2))
            nub = np.sqrt(fw[i] * fw[i + 1]) * 1.2398 / F.FACT
            assert F.filter_weight(f, j, F.ray_frequency(nub)) == fp[i]
            p = rng.uniform(-0.55, 0.55, 3) * rmaxi
            if np.dot(p, ob) < 0:
                p = p - 2 * np.dot(p, ob) * ob
            if np.linalg.norm(p) < 3.0:
                p = 40.0 * ob
            t = o.peel_ray(p, cell_of(m, p), kd, stokes, k, nub)
            ix = int(np.floor((np.dot(p, east) + rmaxi) / (2 * rmaxi / nx)))
            iy = int(np.floor((np.dot(p, north) + rmaxi) / (2 * rmaxi / nx)))
            bi = t["image_b_i"]
            nz = np.argwhere(bi != 0.0)
            assert nz.tolist() == [[k, ix, iy, j]], (k, j, nz)
            assert bi[k, ix, iy, j] == stokes[0] * fp[i]
            # Q, U, V carry the same weight as I (the cube holds them unweighted)
            for s in "quv":
                cube = t[f"img_{s}"][k, ix, iy, :].sum()
                got = t[f"image_b_{s}"][k, ix, iy, j]
                assert got == cube * fp[i], (s, got, cube)
            assert t["img_q"][k, ix, iy, :].sum() != 0.0


def test_filter_image_outside_every_node_range_is_empty():
    m = _filter_model()
    o = Oracle(m, ORACLE)
    f = F.synthetic_filters()
    o.set_filters(f)
    kd, _ = _kap(0.0)
    ob, _, _ = _observer(m, 1)
    p = 40.0 * ob
    for j in range(f.nfilt):
        nf = int(f.nwfilt[j]) - 1
        for nu in (f.filtwave[0, j] * (1 - 1e-9), f.filtwave[nf - 1, j] * (1 + 1e-9)):
            t = o.peel_ray(p, cell_of(m, p), kd, (1.0, 0, 0, 0), 1, nu * 1.2398 / F.FACT)
            assert t["image_b_i"].sum() == 0.0
            assert t["img_i"][..., 0].sum() == 1.0            # the ray was peeled
        nu = f.filtwave[0, j] * (1 + 1e-9)                 # just inside: the first node
        t = o.peel_ray(p, cell_of(m, p), kd, (1.0, 0, 0, 0), 1, nu * 1.2398 / F.FACT)
        assert t["image_b_i"].sum() == f.filtphi[0, j]
    # a ray exactly on an end node: the end nodes' wavelengths are the rays'
    # own 1.2398 / nub, so the frequencies agree to the last bit
    nub_lo, nub_hi = 0.5, 0.6
    w = np.linspace(1.2398 / nub_lo, 1.2398 / nub_hi, 9)
    w[0], w[-1] = 1.2398 / nub_lo, 1.2398 / nub_hi
    g = F.filt([("E", w, np.linspace(1.0, 2.0, 9))])
    o.set_filters(g)
    assert g.filtwave[0, 0] == F.ray_frequency(nub_lo)
    assert g.filtwave[8, 0] == F.ray_frequency(nub_hi)
    for nub in (nub_lo, nub_hi):
        t = o.peel_ray(p, cell_of(m, p), kd, (1.0, 0, 0, 0), 1, nub)
        assert t["image_b_i"].sum() == 0.0
    t = o.peel_ray(p, cell_of(m, p), kd, (1.0, 0, 0, 0), 1, 0.55)
    assert t["image_b_i"].sum() == F.filter_weight(g, 0, F.ray_frequency(0.55)) > 0


def test_filter_image_files_follow_the_edit_descriptors(tmp_path):
    m = _filter_model()
    f = F.synthetic_filters((0.55, 2.2))
    npeel, nx = 3, m.cfg.nximg
    rng = np.random.default_rng(5)
    bi = rng.uniform(0, 1, (npeel, nx, nx, 2))
    bi[0, 3, 4, 1] = 0.0
    bq, bu, bv = (rng.normal(size=bi.shape) * bi for _ in range(3))
    sed = np.zeros((9, m.cfg.nfreq, m.cfg.nmu, m.cfg.nph, m.cfg.nap, 8))
    fl = diskflux(m, 1000.0, sed, filt=(bi, bq, bu, bv))
    # q, u, v as fractions of I (zero where I is), I in intensity units
    assert fl.image_b_q[0, 3, 4, 1] == 0.0
    np.testing.assert_allclose(fl.image_b_q[1], (bq[1] / bi[1]).astype(np.float32), rtol=1e-6)
    pos = bi > 0
    ratio = bi.astype(np.float32)[pos] / fl.image_b_i[pos]
    assert np.ptp(ratio) <= 1e-6 * ratio.mean()
    thete = np.radians([30.0, 87.1, 145.3])
    phie = np.radians([0.0, 45.0, 359.9])
    paths = write_filter_images(str(tmp_path), fl, f.names, thete, phie)
    assert len(paths) == npeel * 2 * 4
    name = os.path.basename(paths[4 * 3 + 2])           # observer 1, filter S2, U
    assert name == "e_S2_87.1_45.0_U_img.dat"
    assert os.path.basename(paths[-1]) == "e_S2_145.3_359.9_V_img.dat"
    rows = open(paths[0]).read().split("\n")[:-1]
    assert len(rows) == nx and all(len(r) == 12 * nx for r in rows)
    # row iy holds ix = 1..nx; ES11.4 then one blank
    for iy in (0, 7, nx - 1):
        vals = [float(rows[iy][12 * ix:12 * ix + 11]) for ix in range(nx)]
        np.testing.assert_allclose(vals, fl.image_b_i[0, :, iy, 0], rtol=6e-5)
        assert rows[iy][11] == " " and rows[iy][12:23].strip()[1] == "."
