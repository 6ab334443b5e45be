"""Accuracy of the deterministic FP64 elementary functions (hc_math.h) that
the plain FP64 GPU path and the oracle's "dm" build share, measured in ulp
against the C library over the argument ranges the transport path uses."""
import ctypes as C

import numpy as np
import pytest

from oracle import pyoracle as po

FN = {"log": 0, "exp": 1, "sin": 2, "cos": 3, "acos": 4, "pow": 5, "atan2": 6, "log10": 7,
      "exp10": 8}


def _eval(name, x, y=None):
    lib = po.load("dm")
    x = np.ascontiguousarray(x, np.float64)
    y = np.ascontiguousarray(np.zeros_like(x) if y is None else y, np.float64)
    out = np.empty_like(x)
    dp = C.POINTER(C.c_double)
    lib.hco_math(FN[name], x.ctypes.data_as(dp), y.ctypes.data_as(dp), out.ctypes.data_as(dp),
                 x.size)
    return out


def _ulp_err(got, ref):
    return np.abs(got - ref) / np.spacing(np.abs(ref))


rng = np.random.default_rng(5)
CASES = [
    ("log", 10.0 ** rng.uniform(-300, 300, 200000), None, np.log, 2.0),
    ("log", rng.uniform(4.7e-10, 1.0, 200000), None, np.log, 2.0),
    ("log", 1.0 + rng.uniform(-1e-3, 1e-3, 100000), None, np.log, 2.0),
    ("exp", rng.uniform(-740, 700, 200000), None, np.exp, 2.0),
    ("exp", rng.uniform(-30, 0, 200000), None, np.exp, 2.0),
    ("sin", rng.uniform(-7, 13, 200000), None, np.sin, 2.0),
    ("cos", rng.uniform(-7, 13, 200000), None, np.cos, 2.0),
    ("acos", rng.uniform(-1, 1, 200000), None, np.arccos, 2.0),
    ("pow", rng.uniform(0.9, 1.2, 200000), rng.uniform(0, 1, 200000), np.power, 2.0),
    ("pow", 10.0 ** rng.uniform(-30, 30, 100000), np.full(100000, 0.25), np.power, 1.0),
    ("log10", 10.0 ** rng.uniform(-300, 300, 200000), None, np.log10, 3.0),
    ("log10", rng.uniform(1e-12, 1.0, 200000), None, np.log10, 3.0),
    # photon energies in eV: 10^y for |y| < 3
    ("exp10", rng.uniform(-3, 3, 200000), None, lambda v: 10.0 ** v, 4.0),
]


@pytest.mark.parametrize("name,x,y,ref,maxulp", CASES, ids=[c[0] for c in CASES])
def test_accuracy(name, x, y, ref, maxulp):
    got = _eval(name, x, y)
    want = ref(x) if y is None else ref(x, y)
    err = _ulp_err(got, want)
    assert np.all(np.isfinite(got))
    assert err.max() <= maxulp, (name, err.max(), x[np.argmax(err)])


def test_sin_cos_near_zeros():
    # relative accuracy where the result is small: near multiples of pi/2
    k = np.arange(-4, 9)
    x = (k[:, None] * (np.pi / 2) + np.linspace(-1e-6, 1e-6, 2001)[None, :]).ravel()
    for name, ref in (("sin", np.sin), ("cos", np.cos)):
        got = _eval(name, x)
        want = ref(x)
        assert np.max(np.abs(got - want)) <= 4e-16 * np.maximum(np.abs(want), 1e-10).max()


def test_acos_endpoints_and_atan2():
    x = np.array([-1.0, -0.5, 0.0, 0.5, 1.0, 1 - 2 ** -52, -1 + 2 ** -52])
    np.testing.assert_allclose(_eval("acos", x), np.arccos(x), rtol=4e-16, atol=1e-22)
    a = rng.uniform(-5, 5, 10000)
    b = rng.uniform(-5, 5, 10000)
    np.testing.assert_allclose(_eval("atan2", a, b), np.arctan2(a, b), rtol=0, atol=1e-14)


def test_libm_and_dm_oracles_agree_on_packets():
    """The two oracle builds differ only in the last bits of elementary
    functions: on the same packets the integer history is the same for nearly
    all packets."""
    from hochunk3d_amd import build_model, config1
    from hochunk3d_amd.testing import synthetic_temperature
    m = build_model(config1(nr=24, ntheta=24, nphi=12))
    m.tdust[...] = synthetic_temperature(m)
    a = po.Oracle(m, "libm").run(1000, trace=True)["trace"]
    b = po.Oracle(m, "dm").run(1000, trace=True)["trace"]
    same = (a["cellhash"] == b["cellhash"]) & (a["nscat"] == b["nscat"])
    assert same.mean() > 0.95
