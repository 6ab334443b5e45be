"""Comparison helpers shared by the parity tests, the smoke check and the
benchmark's self-check (no oracle code here: these compare arrays)."""
from __future__ import annotations

import numpy as np

EXACT_FIELDS = ("status", "ncross", "npeelcross", "nscat", "nabs", "nstar", "cellhash", "ir",
                "it", "ip")
FLOAT_FIELDS = ("x", "y", "z", "cost", "sint", "phi", "I", "Q", "U", "V", "nub", "path",
                "tau_peel")


def synthetic_temperature(model, t_in: float = 1700.0, slope: float = -0.45,
                          tmin: float = 8.0) -> np.ndarray:
    """A dust temperature field (eV) that falls as r^slope from t_in K at the
    inner dust radius: it puts the innermost cells above the 1600 K
    sublimation temperature (exercising the opacity mask) and spans the
    emissivity table elsewhere."""
    r = 0.5 * (model.rarr[:-1] + model.rarr[1:])
    t = np.clip(t_in * (r / model.rarr[1]) ** slope, tmin, 2000.0)
    out = np.full(model.tdust.shape, 0.1 / 11605.0, order="F")
    out[:-1, :-1, :-1, :] = (t / 11605.0)[:, None, None, None]
    return out


def compare_traces(a: np.ndarray, b: np.ndarray, rtol: float = 1e-12) -> dict:
    """Assert two per-packet trace arrays agree: integer fields (status,
    crossing/event counts, cell-sequence hash, final cell) bit-exactly, and the
    floating-point fields to rtol relative (scaled by max(|a|,|b|,1))."""
    assert a.shape == b.shape, (a.shape, b.shape)
    report = {}
    for k in EXACT_FIELDS:
        bad = np.nonzero(a[k] != b[k])[0]
        if bad.size:
            i = bad[0]
            raise AssertionError(f"trace field {k} differs for {bad.size} packets; first p={i}: "
                                 f"{a[i]} vs {b[i]}")
    for k in FLOAT_FIELDS:
        x, y = a[k], b[k]
        scale = np.maximum(np.maximum(np.abs(x), np.abs(y)), 1.0)
        err = np.abs(x - y) / scale
        report[k] = float(err.max()) if err.size else 0.0
        if err.size and err.max() > rtol:
            i = int(np.argmax(err))
            raise AssertionError(f"trace field {k} rel err {err.max():.3e} > {rtol} at p={i}: "
                                 f"{x[i]!r} vs {y[i]!r}")
    return report


def rel_close(a, b, rtol: float, atol: float = 0.0) -> float:
    a = np.asarray(a, float)
    b = np.asarray(b, float)
    err = np.abs(a - b)
    tol = atol + rtol * np.maximum(np.abs(a), np.abs(b))
    bad = err > tol
    if bad.any():
        i = np.unravel_index(int(np.argmax(err - tol)), a.shape)
        raise AssertionError(f"max excess at {i}: {a[i]!r} vs {b[i]!r}")
    return float((err / np.maximum(np.maximum(np.abs(a), np.abs(b)), 1e-300)).max()) if a.size else 0.0


def assert_temperatures_close(actual, desired, same_start: bool):
    """Dust temperatures of two runs over the same packets whose tallies were
    summed in different orders.  From the same starting temperatures the
    update's fixed-point iteration (stopping at 1e-5) bounds the difference
    at 2e-5 in every cell.  After further iterations the histories can
    diverge in a few packets, so at most 0.1 % of the cells may differ by more
    and none by more than 5 %.  A replay of the wrong packets moves most cells
    by the Monte Carlo noise and fails either way."""
    a = np.asarray(actual, dtype=float)
    d = np.asarray(desired, dtype=float)
    rel = np.abs(a - d) / np.maximum(np.abs(d), 1e-300)
    if same_start:
        assert rel.max() <= 2e-5, f"max relative temperature difference {rel.max():.3g}"
    else:
        frac = float(np.mean(rel > 2e-5))
        assert frac <= 1e-3 and rel.max() <= 0.05, (
            f"{frac:.2%} of the cells differ by more than 2e-5, max {rel.max():.3g}")
