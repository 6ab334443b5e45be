"""Hierarchically clumped density multiplier (the fractal_sph construction,
fractal_sph.f90:1-175, with its normalisation in gridset.f90:796-870).

Points are placed in ``levels + 1`` nested generations: ``n`` parents drawn
uniformly in the unit cube, each with ``n`` children offset by uniform
deviates scaled by ``(1/fl)^k`` at generation k.  Fractal dimension
D = ln(n)/ln(fl) (n = 32, fl = 3.792 gives D = 2.6).  Each point is mapped to
the spherical grid (radius scaled to ``2 rmax/(1 + 1/fl)``) and counted; the
counts become a multiplier with mean 1:

    mult = counts / <counts> * densratio + 1 - densratio.

The reference draws the deviates from its legacy generator ``ran_old`` in
depth-first order; this construction draws them from NumPy's PCG64 seeded
with ``seed``, so clump positions are statistically equivalent to, not
identical with, the reference's for the same seed.

Draw order: the 3n parent coordinates, then for each parent in turn, level
by level, the 3 n^k offsets of level k in (node, child, coordinate) order.
The cell of a point is found with comparisons only (no arccos/arctan2), so
that ``hc_fractal_counts`` (hc_grid.hip), which jumps the same PCG64 stream
to each point's draws, counts the same points in the same cells:

* radius: the number of radial walls below r * scale;
* theta: the number of walls with cos(theta_wall) > z/r;
* phi: the number of walls below the point's azimuth, decided per wall by
  the half-plane (y > 0, or y = 0 and x > 0, is [0, pi)) and the sign of
  cos(phi_wall) y - sin(phi_wall) x within the same half-plane.
"""
from __future__ import annotations

from typing import Optional

import numpy as np

PI = 3.1415926535897932384626433


def _cell_counts_of(pts, rarr, mcos_t, cph, sph, scale_r, counts):
    """Bins pts (m,3) into counts (nrg,ntg,npg) (column-major flat index);
    walls as described in the module docstring."""
    nrg, ntg, npg = counts.shape
    x, y, z = pts[:, 0], pts[:, 1], pts[:, 2]
    r = np.sqrt(x * x + y * y + z * z)
    cz = np.clip(z / r, -1.0, 1.0)
    ir = np.searchsorted(rarr, r * scale_r, side="left")
    ir = np.where(ir == 0, 1, ir)
    it = np.maximum(np.searchsorted(mcos_t, -cz, side="left"), 1)
    # azimuth: start from arctan2 and settle on the comparison count
    ph = np.arctan2(y, x)
    ph = np.where(ph < 0.0, ph + 2.0 * PI, ph)
    c = np.searchsorted(np.arange(npg) * (2.0 * PI / (npg - 1)), ph, side="left")  # estimate
    lower = (y < 0.0) | ((y == 0.0) & (x < 0.0))

    def below(j):           # wall j lies below the point's azimuth
        jj = np.clip(j, 0, npg - 1)
        hw = 2 * jj >= npg - 1
        cr = cph[jj] * y - sph[jj] * x
        b = (hw < lower) | ((hw == lower) & (cr > 0.0))
        return np.where(j < 0, True, np.where(j >= npg, False, b))

    # arctan2 is within a wall of the count: one step up or down settles it
    c = np.where(below(c), c + 1, c)
    c = np.where(below(c - 1), c, c - 1)
    if not (np.all(below(c - 1)) and not np.any(below(c))):
        raise RuntimeError("azimuth count did not settle")
    ip = np.maximum(c, 1)
    keep = ir <= nrg - 1
    idx = (ir[keep] - 1) + nrg * ((it[keep] - 1) + ntg * (ip[keep] - 1))
    counts.reshape(-1, order="F")[:] += np.bincount(idx, minlength=nrg * ntg * npg)


def fractal_counts(rarr: np.ndarray, thetarr: np.ndarray, phiarr: np.ndarray, n: int = 32,
                   levels: int = 4, fl: float = 3.792, seed: int = 7,
                   device: Optional[int] = None) -> np.ndarray:
    """Points per cell, (nrg,ntg,npg) column-major, on the host or (with
    `device`) on that GPU; both give the same counts."""
    nrg, ntg, npg = len(rarr), len(thetarr), len(phiarr)
    if np.max(np.abs(phiarr - np.arange(npg) * (2.0 * PI / (npg - 1)))) > 1e-12:
        raise ValueError("the clump construction needs uniform phi walls over [0, 2 pi]")
    rl = 1.0 / fl
    scale_r = rarr[-1] * 2.0 / (1.0 + rl)
    mcos_t = -np.cos(thetarr)
    cph, sph = np.cos(phiarr), np.sin(phiarr)
    rng = np.random.default_rng(seed)
    if device is not None:
        from ._lib import fractal_counts_device
        st = rng.bit_generator.state["state"]
        return fractal_counts_device(rarr, mcos_t, cph, sph, n, levels,
                                     [rl ** k for k in range(1, levels + 1)], scale_r,
                                     int(st["state"]), int(st["inc"]), device)
    counts = np.zeros((nrg, ntg, npg), order="F")
    base = rng.random((n, 3)) - 0.5
    for parent in base:
        pts = parent[None, :]
        for k in range(1, levels + 1):
            off = (rng.random((pts.shape[0], n, 3)) - 0.5) * rl ** k
            pts = (pts[:, None, :] + off).reshape(-1, 3)
        _cell_counts_of(pts, rarr, mcos_t, cph, sph, scale_r, counts)
    return counts


def fractal_multiplier(rarr: np.ndarray, thetarr: np.ndarray, phiarr: np.ndarray, n: int = 32,
                       levels: int = 4, fl: float = 3.792, densratio: float = 0.75,
                       seed: int = 7, device: Optional[int] = None) -> np.ndarray:
    counts = fractal_counts(rarr, thetarr, phiarr, n, levels, fl, seed, device)
    cells = counts[:-1, :-1, :-1]
    av = cells.mean()
    mult = cells / av * densratio + 1.0 - densratio
    return mult
