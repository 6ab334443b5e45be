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
depth-first order; this construction draws them from a seeded NumPy
generator, so clump positions are statistically equivalent to, not identical
with, the reference's for the same seed.
"""
from __future__ import annotations

import numpy as np

PI = 3.1415926535897932384626433


def fractal_multiplier(rarr: np.ndarray, thetarr: np.ndarray, phiarr: np.ndarray, n: int = 32,
                       levels: int = 4, fl: float = 3.792, densratio: float = 0.75,
                       seed: int = 7) -> np.ndarray:
    nrg, ntg, npg = len(rarr), len(thetarr), len(phiarr)
    rng = np.random.default_rng(seed)
    rl = 1.0 / fl
    counts = np.zeros((nrg, ntg, npg))
    base = rng.random((n, 3)) - 0.5
    scale_r = rarr[-1] * 2.0 / (1.0 + rl)
    for parent in base:
        pts = parent[None, :]
        for k in range(1, levels + 1):
            off = (rng.random((pts.shape[0], n, 3)) - 0.5) * rl ** k
            pts = (pts[:, None, :] + off).reshape(-1, 3)
        x, y, z = pts[:, 0], pts[:, 1], pts[:, 2]
        r = np.sqrt(x * x + y * y + z * z)
        th = np.arccos(np.clip(z / r, -1.0, 1.0))
        ph = np.arctan2(y, x)
        ph = np.where(ph < 0.0, ph + 2.0 * PI, ph)
        r = r * scale_r
        ir = np.searchsorted(rarr, r, side="left")        # nr.f locate (1-based)
        ir = np.where(ir == 0, 1, ir)
        it = np.maximum(np.searchsorted(thetarr, th, side="left"), 1)
        ip = np.maximum(np.searchsorted(phiarr, ph, side="left"), 1)
        keep = ir <= nrg - 1
        idx = (ir[keep] - 1) + nrg * ((it[keep] - 1) + ntg * (ip[keep] - 1))
        counts += np.bincount(idx, minlength=nrg * ntg * npg).reshape((nrg, ntg, npg), order="F")
    cells = counts[:-1, :-1, :-1]
    av = cells.mean()
    mult = cells / av * densratio + 1.0 - densratio
    return mult
