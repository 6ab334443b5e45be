"""The diffusion approximation's grid mask (gridset.f90:666-789).

Cells deep in the optically thick disk midplane are not walked: a packet
absorbed there (or an accretion packet born there) is re-emitted from the
surface of the layer (emit_common.f90:80-387, on the GPU and in the oracle).
This module builds the two per-cell arrays of that layer from the model's
parameters, as gridset does:

* ``diffus`` (nrg, ntg, npg) int8: 1 inside the layer, where the cell-centre
  |cos theta| lies below the disk's photosphere mu_disk and the radial
  Rosseland optical depth from the inner edge exceeds taudisk;
* ``diffdir`` (nrg, ntg, npg) int8: the way out, -1 radially inwards (where
  the radial depth is below the vertical one), -2 up (wall theta < pi/2),
  +2 down; 0 beyond the disk's outer radius.

The analytic disk of setup_wave.f90:60-115 sets the optical depths: rho0 from
the disk mass (the reference's normalisation with its (p + 1.5) term), the
radial V-band depth taur of each disk species, and the Rosseland means chiR
(dust.f90:83-111) of setemiss's kappar table (hochunk3d_amd.dust.kappar_species).
``ierfc`` is mathsubs.f90:49-77 (Abramowitz & Stegun 26.2.23).  The scalar
arithmetic uses Python floats and the C library's functions (math), like the
reference's real*8 code.
"""
from __future__ import annotations

import math
from typing import Tuple

import numpy as np

from .dust import ensure_kappar
from .model import MSOL, PI, RSOL, Model

PIHALF = PI * 0.5              # constants.f90:6
R2P = 2.0 * PI
TSUB_EV = 1600.0 / 11605.0


def ierfc(x: float) -> float:
    """Inverse complementary error function, the rational approximation of
    Abramowitz & Stegun 26.2.23 (mathsubs.f90:49-77); |error| < 4.5e-4 in the
    argument of the normal tail, for 0 < x < 1."""
    sqhalf = 0.70710678
    c0, c1, c2 = 2.515517, 0.802853, 0.010328
    d1, d2, d3 = 1.432788, 0.189269, 0.001308
    t = math.sqrt(2.0 * math.log(2.0 / x))
    return (t - (c0 + t * (c1 + t * c2)) / (1.0 + t * (d1 + t * (d2 + t * d3)))) * sqhalf


def chir(kappar: np.ndarray, lTint: np.ndarray, T: float, id_: int) -> float:
    """chiR (dust.f90:83-111): kappar (nT, ndg) at temperature T (eV),
    linear in ln T, the end values outside the table."""
    nT = lTint.size
    lT = math.log(T)
    k = int(float(nT - 1) * (lT - lTint[0]) / (lTint[nT - 1] - lTint[0])) + 1
    if k < 1 or k > nT - 1:
        return float(kappar[0 if k < 1 else nT - 1, id_])
    f = (lT - lTint[k - 1]) / (lTint[k] - lTint[k - 1])
    return (1.0 - f) * float(kappar[k - 1, id_]) + f * float(kappar[k, id_])


def disk_rho0(model: Model, id_: int) -> float:
    """rho0 of disk species id_ (0-based), setup_wave.f90:66-90 (g/cm^3 at
    the stellar radius in the midplane)."""
    cfg = model.cfg
    fm = float(model.fmass8[id_])
    if fm <= 0.0:
        return 0.0
    a, b = cfg.a_exp, cfg.b_exp
    p = b - a
    rmincgs = model.rstar * RSOL
    z1cgs = model.rstar * RSOL * float(model.z1_8[id_])
    const = R2P ** 1.5 * rmincgs ** 2 * z1cgs
    rmaxd, rmind = model.rmaxdmax, model.rddustmin
    mass = cfg.massd * MSOL
    if p != -2.0 and p != -1.5:
        rho0 = mass / const / (1.0 / (p + 2.0) * (rmaxd ** (p + 2.0) - rmind ** (p + 2.0))
                               - 1.0 / (p + 1.5) * (rmaxd ** (p + 1.5) - rmind ** (p + 1.5)))
    elif p == -2.0:
        rho0 = mass / const / (math.log(rmaxd / rmind)
                               - 1.0 / (p + 1.5) * (rmaxd ** (p + 1.5) - rmind ** (p + 1.5)))
    else:
        rho0 = mass / const / (1.0 / (p + 2.0) * (rmaxd ** (p + 2.0) - rmind ** (p + 2.0))
                               - math.log(rmaxd / rmind))
    return rho0 * fm


def disk_taur(model: Model, id_: int) -> float:
    """Radial midplane V-band optical depth of disk species id_ from its
    inner to its outer radius (setup_wave.f90:91-115)."""
    rho0 = disk_rho0(model, id_)
    if rho0 == 0.0:
        return 0.0
    a = model.cfg.a_exp
    a1 = 1.0 - a
    c1 = float(model.dust.kappav[id_]) * rho0 * (model.rstar * RSOL)
    rmaxd, rdd = model.rmaxdmax, model.rddustmin
    if a != 1.0 and a != 0.5:
        return (c1 / a1 * (rmaxd ** a1 - rdd ** a1)
                - c1 / (0.5 - a) * (rmaxd ** (0.5 - a) - rdd ** (0.5 - a)))
    if a == 1.0:
        return (c1 * math.log(rmaxd / rdd)
                - c1 / (0.5 - a) * (rmaxd ** (0.5 - a) - rdd ** (0.5 - a)))
    return c1 / a1 * (rmaxd ** a1 - rdd ** a1) - c1 * math.log(rmaxd / rdd)


def diffusion_mask(model: Model) -> Tuple[np.ndarray, np.ndarray]:
    """(diffus, diffdir) of gridset.f90:666-789 for the model's disk: int8
    arrays (nrg, ntg, npg), column-major, nonzero only in cells of the
    grid."""
    from .output import cell_centres
    cfg = model.cfg
    nrg, ntg, npg = model.nrg, model.ntg, model.npg
    diffus = np.zeros((nrg, ntg, npg), np.int8, order="F")
    diffdir = np.zeros((nrg, ntg, npg), np.int8, order="F")
    if cfg.massd <= 0.0:
        return diffus, diffdir
    kappar = ensure_kappar(model.dust)
    lTint = model.dust.lTint
    rdd = model.rddustmin                 # rddust(1) = rddust(2) = rddustmin here
    rmaxd = model.rmaxdmax
    a, b = cfg.a_exp, cfg.b_exp
    mu0 = tauv = tauzrat = tauros = 0.0
    for id_ in range(2):                                                  # :677-685
        mu0tmp = float(model.fmass8[id_]) * float(model.z1_8[id_]) * rdd ** b / rdd
        mu0 = mu0 + mu0tmp
        tr = disk_taur(model, id_)
        tauv = tauv + tr
        a1 = 1.0 - a
        tauzrat = tauzrat + math.sqrt(PIHALF) * a1 * mu0tmp / ((rmaxd / rdd) ** a1 - 1.0)
        tauros = tauros + tr * chir(kappar, lTint, TSUB_EV, id_)
    taurth = tauzrat * tauros
    taudisk = min(10.0, taurth)
    nr0 = 1 if rdd == model.rarr[0] else 3                                  # :688-694
    rave, tave, _ = cell_centres(model)
    tstar_k = cfg.tstar
    fm1 = float(model.fmass8[0])
    for ir in range(nr0, nrg):                                            # 1-based ir
        r = rave[ir - 1] / rdd
        if r * rdd > rmaxd:
            break
        rs = 1.0 / (rdd * r)
        tdisk = min(TSUB_EV, max(3.0 / 11605.0, (tstar_k / 11605.0) * (max(
            2.0 / 3.0 * rs ** 3, math.asin(rs) - rs * math.sqrt(1.0 - rs ** 2)) / PI) ** 0.25))
        sigmu = fm1 * mu0 * r ** (b - 1.0)                                  # :722-726
        taumid = fm1 * tauzrat * tauv * chir(kappar, lTint, tdisk, 0) / r ** (a - b)
        a1 = 1.0 - a
        mudisk = sigmu * math.sqrt(2.0) * ierfc(taudisk / taumid) if taumid > taudisk else 0.0
        for it in range(1, ntg):
            mu = abs(math.cos(tave[it - 1]))
            taurmu = taumid * math.erfc(mu / (sigmu * math.sqrt(2.0)))
            taurr = (tauros * math.exp(-0.5 * mu * mu / (sigmu * sigmu))
                     * (1.0 - r ** a1) / (1.0 - (rmaxd / rdd) ** a1))
            dif = 1 if (mu < mudisk and taurr > taudisk) else 0
            if taurr < taurmu:
                dd = -1
            else:
                dd = -2 if model.thetarr[it - 1] < PIHALF else 2
            diffus[ir - 1, it - 1, :npg - 1] = dif
            diffdir[ir - 1, it - 1, :npg - 1] = dd
    return diffus, diffdir
