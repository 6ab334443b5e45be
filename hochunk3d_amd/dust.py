"""Dust opacity tables for the transport path.

Host-side setup that produces the arrays the transport pass reads:

* ``synthetic_dust_table`` writes the 7-column dust file format the reference
  reads (``lambda[um], Cext, Csca, kappa, g, pmax, theta``; dust.f90:337).
* ``dustinit`` turns such a table into the opacity arrays of dust_mod
  (dust.f90:314-391): frequencies in eV (``nu = 12398/lambda[A]``), kappa
  normalised to its V-band value, albedo, g, pmax, plus the temperature and
  frequency grids of the emissivity tables.
* ``setemiss`` builds the normalised cumulative emissivity table
  ``kapBint(nT, nnu, ndg)`` and the Planck-mean opacity ``kappap(nT, ndg)``
  (dust.f90:394-474); ``kapdB_table`` builds the same call's
  ``kapdBint(nT, nnu, ndg)``, the cumulative of kappa_abs dB_nu/dT that the
  Bjorkman-Wood re-emission samples (dustfreq, dust.f90:158-311), and
  ``kappar_species`` its Rosseland mean ``kappar(nT, ndg)`` (the diffusion
  approximation's chiR).  The reference integrates each frequency bin with an
  adaptive Simpson rule to 1e-5; here each bin uses 8-point Gauss-Legendre
  quadrature, which is exact to far below that tolerance for this smooth
  integrand.  The tables are boundary inputs: the GPU path and the CPU oracle
  read the same arrays.
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import Optional

import numpy as np

NDG = 8
NLAMMAX = 1001
NT = 1000
NNU = 1025
TMIN_K, TMAX_K = 0.1, 2100.0          # dust.f90:319
NUMAX_T = 32.0                         # dust.f90:324 (units of kT)


def synthetic_dust_table(n: int = 200, lam_min: float = 0.01, lam_max: float = 1.0e4,
                         kappa_v: float = 200.0, albedo_max: float = 0.6, g_max: float = 0.6,
                         pmax: float = 0.5) -> np.ndarray:
    """The synthetic dust model of the benchmark configurations.

    ``n`` log-spaced wavelengths; kappa ~ lambda^-1 below 1 um and
    lambda^-1.7 beyond, scaled so kappa(0.55 um) ~ kappa_v; albedo and g fall
    linearly in log(lambda) from their maxima at lam_min to 0 at lam_max;
    constant peak polarisation.  Returns an (n, 7) array in the reference's
    column order.
    """
    lam = np.logspace(np.log10(lam_min), np.log10(lam_max), n)
    k1 = kappa_v * 0.55  # kappa(1um) such that kappa(0.55) = kappa_v for lambda^-1
    kap = np.where(lam <= 1.0, k1 / lam, k1 * lam ** -1.7)
    w = np.clip((np.log10(lam_max) - np.log10(lam)) / (np.log10(lam_max) - np.log10(lam_min)),
                0.0, 1.0)
    alb = albedo_max * w
    g = g_max * w
    cext = kap.copy()
    csca = alb * cext
    out = np.zeros((n, 7))
    out[:, 0] = lam
    out[:, 1] = cext
    out[:, 2] = csca
    out[:, 3] = kap
    out[:, 4] = g
    out[:, 5] = pmax
    return out


@dataclass
class DustTables:
    nlambda: np.ndarray        # (ndg,) int32
    nudust: np.ndarray         # (nlammax, ndg) F-order
    kappad: np.ndarray
    adust: np.ndarray
    gdust: np.ndarray
    pldust: np.ndarray
    kappav: np.ndarray         # (ndg,)
    lTint: np.ndarray          # (nT,)
    nuint: np.ndarray          # (nnu,)
    kapBint: np.ndarray        # (nT, nnu, ndg) F-order
    kappap: np.ndarray         # (nT, ndg) F-order
    kapdBint: Optional[np.ndarray] = None  # (nT, nnu, ndg): see ensure_kapdB
    kappar: Optional[np.ndarray] = None    # (nT, ndg) Rosseland mean: see ensure_kappar


def _species_arrays(table: np.ndarray):
    """dust.f90:330-374 for one species."""
    lam = table[:, 0]
    nl = len(lam)
    if nl > NLAMMAX:
        raise ValueError("too many wavelengths in dust opacity file")
    lamdust = 10000.0 * lam
    nud = 12398.0 / lamdust
    kd = table[:, 3].copy()
    ad = table[:, 2] / table[:, 1]
    gd = table[:, 4].copy()
    pd = table[:, 5].copy()
    kd[0] = float(np.float32(0.00001)) * kd[0]  # dust.f90:353, a default-real literal
    iv = int(np.nonzero(lam <= 0.55)[0][-1])  # 0-based index of the last lambda <= 0.55
    f = (5500.0 - lamdust[iv]) / (lamdust[iv + 1] - lamdust[iv])
    kv = (1.0 - f) * kd[iv] + f * kd[iv + 1]
    kd = kd / kv
    return nl, nud, kd, ad, gd, pd, kv


def opacset_np(nu: np.ndarray, nud: np.ndarray, kd: np.ndarray, ad: np.ndarray):
    """Vectorised opacset (dust.f90:13-60) for kappa and albedo."""
    x = nud[::-1]
    kap = np.interp(nu, x, kd[::-1])
    alb = np.interp(nu, x, ad[::-1])
    return kap, alb


def _planck_norm(nu: np.ndarray, T: np.ndarray) -> np.ndarray:
    """Normalised Planck function (B_nu/B)(kT/h), mathsubs.f90:1-24."""
    x = nu / T
    with np.errstate(divide="ignore", invalid="ignore"):
        b = 0.1539897 * x ** 3 / np.expm1(x)
    return np.where(x == 0.0, 0.0, b)


def _dplanck_norm(nu: np.ndarray, T: np.ndarray) -> np.ndarray:
    """Temperature derivative of the normalised Planck function,
    (dB_nu/dT)/(dB/dT) (kT/h) = (15/pi^4) x^4 e^x / (4 (e^x - 1)^2),
    mathsubs.f90:26-49."""
    x = nu / T
    with np.errstate(divide="ignore", invalid="ignore", over="ignore"):
        em = np.expm1(x)
        b = 0.1539897 * x ** 4 * np.exp(x) / (4.0 * em * em)
    return np.where((x == 0.0) | ~np.isfinite(b), 0.0, b)


def _cumulative_table(nud, kd, ad, lTint, nuint, weight, nquad: int = 8):
    """setemiss's cumulative bin integrals of kappa_abs * weight(nu, T) on
    the (nT, nnu) grid (frequencies in units of T), unnormalised, with the
    reference's cut of bins beyond nu0/T = 50 (dust.f90:425-436)."""
    xg, wg = np.polynomial.legendre.leggauss(nquad)
    fq = 0.5 * (xg + 1.0)
    wq = 0.5 * wg
    T = np.exp(lTint)[:, None]
    nu1 = nuint[None, :] * T
    nu0 = np.concatenate([np.zeros((len(lTint), 1)), nu1[:, :-1]], axis=1)
    dnu = nu1 - nu0
    dx = dnu / T
    k1, a1 = opacset_np(nu1, nud, kd, ad)
    kap1 = (1.0 - a1) * k1
    k0, a0 = opacset_np(nu0, nud, kd, ad)
    kap0 = (1.0 - a0) * k0
    acc = np.zeros_like(nu1)
    for f, w in zip(fq, wq):
        nu = nu0 + f * dnu
        acc += w * ((1.0 - f) * kap0 + f * kap1) * weight(nu, T)
    acc = np.where(nu0 / T < 50.0, acc, 0.0)
    return np.cumsum(dx * acc, axis=1)


def kapdB_species(nud, kd, ad, lTint, nuint, nquad: int = 8) -> np.ndarray:
    """kapdBint (nT, nnu) of one species: the cumulative of kappa_abs
    dB_nu/dT (setemiss, dust.f90:426-462), normalised to 1 at the last
    node."""
    cum = _cumulative_table(nud, kd, ad, lTint, nuint, _dplanck_norm, nquad)
    out = cum / cum[:, -1:]
    out[:, -1] = 1.0
    return out


def kapdB_table(dust: "DustTables") -> np.ndarray:
    """kapdBint (nT, nnu, ndg) for the species of `dust` (species with the
    same opacity columns share one computation)."""
    nT, nnu = dust.lTint.size, dust.nuint.size
    out = np.zeros((nT, nnu, NDG), order="F")
    cache = {}
    for id_ in range(NDG):
        nl = int(dust.nlambda[id_])
        cols = (dust.nudust[:nl, id_], dust.kappad[:nl, id_], dust.adust[:nl, id_])
        key = b"".join(np.ascontiguousarray(c).tobytes() for c in cols)
        if key not in cache:
            cache[key] = kapdB_species(*cols, dust.lTint, dust.nuint)
        out[:, :, id_] = cache[key]
    return out


def ensure_kapdB(dust: "DustTables") -> np.ndarray:
    """dust.kapdBint, built on first use (Bjorkman-Wood re-emission only)."""
    if dust.kapdBint is None:
        dust.kapdBint = kapdB_table(dust)
    return dust.kapdBint


def kappar_species(nud, kd, ad, lTint, nuint, nquad: int = 8) -> np.ndarray:
    """kappar (nT,) of one species, the Rosseland mean in units of kappa_V
    (setemiss, dust.f90:414-456): per frequency bin the integral of
    (dB_nu/dT) / chi_nu with the extinction chi interpolated linearly across
    the bin (chidB, dust.f90:515-527), bins beyond nu0/T = 50 cut, the bins
    summed, and the sum inverted."""
    xg, wg = np.polynomial.legendre.leggauss(nquad)
    fq = 0.5 * (xg + 1.0)
    wq = 0.5 * wg
    T = np.exp(lTint)[:, None]
    nu1 = nuint[None, :] * T
    nu0 = np.concatenate([np.zeros((len(lTint), 1)), nu1[:, :-1]], axis=1)
    dnu = nu1 - nu0
    dx = dnu / T
    c1, _ = opacset_np(nu1, nud, kd, ad)
    c0, _ = opacset_np(nu0, nud, kd, ad)
    acc = np.zeros_like(nu1)
    for f, w in zip(fq, wq):
        nu = nu0 + f * dnu
        acc += w * _dplanck_norm(nu, T) / ((1.0 - f) * c0 + f * c1)
    acc = np.where(nu0 / T < 50.0, acc, 0.0)
    return 1.0 / np.sum(dx * acc, axis=1)


def ensure_kappar(dust: "DustTables") -> np.ndarray:
    """dust.kappar (nT, ndg), built on first use (the diffusion
    approximation's chiR only)."""
    if dust.kappar is None:
        nT = dust.lTint.size
        out = np.zeros((nT, NDG), order="F")
        cache = {}
        for id_ in range(NDG):
            nl = int(dust.nlambda[id_])
            cols = (dust.nudust[:nl, id_], dust.kappad[:nl, id_], dust.adust[:nl, id_])
            key = b"".join(np.ascontiguousarray(c).tobytes() for c in cols)
            if key not in cache:
                cache[key] = kappar_species(*cols, dust.lTint, dust.nuint)
            out[:, id_] = cache[key]
        dust.kappar = out
    return dust.kappar


def setemiss_species(nud, kd, ad, lTint, nuint, nquad: int = 8, device=None):
    """kapBint (nT, nnu) normalised and kappap (nT,) for one species; with
    `device`, computed by hc_build_emissivity on that GPU."""
    xg, wg = np.polynomial.legendre.leggauss(nquad)
    fq = 0.5 * (xg + 1.0)
    wq = 0.5 * wg
    if device is not None:
        from ._lib import build_emissivity
        return build_emissivity(nud, kd, ad, lTint, nuint, fq, wq, device)
    T = np.exp(lTint)[:, None]                      # (nT,1)
    nu1 = nuint[None, :] * T                        # (nT,nnu)
    nu0 = np.concatenate([np.zeros((len(lTint), 1)), nu1[:, :-1]], axis=1)
    dnu = nu1 - nu0
    dx = dnu / T
    k1, a1 = opacset_np(nu1, nud, kd, ad)
    kap1 = (1.0 - a1) * k1
    k0, a0 = opacset_np(nu0, nud, kd, ad)
    kap0 = (1.0 - a0) * k0
    acc = np.zeros_like(nu1)
    for f, w in zip(fq, wq):
        nu = nu0 + f * dnu
        acc += w * ((1.0 - f) * kap0 + f * kap1) * _planck_norm(nu, T)
    cum = np.cumsum(dx * acc, axis=1)
    kappap = cum[:, -1].copy()
    kapB = cum / cum[:, -1:]
    kapB[:, -1] = 1.0
    return kapB, kappap


def dustinit(tables, device=None) -> DustTables:
    """Build dust_mod's arrays from ndg tables (a list of (n,7) arrays, or one
    table used for every species); `device` computes the emissivity tables
    on that GPU."""
    if isinstance(tables, np.ndarray):
        tables = [tables] * NDG
    if len(tables) != NDG:
        raise ValueError(f"need {NDG} dust tables")
    nlambda = np.zeros(NDG, np.int32)
    arrs = {k: np.zeros((NLAMMAX, NDG), order="F") for k in ("nud", "kd", "ad", "gd", "pd")}
    kappav = np.zeros(NDG)
    lTint = np.log(TMIN_K / 11605.0) + np.arange(NT) * np.log(TMAX_K / TMIN_K) / (NT - 1)
    numin = NUMAX_T / 1024.0
    nuint = numin + np.arange(NNU) * (NUMAX_T - numin) / (NNU - 1)
    kapBint = np.zeros((NT, NNU, NDG), order="F")
    kappap = np.zeros((NT, NDG), order="F")
    cache = {}
    for id_, tab in enumerate(tables):
        nl, nud, kd, ad, gd, pd, kv = _species_arrays(np.asarray(tab, float))
        nlambda[id_] = nl
        arrs["nud"][:nl, id_] = nud
        arrs["kd"][:nl, id_] = kd
        arrs["ad"][:nl, id_] = ad
        arrs["gd"][:nl, id_] = gd
        arrs["pd"][:nl, id_] = pd
        kappav[id_] = kv
        key = id(tab)
        if key not in cache:
            cache[key] = setemiss_species(nud, kd, ad, lTint, nuint, device=device)
        kapBint[:, :, id_], kappap[:, id_] = cache[key]
    return DustTables(nlambda, arrs["nud"], arrs["kd"], arrs["ad"], arrs["gd"], arrs["pd"],
                      kappav, lTint, nuint, kapBint, kappap)
