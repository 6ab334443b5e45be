"""Stellar spectrum and limb-darkening tables of the stellar launch.

* ``read_atmosphere`` reads a model-atmosphere file in the format atmosinit
  reads (starbarb.f90:28-47): one header line, then ``lambda[um] F_nu`` rows.
* ``atmosinit`` builds the atmos_interp arrays from it (starbarb.f90:3-118):
  the rows reordered to increasing frequency ``atmosnu = 1.2398 / lambda``
  (eV), the cumulative flux ``hnuint`` by the trapezoid rule in 1/lambda,
  normalised to 1 at the highest frequency, and the base-10 logarithms
  ``lognu``, ``loghnu`` that atmosfreq interpolates in.  With
  ``iplanckst`` (CPLANCKST=YES) the fluxes are replaced by the normalised
  Planck function at tstar on the file's wavelengths first, as the
  reference does; its stellar photons then sample that table
  (newdisktherm.F90:283-287 sets iplanckst = 0 in the photon loop).
* ``limb_table`` is the Eddington limb-darkening table of ``table``
  (samptabl.f90:1-22): the cumulative probability (mu^3 + mu^2) / 2 of the
  emission cosine on ntab nodes, with its first node at mu = 0 and
  probability 0.

These arrays cross the C ABI as they are (hc_set_atmosphere,
hc_set_limb_table), so the GPU and the oracle sample the same tables.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

NHNUMAX = 60000  # atmos_interp_mod.f90:7


@dataclass
class Atmosphere:
    atmosnu: np.ndarray   # (nhnu,) eV, increasing
    hnuint: np.ndarray    # (nhnu,) cumulative, hnuint[0] = 0, hnuint[-1] = 1
    lognu: np.ndarray     # log10(atmosnu)
    loghnu: np.ndarray    # log10(hnuint), loghnu[0] = 0 (never read)

    @property
    def nhnu(self) -> int:
        return int(self.atmosnu.size)


def read_atmosphere(path: str):
    """(wavelength um, F_nu) rows of a model-atmosphere file: a header line,
    then two columns, read until the first row that does not parse."""
    wave, flux = [], []
    with open(path) as f:
        f.readline()
        for line in f:
            parts = line.split()
            try:
                w, fl = float(parts[0]), float(parts[1])
            except (IndexError, ValueError):
                break
            wave.append(w)
            flux.append(fl)
            if len(wave) >= NHNUMAX:
                raise ValueError("atmosinit: too many frequencies in model atmosphere")
    return np.array(wave), np.array(flux)


def bnu(nu, T):
    """The normalised Planck function (B_nu / B)(kT/h) of mathsubs.f90:4-22,
    nu and T in eV."""
    x = np.asarray(nu, float) / T
    with np.errstate(divide="ignore", invalid="ignore"):
        b = 0.1539897 * x * x * x / np.expm1(x)
    return np.where(x == 0.0, 0.0, b)


def atmosinit(wave, flux, tstar_ev: float = 0.0, iplanckst: bool = False) -> Atmosphere:
    """The atmos_interp arrays of starbarb.f90:3-118 from the file's rows."""
    wave = np.asarray(wave, float)
    f = np.asarray(flux, float).copy()
    if wave.size < 2:
        raise ValueError("atmosinit: need at least two rows")
    if iplanckst:
        f = bnu(1.2398 / wave, tstar_ev)
    winv = (1.0 / wave)[::-1].copy()
    f = f[::-1].copy()
    atmosnu = 1.2398 * winv
    lognu = np.log10(atmosnu)
    steps = 0.5 * (f[1:] + f[:-1]) * (winv[1:] - winv[:-1])
    hnuint = np.concatenate([[0.0], np.add.accumulate(steps)])
    fbol = hnuint[-1]
    if fbol == 0.0:
        raise ValueError("atmosinit: the atmosphere file is all zeros")
    hnuint = hnuint / fbol
    loghnu = np.zeros_like(hnuint)
    with np.errstate(divide="ignore"):
        loghnu[1:] = np.log10(hnuint[1:])
    if not np.all(np.diff(atmosnu) > 0.0):
        raise ValueError("atmosinit: wavelengths must increase")
    return Atmosphere(atmosnu, hnuint, lognu, loghnu)


def limb_table(ntab: int = 1000):
    """(xmu, prob) of samptabl.f90:1-22 for ntab nodes (tabl_mod, which sets
    ntab, is absent from the reference tree; ntab is the caller's)."""
    if ntab < 3:
        raise ValueError("limb table needs ntab >= 3")
    delmu = 1.0 / float(ntab)
    xmu = np.zeros(ntab)
    prob = np.zeros(ntab)
    i = np.arange(2, ntab + 1, dtype=float)
    xmu[1:] = delmu * i
    prob[1:] = 0.5 * (np.power(xmu[1:], 3.0) + np.power(xmu[1:], 2.0))
    return xmu, prob
