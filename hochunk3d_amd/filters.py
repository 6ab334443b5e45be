"""Filter response curves for the filter-convolved peel-off images.

``filt`` restates filt_nu.f90:1-118: each filter file is a header line then
rows of wavelength (microns) and response; the wavelengths become frequencies
``filtwave = 2.9979e14 / w`` (Hz, increasing down the file, whose wavelengths
decrease), and the response is divided by its trapezoidal integral over
frequency, so that every curve integrates to 1.  The arrays have filt_mod's
layout: ``nwfilt[nfilt]`` counts the file's lines (header included, as
filt_mod.f90:17 lists them), ``filtwave`` and ``filtphi`` are
(nwfiltmax, nfilt) column-major.

The peel-off tally (peeloff_3d.f90:182-211) weights a ray at frequency nu
by the response at the last node below nu, for every filter whose node range
holds nu strictly; ``filter_weight`` is that rule, for the tests.
"""
from __future__ import annotations

import os
from dataclasses import dataclass
from typing import Optional, Sequence, Tuple

import numpy as np

FACT = 2.9979e14  # c in micron/s (filt_nu.f90:13)
NWFILTMAX = 1200  # filt_mod.f90:8
# filt_mod.f90:19-24
FILTNAMES = ("VV", "RR", "II", "2J", "2H", "2K", "I1", "I2", "I3", "I4", "M1", "M2", "M3",
             "H1", "H2", "H3", "H4", "H5", "H6", "Z1", "Z2")
# filt_mod.f90:17: the line count of each file, header included; filt reads
# nwfilt - 1 rows (one row short of the 2K file's 76)
NWFILT = (25, 25, 24, 108, 59, 76, 506, 428, 371, 422, 129, 112, 401, 69, 98, 123, 125, 179,
          135, 7, 7)


@dataclass
class Filters:
    names: Tuple[str, ...]
    nwfilt: np.ndarray     # (nfilt,) int32: nodes + 1 (filt_mod's line counts)
    filtwave: np.ndarray   # (nwfiltmax, nfilt) F-order, Hz, increasing per filter
    filtphi: np.ndarray    # (nwfiltmax, nfilt) F-order, response / its integral over Hz

    @property
    def nfilt(self) -> int:
        return len(self.names)


def read_filter_file(path: str, nrows: Optional[int] = None) -> Tuple[np.ndarray, np.ndarray]:
    """(wavelength um, response) of the first nrows rows of a filter file
    (all rows when nrows is None) after its header line."""
    with open(path) as f:
        f.readline()
        rows = [ln.split() for ln in f if ln.strip()]
    if nrows is not None:
        if len(rows) < nrows:
            raise ValueError(f"{path}: {len(rows)} rows, {nrows} needed")
        rows = rows[:nrows]
    a = np.array([[float(r[0]), float(r[1])] for r in rows], dtype=np.float64)
    return a[:, 0], a[:, 1]


def filt(curves: Sequence[Tuple[str, np.ndarray, np.ndarray]],
         nwfiltmax: int = NWFILTMAX) -> Filters:
    """filt_nu.f90:1-118 on (name, wavelength um, response) curves."""
    nfilt = len(curves)
    fw = np.zeros((nwfiltmax, nfilt), order="F")
    fp = np.zeros((nwfiltmax, nfilt), order="F")
    nw = np.zeros(nfilt, np.int32)
    for k, (_, w, phi) in enumerate(curves):
        nf = len(w)
        if nf < 2 or nf > nwfiltmax:
            raise ValueError("a filter needs 2..nwfiltmax nodes")
        fw[:nf, k] = FACT / np.asarray(w, np.float64)
        fp[:nf, k] = np.asarray(phi, np.float64)
        fint = 0.0
        for i in range(nf - 1):  # the reference's summation order
            fint = 0.5 * (fp[i + 1, k] + fp[i, k]) * (fw[i + 1, k] - fw[i, k]) + fint
        fp[:nf, k] = fp[:nf, k] / fint
        nw[k] = nf + 1
    return Filters(tuple(c[0] for c in curves), nw, fw, fp)


def reference_filters(filter_dir: str, names: Optional[Sequence[str]] = None) -> Filters:
    """The filters of filt_mod.f90 found in filter_dir (<name>.txt), each
    read to filt_mod's line count."""
    curves = []
    for n in (names or FILTNAMES):
        p = os.path.join(filter_dir, f"{n}.txt")
        if os.path.exists(p):
            w, phi = read_filter_file(p, NWFILT[FILTNAMES.index(n)] - 1)
            curves.append((n, w, phi))
    return filt(curves)


def synthetic_filters(centres_um: Sequence[float] = (0.55, 2.2, 10.0, 70.0),
                      width: float = 0.15, n: int = 60) -> Filters:
    """Box-like filters with smooth edges around the given wavelengths
    (relative width `width`), wavelengths decreasing down the table like the
    reference's files."""
    curves = []
    for c in centres_um:
        w = np.linspace(c * (1 + 1.5 * width), c * (1 - 1.5 * width), n)
        x = (w / c - 1.0) / width
        phi = 1.0 / (1.0 + np.exp(8.0 * (np.abs(x) - 1.0)))
        curves.append((f"S{len(curves) + 1}", w, phi))
    return filt(curves)


def filter_weight(f: Filters, k: int, freq: float) -> Optional[float]:
    """peeloff_3d.f90:188-200: the weight of a ray of frequency freq (Hz) in
    filter k, or None when freq is outside the filter's node range."""
    nf = int(f.nwfilt[k]) - 1
    fw = f.filtwave[:nf, k]
    if not (freq > fw[0] and freq < fw[nf - 1]):
        return None
    iw = 0
    while freq > fw[iw]:
        iw += 1
    return float(f.filtphi[iw - 1, k])


def ray_frequency(nub: float) -> float:
    """The peel ray's frequency in Hz from its energy in eV, in the
    reference's two steps (wave = 1.2398/nub, newdisktherm.F90:306;
    freq = 2.9979e14/wave, peeloff_3d.f90:180)."""
    wave = 1.2398 / nub
    return FACT / wave
