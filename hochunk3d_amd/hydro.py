"""Hydro-like density input of the large 3-D model: a log-normal density
field and the reference's density file.

* ``lognormal_field`` is a log-normal multiplier exp(sigma g - sigma^2/2)
  of the cells: g a Gaussian random field with power spectrum P(k) ~ k^slope
  (k^-11/3, the 3-D Kolmogorov spectrum) on the cell lattice (ln r, theta,
  phi) -- scale-free in radius, periodic in phi -- normalised to zero mean
  and unit variance over the grid, drawn from NumPy's PCG64 stream of
  `seed`.  The multiplier's mean is 1 in the ensemble, its logarithm has
  standard deviation sigma.
* ``write_densarr_unf`` / ``read_densarr_unf`` are the density file of
  DENSINPUT=YES (gridsethydro.f90:346-362, instructions.txt:283-292): Fortran
  sequential unformatted records (4-byte little-endian length markers) of
  (nrg, ntg, npg, ndg) as int32, then rarr, thetarr, phiarr and
  densarr(nrg,ntg,npg,ndg) column-major in real*8.  The reference skips the
  first record and reads species 1..2 of the last; the envelope and cavity
  are computed analytically and added (gridsethydro.f90:1019-1031).
"""
from __future__ import annotations

import struct

import numpy as np

_MAX_RECORD = (1 << 31) - 1


def lognormal_field(shape, sigma: float = 1.5, slope: float = -11.0 / 3.0,
                    seed: int = 11) -> np.ndarray:
    """(nr, ntheta, nphi) float64 multiplier, C order over (ir, it, ip)."""
    shape = tuple(int(n) for n in shape)
    rng = np.random.default_rng(seed)
    white = rng.standard_normal(shape)
    f = np.fft.rfftn(white)
    k2 = np.zeros(f.shape)
    for ax, n in enumerate(shape):
        k = np.fft.rfftfreq(n) if ax == len(shape) - 1 else np.fft.fftfreq(n)
        s = [1] * len(shape)
        s[ax] = k.size
        k2 = k2 + (k * n).reshape(s) ** 2          # lattice wave numbers
    amp = np.zeros_like(k2)
    nz = k2 > 0
    amp[nz] = k2[nz] ** (slope / 4.0)              # sqrt(P(k)) = k^(slope/2)
    f *= amp
    g = np.fft.irfftn(f, s=shape, axes=tuple(range(len(shape))))
    g -= g.mean()
    g /= g.std()
    return np.exp(sigma * g - 0.5 * sigma * sigma)


def _record(f, payload: bytes) -> None:
    n = len(payload)
    if n > _MAX_RECORD:
        raise ValueError("a Fortran record above 2 GiB needs subrecords: write fewer species")
    f.write(struct.pack("<i", n))
    f.write(payload)
    f.write(struct.pack("<i", n))


def write_densarr_unf(path: str, rarr, thetarr, phiarr, densarr) -> None:
    """densarr.unf for DENSINPUT=YES: densarr (nrg, ntg, npg, ndg)."""
    d = np.asarray(densarr, dtype="<f8")
    nrg, ntg, npg, ndg = d.shape
    with open(path, "wb") as f:
        _record(f, struct.pack("<4i", nrg, ntg, npg, ndg))
        for a, n in ((rarr, nrg), (thetarr, ntg), (phiarr, npg)):
            a = np.asarray(a, dtype="<f8")
            assert a.shape == (n,)
            _record(f, a.tobytes())
        _record(f, np.asfortranarray(d).tobytes(order="F"))


def read_densarr_unf(path: str):
    """(rarr, thetarr, phiarr, densarr (nrg, ntg, npg, ndg) F order) of a
    densarr.unf; the record markers are checked."""
    def rec(f):
        head = f.read(4)
        if len(head) != 4:
            raise ValueError("densarr.unf: truncated")
        (n,) = struct.unpack("<i", head)
        body = f.read(n)
        tail = f.read(4)
        if len(body) != n or len(tail) != 4 or struct.unpack("<i", tail)[0] != n:
            raise ValueError("densarr.unf: bad record markers")
        return body
    with open(path, "rb") as f:
        nrg, ntg, npg, ndg = struct.unpack("<4i", rec(f))
        rarr = np.frombuffer(rec(f), "<f8").copy()
        thetarr = np.frombuffer(rec(f), "<f8").copy()
        phiarr = np.frombuffer(rec(f), "<f8").copy()
        d = np.frombuffer(rec(f), "<f8")
    if (rarr.size, thetarr.size, phiarr.size, d.size) != (nrg, ntg, npg, nrg * ntg * npg * ndg):
        raise ValueError("densarr.unf: record sizes disagree with the header")
    return rarr, thetarr, phiarr, d.reshape((nrg, ntg, npg, ndg), order="F").copy(order="F")
