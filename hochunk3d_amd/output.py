"""Output files of a run, written from the transport tallies in the layouts
and formats of the reference's ASCII build (output.F90, flux.f90).

* `diskflux` restates flux.f90:23-121: Monte Carlo errors from the second
  moments and counts, then the normalisation of the SEDs to L_sun and of the
  image cube to intensity.  The reference keeps these arrays in real*4
  (initarr.f90:30-40, 110-125) and does this arithmetic in real*4; so does
  `diskflux` (NumPy float32, one rounding per operation, in the reference's
  order of evaluation).  The tallies themselves are FP64 sums narrowed once.
* `write_flux`, `write_peel_flux` write flux*.dat and peel_flux*.dat
  (output.F90:233-329); `write_grid_unf` the Fortran unformatted grid files
  (output.F90:38-158); `write_tmidplane`, `write_tave` the midplane and
  shell-averaged temperatures (output.F90:466-501).
* `write_image_cube` has no counterpart in the ASCII build, which never
  writes the IMCUBE cube (only output_FITS.F90 does, through cfitsio): it
  writes the normalised cube as one Fortran unformatted file.
* `write_filter_images` writes the filter-convolved peel images
  e_<filter>_<theta>_<phi>_{I,Q,U,V}_img.dat (output.F90:380-411), after
  `diskflux` has divided Q, U, V by I and put I in intensity units
  (flux.f90:87-102).
* `write_run_outputs` writes what output() writes for one pass.

Formatted output follows the Fortran edit descriptors (Fw.d, ESw.d, Iw) and
gfortran's list-directed layout (one leading blank; integers right-justified
in 11 columns, one blank between items).  Unformatted files are gfortran's
sequential records: a 4-byte length, the data, the length again, native
(little-endian) byte order.
"""
from __future__ import annotations

import math
import os
from dataclasses import dataclass
from typing import Optional, Sequence

import numpy as np

from .model import LSOL, PI, Model

RAD2DEG = 180.0 / PI             # constants.f90:9
PC2CM = 3.0857e18                # constants.f90:26
C_CGS = 2.99792458e10            # constants.f90:22
EV_TO_K = 11605.0                # kT in eV to K (output.F90:474,498)
NUMAX = 1.2398 / 0.01            # setup.f90:39 (lambda = 0.01 micron)

# file-name suffix of each origin index io (output.F90:223-230)
SUFFIX = ("", "_star", "_disk", "_enve", "_outi", "_dire", "_scat", "_ther")

_F4 = np.float32


# --------------------------------------------------------------------------
# Fortran edit descriptors
# --------------------------------------------------------------------------

def _inf_nan(x: float, w: int) -> str:
    """gfortran's text for a non-finite value in a field of width w."""
    if math.isnan(x):
        s = "NaN"
    elif x > 0:
        s = "Infinity" if w >= 8 else "Inf"
    else:
        s = "-Infinity" if w >= 9 else "-Inf"
    return s.rjust(w) if len(s) <= w else "*" * w


def fmt_f(x: float, w: int, d: int) -> str:
    """Fw.d: fixed point, d decimals, right-justified; asterisks if it does
    not fit (the leading zero of |x| < 1 is dropped first)."""
    x = float(x)
    if not math.isfinite(x):
        return _inf_nan(x, w)
    s = f"{x:.{d}f}"
    if len(s) > w and (s.startswith("0.") or s.startswith("-0.")):
        s = s.replace("0.", ".", 1)
    return s.rjust(w) if len(s) <= w else "*" * w


def fmt_es(x: float, w: int, d: int) -> str:
    """ESw.d: scientific, one digit before the point, d after; exponent
    E+dd, or +ddd without the letter when |exponent| > 99."""
    x = float(x)
    if not math.isfinite(x):
        return _inf_nan(x, w)
    mant, ex = f"{x:.{d}E}".split("E")
    e = int(ex)
    if abs(e) <= 99:
        s = f"{mant}E{e:+03d}"
    elif abs(e) <= 999:
        s = f"{mant}{e:+04d}"
    else:
        return "*" * w
    return s.rjust(w) if len(s) <= w else "*" * w


def fmt_i(n: int, w: int) -> str:
    """Iw: integer right-justified in w columns."""
    s = str(int(n))
    return s.rjust(w) if len(s) <= w else "*" * w


def list_directed(*items) -> str:
    """One record of gfortran list-directed output of integers and strings."""
    out = []
    prev_char = False
    for k, it in enumerate(items):
        is_char = isinstance(it, str)
        if k == 0:
            out.append(" ")
        elif not (is_char and prev_char):
            out.append(" ")
        out.append(it if is_char else fmt_i(it, 11))
        prev_char = is_char
    return "".join(out)


def _flux_row(lam: float, vals: Sequence[float]) -> str:
    """'(F10.4,8(2X,ES11.4))' (output.F90:266,318)."""
    return fmt_f(lam, 10, 4) + "".join("  " + fmt_es(v, 11, 4) for v in vals)


# '("#",1X,"Lambda",9X,"I",10X,"I_err",9X,"Q/I",7X,"(Q/I)_err",7X,"U/I",7X,
#   "(U/I)_err",7X,"V/I",7X,"(V/I)_err")' (output.F90:260,314)
_FLUX_HEADER = ("# Lambda" + " " * 9 + "I" + " " * 10 + "I_err" + " " * 9 + "Q/I" + " " * 7
                + "(Q/I)_err" + " " * 7 + "U/I" + " " * 7 + "(U/I)_err" + " " * 7 + "V/I"
                + " " * 7 + "(V/I)_err")


# --------------------------------------------------------------------------
# flux.f90: errors and normalisation
# --------------------------------------------------------------------------

@dataclass
class Fluxes:
    """The SED, peel-off and cube arrays after diskflux, real*4 like the
    reference's.  si..sv: (nfreq,nmu,nph,nap,no) normalised Stokes sums;
    si2..sv2 their errors (si2 relative, sq2..sv2 in units of I);
    ti..tv, ti2..tv2: (nfreq,npeel,nap,no) likewise for the peeled SED;
    image_m_*: the cube, I in intensity units, Q/U/V as fractions of I."""
    nunorm: float
    si: np.ndarray
    sq: np.ndarray
    su: np.ndarray
    sv: np.ndarray
    si2: np.ndarray
    sq2: np.ndarray
    su2: np.ndarray
    sv2: np.ndarray
    si_init: Optional[np.ndarray] = None
    ti: Optional[np.ndarray] = None
    tq: Optional[np.ndarray] = None
    tu: Optional[np.ndarray] = None
    tv: Optional[np.ndarray] = None
    ti2: Optional[np.ndarray] = None
    tq2: Optional[np.ndarray] = None
    tu2: Optional[np.ndarray] = None
    tv2: Optional[np.ndarray] = None
    image_m_i: Optional[np.ndarray] = None
    image_m_q: Optional[np.ndarray] = None
    image_m_u: Optional[np.ndarray] = None
    image_m_v: Optional[np.ndarray] = None
    image_b_i: Optional[np.ndarray] = None
    image_b_q: Optional[np.ndarray] = None
    image_b_u: Optional[np.ndarray] = None
    image_b_v: Optional[np.ndarray] = None


def nunorm_of(nfreq: int, numin: float, numax: float = NUMAX) -> float:
    """Width of one SED bin in ln(nu) units, nurat^(1/2n) - nurat^(-1/2n)
    (setup.f90:41-43)."""
    nurat = numax / numin
    return nurat ** (0.5 / float(nfreq)) - nurat ** (-0.5 / float(nfreq))


def output_wavelengths(nfreq: int, numin: float, numax: float = NUMAX) -> np.ndarray:
    """Bin-centre wavelengths (micron) of the written SEDs, output.F90:233-235:
    the exponent (real(inub)-0.5)/nfreq is real*4 there."""
    nurat = numax / numin
    inub = np.arange(1, nfreq + 1, dtype=_F4)
    e32 = (inub - _F4(0.5)) / _F4(nfreq)
    return 1.2398 / (numin * nurat ** e32.astype(np.float64))


def _bin_frequencies(nfreq: int, numin: float, numax: float) -> np.ndarray:
    """freqarr of flux.f90:41-44 (Hz), for bins 1..nfreq+1."""
    nurat = numax / numin
    inub = np.arange(1, nfreq + 2, dtype=np.float64)
    wav = 1.2398 / (numin * nurat ** ((inub - 0.5) / float(nfreq)))
    return C_CGS * 1.0e4 / wav


def _sed_error(si, x, x2, nums, first: bool):
    """flux.f90:47-56: where(nums > 1.) si2 = sqrt((si2-si**2/nums)/(nums-1.)
    *nums)/si+1./nums (and without +1./nums for Q,U,V, over si)."""
    one = _F4(1.0)
    v = np.sqrt((x2 - x * x / nums) / (nums - one) * nums) / si
    if first:
        v = v + one / nums
    return np.where(nums > one, v, _F4(0.0)).astype(_F4)


def diskflux(model: Model, np_: float, sed: np.ndarray, peel: Optional[np.ndarray] = None,
             si_init: Optional[np.ndarray] = None, cube=None,
             numax: float = NUMAX, filt=None) -> Fluxes:
    """flux.f90:23-121 on one pass's tallies.

    sed: (9, nfreq, nmu, nph, nap, no) I,Q,U,V,I^2,Q^2,U^2,V^2,N sums;
    peel: (9, nfreq, npeel, nap, no) likewise; si_init: (nfreq, 5);
    cube: (img_i (npeel,nx,nx,nfreq,no), img_q, img_u, img_v (npeel,nx,nx,nfreq));
    filt: the filter images (image_b_i, q, u, v), each (npeel,nx,nx,nfilt).
    np_ is the pass's packet count (rnp)."""
    sed = np.asarray(sed)
    nfreq, nmu, nph = sed.shape[1:4]
    rnp = float(np_)
    ltot = float(model.ltot)
    rnorm = rnp / float(nmu) / float(nph) / (ltot / LSOL)                     # :29
    image_norm = rnp / (ltot / LSOL)                                         # :30
    f = [a.astype(_F4) for a in sed]
    si, sq, su, sv, si2, sq2, su2, sv2, nums = f
    with np.errstate(all="ignore"):
        e = [_sed_error(si, x, x2, nums, k == 0)
             for k, (x, x2) in enumerate(((si, si2), (sq, sq2), (su, su2), (sv, sv2)))]
        r = _F4(rnorm)
        out = Fluxes(nunorm_of(nfreq, model.numin, numax), si / r, sq / r, su / r, sv / r, *e)
        if si_init is not None:
            out.si_init = np.asarray(si_init).astype(_F4) / _F4(image_norm)  # :66
        if peel is not None:
            ti, tq, tu, tv, ti2, tq2, tu2, tv2, numt = [a.astype(_F4) for a in np.asarray(peel)]
            one = _F4(1.0)

            def perr(x, x2):                                                 # :70-78
                v = np.sqrt((x2 * numt / (x * x) - one) / (numt - one))
                return np.where(numt > one, v, _F4(0.0)).astype(_F4)

            q = _F4(image_norm)
            out.ti2, out.tq2, out.tu2, out.tv2 = (perr(ti, ti2), perr(tq, tq2), perr(tu, tu2),
                                                  perr(tv, tv2))
            out.ti, out.tq, out.tu, out.tv = ti / q, tq / q, tu / q, tv / q       # :80-83
        def image_norm_intensity(nx):                                        # :37
            return (rnp / ltot * 4.0 * PI / 1.0e17
                    * (2.0 * model.rmaxi / model.autors / float(nx)) ** 2
                    * PC2CM ** 2 / 206265.0 ** 2)

        if filt is not None:
            bi, bq, bu, bv = (np.asarray(a).astype(_F4) for a in filt)
            pos = bi > _F4(0.0)
            out.image_b_q = np.where(pos, bq / bi, _F4(0.0)).astype(_F4)     # :87-98
            out.image_b_u = np.where(pos, bu / bi, _F4(0.0)).astype(_F4)
            out.image_b_v = np.where(pos, bv / bi, _F4(0.0)).astype(_F4)
            out.image_b_i = bi / _F4(image_norm_intensity(bi.shape[1]))      # :101
        if cube is not None:
            mi, mq, mu_, mv = (np.asarray(a).astype(_F4) for a in cube)
            nx = mi.shape[1]
            inorm = image_norm_intensity(nx)
            i1 = mi[..., 0]
            pos = i1 > _F4(0.0)
            mq = np.where(pos, mq / i1, _F4(0.0)).astype(_F4)                # :107-115
            mu_ = np.where(pos, mu_ / i1, _F4(0.0)).astype(_F4)
            mv = np.where(pos, mv / i1, _F4(0.0)).astype(_F4)
            # :117-119 divides bin inub by freqarr(inub+1)-freqarr(inub); for
            # the last bin that index is past the end of freqarr, so the
            # geometric sequence is continued by one bin here
            fr = _bin_frequencies(nfreq, model.numin, numax)
            dnu = (fr[1:] - fr[:-1]).astype(_F4)
            mi = (mi / dnu[None, None, None, :, None]) / _F4(inorm)
            out.image_m_i, out.image_m_q, out.image_m_u, out.image_m_v = mi, mq, mu_, mv
    return out


# --------------------------------------------------------------------------
# ASCII writers (output.F90)
# --------------------------------------------------------------------------

def _stokes_columns(i, q, u, v, ie, qe, ue, ve, nunorm):
    """output.F90:239-251 (and 282-300): where I is zero all eight columns
    are zero; else I/nunorm (real*8), Q/I, U/I, V/I (real*4 quotients),
    I_err*I/nunorm, and the Q, U, V errors as they are."""
    with np.errstate(all="ignore"):
        zero = i == _F4(0.0)
        cols = (i.astype(np.float64) / nunorm,
                (ie * i).astype(np.float64) / nunorm,
                (q / i).astype(np.float64), qe.astype(np.float64),
                (u / i).astype(np.float64), ue.astype(np.float64),
                (v / i).astype(np.float64), ve.astype(np.float64))
        cols = [np.where(zero, 0.0, c) for c in cols]
    # column order: I, I_err, Q/I, (Q/I)_err, U/I, (U/I)_err, V/I, (V/I)_err
    return cols


def write_flux(dirpath: str, fl: Fluxes, numin: float, numax: float = NUMAX) -> list:
    """flux<suffix>.dat for each origin io (output.F90:233-276).  Returns the
    paths written."""
    nfreq, nmu, nph, nap, no = fl.si.shape
    wav = output_wavelengths(nfreq, numin, numax)
    cols = _stokes_columns(fl.si, fl.sq, fl.su, fl.sv, fl.si2, fl.sq2, fl.su2, fl.sv2,
                           fl.nunorm)
    paths = []
    for io in range(no):
        path = os.path.join(dirpath, f"flux{SUFFIX[io]}.dat")
        lines = [list_directed(nfreq, nph, nmu, nap, "   nfreq,nph,nmu,nap"), _FLUX_HEADER]
        for ia in range(nap):
            for imu in range(nmu):
                for ip in range(nph):
                    for inub in range(nfreq):
                        lines.append(_flux_row(wav[inub],
                                               [c[inub, imu, ip, ia, io] for c in cols]))
        with open(path, "w") as fh:
            fh.write("\n".join(lines) + "\n")
        paths.append(path)
    return paths


def angle_label(rad: float) -> str:
    """trim(adjustl(x)) of x written with F5.1 in degrees (output.F90:305-306)."""
    return fmt_f(rad * RAD2DEG, 5, 1).strip()


def write_peel_flux(dirpath: str, fl: Fluxes, thete: Sequence[float], phie: Sequence[float],
                    numin: float, numax: float = NUMAX, partial: bool = False) -> list:
    """peel_flux<suffix>[_partial]_<theta>_<phi>.dat for each origin and
    observer (output.F90:280-329); thete/phie are the observers' angles in
    radians."""
    if fl.ti is None:
        raise ValueError("no peel-off tallies (run the pass with ipeel=1)")
    nfreq, npeel, nap, no = fl.ti.shape
    if len(thete) < npeel or len(phie) < npeel:
        raise ValueError("one (theta, phi) per observer is needed")
    wav = output_wavelengths(nfreq, numin, numax)
    cols = _stokes_columns(fl.ti, fl.tq, fl.tu, fl.tv, fl.ti2, fl.tq2, fl.tu2, fl.tv2, fl.nunorm)
    psuf = "_partial" if partial else ""
    paths = []
    for io in range(no):
        for pk in range(npeel):
            name = (f"peel_flux{SUFFIX[io]}{psuf}_{angle_label(thete[pk])}"
                    f"_{angle_label(phie[pk])}.dat")
            path = os.path.join(dirpath, name)
            lines = [list_directed(nfreq, npeel, nap, "   nfreq,npeel,nap"), _FLUX_HEADER]
            for ia in range(nap):
                for inub in range(nfreq):
                    lines.append(_flux_row(wav[inub], [c[inub, pk, ia, io] for c in cols]))
            with open(path, "w") as fh:
                fh.write("\n".join(lines) + "\n")
            paths.append(path)
    return paths


def cell_centres(model: Model):
    """ravearr, thetavearr, phiavearr (gridset.f90:493-498, 606-619, 652-655):
    geometric mean of the radial walls, (mean of theta^2)^(1/2) (the real*4
    exponent 1./pexp is exactly 0.5; x**0.5 is pow there), mean of the phi
    walls."""
    r = np.log10(model.rarr)
    rave = 10.0 ** (0.5 * (r[:-1] + r[1:]))
    th2 = model.thetarr * model.thetarr
    tave = np.power(0.5 * (th2[:-1] + th2[1:]), np.full(model.ntg - 1, 0.5))
    pave = 0.5 * (model.phiarr[:-1] + model.phiarr[1:])
    return rave, tave, pave


def write_tmidplane(dirpath: str, model: Model, tdust: np.ndarray, tdust2: np.ndarray,
                    tauarr: Optional[np.ndarray] = None) -> list:
    """tmidplane<id>.dat (output.F90:466-478): the temperatures of the cells
    at theta index (ntg+1)/2 and phi index 1, in K.  tauarr (nrg) is a
    column the reference allocates but never assigns in its tree
    (grid_mod.f90:89); zeros unless given."""
    rave, _, _ = cell_centres(model)
    it = (model.ntg + 1) // 2 - 1
    tau = np.zeros(model.nrg) if tauarr is None else np.asarray(tauarr, np.float64)
    paths = []
    for id_ in range(tdust.shape[3]):
        path = os.path.join(dirpath, f"tmidplane{id_ + 1}.dat")
        lines = [fmt_f(model.thetarr[it] * RAD2DEG, 9, 5)
                 + " / theta angle at which these values are quoted",
                 " " * 15 + "r(rstar)" + " " * 9 + "r(au)" + " " * 11 + "tau" + " " * 8 + "Tdust"
                 + " " * 5 + "Tdust2"]
        for ir in range(model.nrg - 1):
            # format 900: (i6,3(1x,f15.5),2(1x,f10.5))
            lines.append(fmt_i(ir + 1, 6) + " " + fmt_f(rave[ir], 15, 5) + " "
                         + fmt_f(rave[ir] / model.autors, 15, 5) + " " + fmt_f(tau[ir], 15, 5)
                         + " " + fmt_f(tdust[ir, it, 0, id_] * EV_TO_K, 10, 5)
                         + " " + fmt_f(tdust2[ir, it, 0, id_] * EV_TO_K, 10, 5))
        with open(path, "w") as fh:
            fh.write("\n".join(lines) + "\n")
        paths.append(path)
    return paths


def write_tave(dirpath: str, model: Model, tdust: np.ndarray, tdust2: np.ndarray) -> list:
    """tave<id>.dat (output.F90:480-501): the mean temperature of each radial
    shell over its (ntg-1)(npg-1) cells, summed in the reference's loop
    order (theta outer, phi inner), in K."""
    rave, _, _ = cell_centres(model)
    nt, npc = model.ntg - 1, model.npg - 1
    paths = []
    for id_ in range(tdust.shape[3]):
        path = os.path.join(dirpath, f"tave{id_ + 1}.dat")
        lines = [list_directed("r(rstar),r(au),  Tave, Tave2")]
        for ir in range(model.nrg - 1):
            means = []
            for t in (tdust, tdust2):
                # row-major over (it, ip) = the reference's nesting; accumulate
                # is a left fold, so the rounding matches the Fortran loop
                shell = np.ascontiguousarray(t[ir, :nt, :npc, id_]).ravel()
                means.append(np.add.accumulate(shell)[-1] / float(nt * npc))
            # format 901: (i10,2(1x,f15.5),2(1x,f10.5))
            lines.append(fmt_i(ir + 1, 10) + " " + fmt_f(rave[ir], 15, 5) + " "
                         + fmt_f(rave[ir] / model.autors, 15, 5) + " "
                         + fmt_f(means[0] * EV_TO_K, 10, 5) + " "
                         + fmt_f(means[1] * EV_TO_K, 10, 5))
        with open(path, "w") as fh:
            fh.write("\n".join(lines) + "\n")
        paths.append(path)
    return paths


# --------------------------------------------------------------------------
# Fortran unformatted files
# --------------------------------------------------------------------------

def _record(fh, arr: np.ndarray) -> None:
    """One sequential record of arr's elements in column-major order."""
    b = np.asarray(arr).tobytes(order="F")
    n = np.array([len(b)], "<i4").tobytes()
    fh.write(n + b + n)


def write_grid_unf(path: str, model: Model, array: np.ndarray) -> None:
    """<prefix>.unf (output.F90:38-158): a record (nrg-1, ntg-1, npg-1,
    lastdim) of int32, then ravearr/autors, thetavearr, phiavearr (float64),
    then the array over the cells (column-major, ir fastest): float64 for
    real*8, int32 for integer, int32 0/1 for logical.  array is
    (nrg,ntg,npg) or (nrg,ntg,npg,ndg)."""
    a = np.asarray(array)
    if a.ndim not in (3, 4) or a.shape[:3] != (model.nrg, model.ntg, model.npg):
        raise ValueError(f"array shape {a.shape} is not (nrg,ntg,npg[,ndg])")
    last = 1 if a.ndim == 3 else a.shape[3]
    cells = a[: model.nrg - 1, : model.ntg - 1, : model.npg - 1]
    if a.dtype == np.bool_:
        data = cells.astype("<i4")
    elif np.issubdtype(a.dtype, np.integer):
        data = cells.astype("<i4")
    else:
        data = cells.astype("<f8")
    rave, tave, pave = cell_centres(model)
    with open(path, "wb") as fh:
        _record(fh, np.array([model.nrg - 1, model.ntg - 1, model.npg - 1, last], "<i4"))
        _record(fh, (rave / model.autors).astype("<f8"))
        _record(fh, tave.astype("<f8"))
        _record(fh, pave.astype("<f8"))
        _record(fh, data)


def read_unf_records(path: str) -> list:
    """The records of a sequential unformatted file, as raw bytes."""
    out = []
    with open(path, "rb") as fh:
        buf = fh.read()
    k = 0
    while k < len(buf):
        n = int(np.frombuffer(buf, "<i4", 1, k)[0])
        out.append(buf[k + 4:k + 4 + n])
        if int(np.frombuffer(buf, "<i4", 1, k + 4 + n)[0]) != n:
            raise ValueError(f"{path}: record markers disagree at byte {k}")
        k += n + 8
    return out


def write_image_cube(path: str, fl: Fluxes) -> None:
    """The normalised IMCUBE cube as one sequential unformatted file: a
    record (npeel, nx, nx, nfreq, no) of int32, then image_m_i
    (npeel,nx,nx,nfreq,no) and image_m_q, _u, _v (npeel,nx,nx,nfreq) as
    real*4, column-major.  No counterpart in the ASCII build."""
    if fl.image_m_i is None:
        raise ValueError("no image cube (run the pass with ipeel=1 and imcube)")
    mi = fl.image_m_i
    with open(path, "wb") as fh:
        _record(fh, np.array(mi.shape, "<i4"))
        for a in (mi, fl.image_m_q, fl.image_m_u, fl.image_m_v):
            _record(fh, a.astype("<f4"))


def write_filter_images(dirpath: str, fl: Fluxes, names: Sequence[str],
                        thete: Sequence[float], phie: Sequence[float],
                        partial: bool = False) -> list:
    """e_<filter>_<theta>_<phi>_{I,Q,U,V}_img<suffix>.dat for each observer
    and filter (output.F90:380-411): one row per iy, the nx values of ix in
    '(nx(ES11.4,1X))'; thete/phie are the observers' angles in radians."""
    if fl.image_b_i is None:
        raise ValueError("no filter images (run a peeled pass with filters set)")
    npeel, nx, _, nfilt = fl.image_b_i.shape
    if len(names) != nfilt:
        raise ValueError("one name per filter is needed")
    if len(thete) < npeel or len(phie) < npeel:
        raise ValueError("one (theta, phi) per observer is needed")
    psuf = "_partial" if partial else ""
    paths = []
    for pk in range(npeel):
        for ift in range(nfilt):
            beg = f"e_{names[ift].strip()}_{angle_label(thete[pk])}_{angle_label(phie[pk])}"
            for tag, a in (("I", fl.image_b_i), ("Q", fl.image_b_q), ("U", fl.image_b_u),
                           ("V", fl.image_b_v)):
                path = os.path.join(dirpath, f"{beg}_{tag}_img{psuf}.dat")
                with open(path, "w") as fh:
                    for iy in range(nx):
                        fh.write("".join(fmt_es(a[pk, ix, iy, ift], 11, 4) + " "
                                         for ix in range(nx)) + "\n")
                paths.append(path)
    return paths


def write_run_outputs(dirpath: str, model: Model, res, np_: float, tdust: np.ndarray,
                      tdust2: np.ndarray, dsg: Optional[np.ndarray] = None,
                      partial: bool = False,
                      filter_names: Optional[Sequence[str]] = None) -> list:
    """What output() writes (output.F90:167-503) for one pass `res` (a
    PassResult): flux*.dat, peel_flux*.dat when the pass peeled, nabs.unf,
    dsgratio.unf (dsg/(dbig+dsg); dsg counts the small-grain events, which
    this transport path does not have, so it defaults to zeros), tmidplane*.dat,
    tave*.dat; imcube.unf when the pass filled the cube; and the filter
    images when the pass tallied them and filter_names names the filters.  tdust2 is the
    temperature after the update, tdust the one the pass ran with
    (initarr.f90:82 copies tdust2 into tdust for the next pass)."""
    os.makedirs(dirpath, exist_ok=True)
    cube = None
    if getattr(res, "img_i", None) is not None:
        cube = (res.img_i, res.img_q, res.img_u, res.img_v)
    filt = getattr(res, "filt", None) if filter_names is not None else None
    fl = diskflux(model, np_, res.sed, res.peel, res.si_init, cube, filt=filt)
    paths = write_flux(dirpath, fl, model.numin)
    if fl.ti is not None:
        thete = np.arctan2(model.sinte, model.coste)
        phie = np.arctan2(model.sinpe, model.cospe)
        paths += write_peel_flux(dirpath, fl, thete, phie, model.numin, partial=partial)
        if fl.image_b_i is not None:
            paths += write_filter_images(dirpath, fl, filter_names, thete, phie,
                                         partial=partial)
    p = os.path.join(dirpath, "nabs.unf")
    write_grid_unf(p, model, res.nabs)
    paths.append(p)
    nsg = np.zeros(res.dbig.shape) if dsg is None else np.asarray(dsg, np.float64)
    tot = res.dbig.astype(np.float64) + nsg
    with np.errstate(all="ignore"):
        ratio = np.where(tot == 0, 0.0, nsg / tot)                       # output.F90:441-447
    p = os.path.join(dirpath, "dsgratio.unf")
    write_grid_unf(p, model, ratio)
    paths.append(p)
    paths += write_tmidplane(dirpath, model, tdust, tdust2)
    paths += write_tave(dirpath, model, tdust, tdust2)
    if fl.image_m_i is not None:
        p = os.path.join(dirpath, "imcube.unf")
        write_image_cube(p, fl)
        paths.append(p)
    return paths
