"""ctypes binding of the C ABI in include/hc.h (libhochunk3d_amd.so).

This is the same binding a Fortran driver makes with ISO_C_BINDING: plain
pointers and sizes, arrays in the reference's column-major layout.  Loading
fails loudly when the library is missing: there is no fallback path.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass
from typing import Optional

import numpy as np

LIBNAME = "libhochunk3d_amd.so"
_HERE = os.path.dirname(os.path.abspath(__file__))

HC_FP64, HC_FP32 = 0, 1
NO = 8

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)

# Every symbol include/hc.h declares.
EXPORTS = (
    "hc_create", "hc_set_grid", "hc_set_temperature", "hc_get_temperature", "hc_set_dust",
    "hc_set_sources", "hc_set_atmosphere", "hc_set_limb_table", "hc_set_bw",
    "hc_set_filters", "hc_get_filter_images", "hc_set_diffusion",
    "hc_run_iteration", "hc_temperature", "hc_run_lucy", "hc_run_batch", "hc_set_option",
    "hc_dropped_packets", "hc_destroy",
    "hc_last_error", "hc_build_density", "hc_build_density_noalloc", "hc_build_emissivity",
    "hc_fractal_counts", "hc_abi_version",
)


class HcDims(C.Structure):
    _fields_ = [(n, C.c_int32) for n in (
        "nrg", "ntg", "npg", "ndg", "nfreq", "nmu", "nph", "nap", "npeel", "nxhst",
        "nlammax", "nT", "nnu", "precision", "imcube")]


class HcSources(C.Structure):
    _fields_ = [
        ("rstar", C.c_double), ("tstar", C.c_double), ("tshock", C.c_double),
        ("limb", C.c_int32), ("ips", C.c_int32), ("ispot", C.c_int32), ("nspot", C.c_int32),
    ] + [(n, C.c_double) for n in (
        "fspot", "csprad", "cspot1", "sspot1", "scspot1", "ssspot1", "cspot2", "sspot2",
        "scspot2", "ssspot2", "accfrac", "accfrac2", "numin", "lnurat", "rmaxi", "ltot")] + [
        ("iintmax", C.c_int32), ("pad", C.c_int32),
        ("u", _dp), ("aperture2", _dp), ("coste", _dp), ("sinte", _dp), ("cospe", _dp),
        ("sinpe", _dp),
        ("rddustmin", C.c_double), ("rmaxdmax", C.c_double), ("rmax", C.c_double),
        ("massc", C.c_double), ("mdot", C.c_double),
        ("fmass", C.c_double * 8), ("z1", C.c_double * 8), ("bflare", C.c_double * 8),
    ]


class HcTallies(C.Structure):
    _fields_ = [
        ("dtauabs", _dp), ("nabs", _dp), ("jmean", _dp),
        ("killed", _ip), ("dbig", _ip), ("dsg", _ip),
        ("sed", _dp), ("peel", _dp),
        ("img_i", _dp), ("img_q", _dp), ("img_u", _dp), ("img_v", _dp), ("si_init", _dp),
    ] + [(n, C.c_double) for n in (
        "tot", "flux", "sflux", "nscat", "scount", "abflux", "aflux", "sumsub", "sumall",
        "hit", "htot")] + [
        ("n_err_nowall", C.c_int64), ("n_err_dust", C.c_int64), ("n_warn", C.c_int64),
        ("kernel_ms", C.c_double),
    ] + [(n, C.c_double) for n in ("ncross", "nscat_ev", "nabs_ev", "npeelray", "npeelcross",
                                   "max_packet_work", "ntally")]


class HcDensityModel(C.Structure):
    _fields_ = [
        ("nrg", C.c_int32), ("ntg", C.c_int32), ("npg", C.c_int32), ("ndisk", C.c_int32),
        ("rarr", _dp), ("thetarr", _dp), ("phiarr", _dp), ("fractal", _dp),
        ("rmincgs", C.c_double), ("disk_id", C.c_int32 * 8),
        ("disk_fmass", C.c_double * 8), ("disk_z1", C.c_double * 8), ("disk_b", C.c_double * 8),
    ] + [(n, C.c_double) for n in ("a_exp", "rmind", "rmaxd", "massd", "rc", "rhoe0", "rmine",
                                   "rmax", "rhoamb")] + [
        ("env_id", C.c_int32), ("cav_id", C.c_int32),
    ] + [(n, C.c_double) for n in ("c1e", "cavity_exp", "rho_cavity", "r1000")]


TRACE_DTYPE = np.dtype([
    ("x", "f8"), ("y", "f8"), ("z", "f8"), ("cost", "f8"), ("sint", "f8"), ("phi", "f8"),
    ("I", "f8"), ("Q", "f8"), ("U", "f8"), ("V", "f8"), ("nub", "f8"), ("path", "f8"),
    ("tau_peel", "f8"), ("ncross", "i8"), ("npeelcross", "i8"), ("nscat", "i8"), ("nabs", "i8"),
    ("nstar", "i8"), ("ntrial", "i8"), ("cellhash", "u8"), ("status", "i4"), ("ir", "i4"),
    ("it", "i4"), ("ip", "i4"), ("pad", "i4"),
], align=True)

SCALARS = ("tot", "flux", "sflux", "nscat", "scount", "abflux", "aflux", "sumsub", "sumall",
           "hit", "htot", "n_err_nowall", "n_err_dust", "n_warn", "ncross", "nscat_ev",
           "nabs_ev", "npeelray", "npeelcross", "max_packet_work", "ntally")


def library_path() -> str:
    return os.environ.get("HOCHUNK3D_AMD_LIB", os.path.join(_HERE, LIBNAME))


_lib = None


def load() -> C.CDLL:
    """Load the library (raises OSError if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    path = library_path()
    if not os.path.exists(path):
        raise OSError(f"{path} is not built; run __graft_entry__.build()")
    lib = C.CDLL(path)
    lib.hc_create.argtypes = [C.POINTER(HcDims), C.c_int, C.POINTER(C.c_void_p)]
    lib.hc_set_grid.argtypes = [C.c_void_p] + [_dp] * 8 + [C.c_double]
    lib.hc_set_temperature.argtypes = [C.c_void_p, _dp]
    lib.hc_get_temperature.argtypes = [C.c_void_p, _dp]
    lib.hc_set_dust.argtypes = [C.c_void_p, _ip] + [_dp] * 10
    lib.hc_set_sources.argtypes = [C.c_void_p, C.POINTER(HcSources)]
    lib.hc_set_atmosphere.argtypes = [C.c_void_p, C.c_int32, _dp, _dp, _dp, _dp]
    lib.hc_set_limb_table.argtypes = [C.c_void_p, C.c_int32, _dp, _dp]
    lib.hc_set_bw.argtypes = [C.c_void_p, _dp, _dp]
    lib.hc_set_filters.argtypes = [C.c_void_p, C.c_int32, C.c_int32, _ip, _dp, _dp]
    lib.hc_get_filter_images.argtypes = [C.c_void_p, _dp, _dp, _dp, _dp]
    lib.hc_set_diffusion.argtypes = [C.c_void_p, C.POINTER(C.c_int8), C.POINTER(C.c_int8), _dp,
                                     _dp]
    lib.hc_run_iteration.argtypes = [C.c_void_p, C.c_int64, C.c_int32, C.c_int32, C.c_int32,
                                     C.c_int32, C.c_uint64, C.POINTER(HcTallies), C.c_void_p]
    lib.hc_temperature.argtypes = [C.c_void_p, _dp, C.c_double, C.c_double, _dp,
                                   C.POINTER(C.c_int64)]
    lib.hc_run_lucy.argtypes = [C.c_void_p, C.c_int32, C.c_int64, C.c_int32, C.c_int32,
                                C.c_int32, C.c_uint64, _dp, C.c_double, C.POINTER(HcTallies),
                                C.POINTER(C.c_int64)]
    lib.hc_run_batch.argtypes = [C.c_int32, C.POINTER(C.c_void_p), C.c_int32,
                                 C.POINTER(C.c_int64), C.c_int32, C.c_int32, C.c_int32,
                                 C.POINTER(C.c_uint64), C.POINTER(_dp), _dp, C.c_int32,
                                 C.POINTER(HcTallies), C.POINTER(C.c_int64),
                                 C.POINTER(C.c_double)]
    lib.hc_set_option.argtypes = [C.c_void_p, C.c_char_p, C.c_int64]
    lib.hc_dropped_packets.argtypes = [C.c_void_p, C.POINTER(C.c_int64), C.c_int32,
                                       C.POINTER(C.c_int64)]
    lib.hc_build_density.argtypes = [C.POINTER(HcDensityModel), C.c_int, _dp, _dp, _dp]
    lib.hc_build_density_noalloc.argtypes = [C.POINTER(HcDensityModel), C.c_int, C.c_void_p,
                                             C.c_void_p, C.c_void_p, C.c_void_p]
    lib.hc_build_emissivity.argtypes = [C.c_int, C.c_int32, _dp, _dp, _dp, C.c_int32, _dp,
                                        C.c_int32, _dp, C.c_int32, _dp, _dp, _dp, _dp]
    lib.hc_fractal_counts.argtypes = [C.c_int, C.c_int32, C.c_int32, C.c_int32, _dp, _dp, _dp,
                                      _dp, C.c_int32, C.c_int32, _dp, C.c_double,
                                      C.POINTER(C.c_uint64), C.POINTER(C.c_uint32)]
    lib.hc_destroy.argtypes = [C.c_void_p]
    lib.hc_last_error.argtypes = [C.c_void_p]
    lib.hc_last_error.restype = C.c_char_p
    lib.hc_abi_version.argtypes = []
    for name in EXPORTS:
        if name not in ("hc_last_error",):
            getattr(lib, name).restype = C.c_int
    _lib = lib
    return lib


def _ptr(a: Optional[np.ndarray], ct=C.c_double):
    if a is None:
        return None
    return a.ctypes.data_as(C.POINTER(ct))


def _f64(a) -> np.ndarray:
    return np.asfortranarray(np.asarray(a, dtype=np.float64))


@dataclass
class PassResult:
    """Tallies of one transport pass (reference layouts, FP64)."""
    dtauabs: np.ndarray
    nabs: np.ndarray
    jmean: np.ndarray
    killed: np.ndarray
    dbig: np.ndarray
    sed: np.ndarray        # (9, nfreq, nmu, nph, nap, no) I,Q,U,V,I2,Q2,U2,V2,N
    peel: Optional[np.ndarray]
    img_i: Optional[np.ndarray]
    img_q: Optional[np.ndarray]
    img_u: Optional[np.ndarray]
    img_v: Optional[np.ndarray]
    si_init: np.ndarray
    scalars: dict
    trace: Optional[np.ndarray]
    kernel_ms: float
    # (image_b_i, q, u, v), each (npeel, nx, nx, nfilt): the filter images of
    # a peeled pass when filters are set
    filt: Optional[tuple] = None


class HcError(RuntimeError):
    pass


class Transport:
    """One GPU context holding a model's grid, dust and sources."""

    def __init__(self, model, device: int = 0, precision: int = HC_FP64, imcube=None,
                 cube_fp32: bool = False):
        """cube_fp32: accumulate the image cube in FP32 (hc_dims.imcube = 2),
        returned as float32 arrays, as the reference's real*4 images."""
        self.precision = precision
        self.lib = load()
        m = model
        self.model = m
        cfg = m.cfg
        self.imcube = bool(cfg.imcube if imcube is None else imcube)
        self.cube_fp32 = bool(cube_fp32 and self.imcube)
        self.dims = HcDims(m.nrg, m.ntg, m.npg, 8, cfg.nfreq, cfg.nmu, cfg.nph, cfg.nap,
                           cfg.npeel, cfg.nximg, m.dust.nudust.shape[0], m.dust.lTint.size,
                           m.dust.nuint.size, precision,
                           (2 if self.cube_fp32 else 1) if self.imcube else 0)
        h = C.c_void_p()
        rc = self.lib.hc_create(C.byref(self.dims), device, C.byref(h))
        if rc != 0:
            raise HcError(f"hc_create failed ({rc})")
        self.h = h
        self._keep = []
        self._filters = None
        self._check(self.lib.hc_set_grid(
            h, _ptr(_f64(m.r2arr)), _ptr(_f64(m.tan2arr)), _ptr(_f64(m.thetarr)),
            _ptr(_f64(m.aarr)), _ptr(_f64(m.barr)), _ptr(_f64(m.phiarr)),
            _ptr(_f64(m.densarr)), _ptr(_f64(m.tdust)), float(cfg.tsub)), "hc_set_grid")
        d = m.dust
        nl = np.ascontiguousarray(d.nlambda, dtype=np.int32)
        self._check(self.lib.hc_set_dust(
            h, _ptr(nl, C.c_int32), _ptr(_f64(d.nudust)), _ptr(_f64(d.kappad)),
            _ptr(_f64(d.adust)), _ptr(_f64(d.gdust)), _ptr(_f64(d.pldust)),
            _ptr(_f64(d.kappav)), _ptr(_f64(d.lTint)), _ptr(_f64(d.nuint)),
            _ptr(_f64(d.kapBint)), _ptr(_f64(d.kappap))), "hc_set_dust")
        self.set_sources()
        if getattr(m, "diffus", None) is not None:  # DIFFUSION=YES
            self.set_diffusion()

    def set_sources(self, **over):
        m = self.model
        arrs = [_f64(a) for a in (m.u, m.aperture2, m.coste, m.sinte, m.cospe, m.sinpe)]
        self._keep = arrs
        s = HcSources()
        s.rstar = m.rstar
        s.tstar = m.tstar_ev
        s.tshock = m.tshock_ev
        s.accfrac2 = m.cfg.accfrac2
        s.numin = m.numin
        s.lnurat = m.lnurat
        s.rmaxi = m.rmaxi
        s.ltot = m.ltot
        s.iintmax = 10000
        s.nspot = 1
        s.limb = int(m.cfg.limb)
        s.accfrac = m.accfrac
        s.rddustmin, s.rmaxdmax, s.rmax = m.rddustmin, m.rmaxdmax, m.rarr[-1]
        s.massc, s.mdot = m.cfg.massc, m.mdot
        for i in range(8):
            s.fmass[i], s.z1[i], s.bflare[i] = m.fmass8[i], m.z1_8[i], m.bflare8[i]
        for k, v in over.items():
            setattr(s, k, v)
        s.u, s.aperture2, s.coste, s.sinte, s.cospe, s.sinpe = [_ptr(a) for a in arrs]
        if m.limb_tab is not None:
            xmu, prob = (_f64(a) for a in m.limb_tab)
            self._check(self.lib.hc_set_limb_table(self.h, xmu.size, _ptr(xmu), _ptr(prob)),
                        "hc_set_limb_table")
        self._check(self.lib.hc_set_sources(self.h, C.byref(s)), "hc_set_sources")
        self.set_atmosphere(m.atmosphere)

    def set_atmosphere(self, atm):
        """The stellar spectrum table (hochunk3d_amd.atmos.Atmosphere), or
        None for the Planck function at tstar."""
        if atm is None:
            self._check(self.lib.hc_set_atmosphere(self.h, 0, None, None, None, None),
                        "hc_set_atmosphere")
            return
        a = [_f64(v) for v in (atm.atmosnu, atm.hnuint, atm.lognu, atm.loghnu)]
        self._check(self.lib.hc_set_atmosphere(self.h, a[0].size, *(_ptr(v) for v in a)),
                    "hc_set_atmosphere")

    def set_bw(self, on: bool = True):
        """Bjorkman-Wood re-emission (hc_set_bw) with the model's kapdBint
        (built on first use) and massarr; on=False returns to Lucy's."""
        if not on:
            self._check(self.lib.hc_set_bw(self.h, None, None), "hc_set_bw")
            return
        from .dust import ensure_kapdB
        kdb = _f64(ensure_kapdB(self.model.dust))
        mass = _f64(self.model.massarr)
        self._check(self.lib.hc_set_bw(self.h, _ptr(kdb), _ptr(mass)), "hc_set_bw")

    def set_filters(self, filters):
        """Filter curves (hochunk3d_amd.filters.Filters) for the filter
        images of peeled passes (hc_set_filters); None switches them off."""
        if filters is None:
            self._filters = None
            self._check(self.lib.hc_set_filters(self.h, 0, 0, None, None, None),
                        "hc_set_filters")
            return
        nw = np.ascontiguousarray(filters.nwfilt, np.int32)
        fw, fp = _f64(filters.filtwave), _f64(filters.filtphi)
        self._check(self.lib.hc_set_filters(self.h, filters.nfilt, fw.shape[0],
                                            _ptr(nw, C.c_int32), _ptr(fw), _ptr(fp)),
                    "hc_set_filters")
        self._filters = filters

    def set_diffusion(self, on: bool = True):
        """The diffusion approximation with the model's layer (model.diffus,
        model.diffdir from hochunk3d_amd.diffusion) and Rosseland table
        (hc_set_diffusion); on=False switches it off."""
        if not on:
            self._check(self.lib.hc_set_diffusion(self.h, None, None, None, None),
                        "hc_set_diffusion")
            return
        from .dust import ensure_kappar
        m = self.model
        if m.diffus is None:
            raise HcError("set_diffusion: the model has no diffusion layer")
        dif = np.asfortranarray(m.diffus, np.int8)
        ddir = np.asfortranarray(m.diffdir, np.int8)
        self._check(self.lib.hc_set_diffusion(
            self.h, _ptr(dif, C.c_int8), _ptr(ddir, C.c_int8), _ptr(_f64(m.rarr)),
            _ptr(_f64(ensure_kappar(m.dust)))), "hc_set_diffusion")

    def filter_images(self):
        """(image_b_i, q, u, v) of the last peeled pass, (npeel, nx, nx, nfilt)."""
        if self._filters is None:
            raise HcError("filter_images: no filters set")
        cfg = self.model.cfg
        shp = (max(cfg.npeel, 1), cfg.nximg, cfg.nximg, self._filters.nfilt)
        out = [np.zeros(shp, order="F") for _ in range(4)]
        self._check(self.lib.hc_get_filter_images(self.h, *(_ptr(a) for a in out)),
                    "hc_get_filter_images")
        return tuple(out)

    def set_option(self, name: str, value: int):
        """Launch-shape knob: regen_lanes, regen_run_lanes, peel_queue,
        blocks_per_cu or max_work (include/hc.h documents each)."""
        self._check(self.lib.hc_set_option(self.h, name.encode(), int(value)), "hc_set_option")

    def dropped_packets(self):
        """(count, packet numbers) of the last pass's packets dropped at
        max_work (at most 16 numbers)."""
        pids = (C.c_int64 * 16)()
        n = C.c_int64(0)
        self._check(self.lib.hc_dropped_packets(self.h, pids, 16, C.byref(n)),
                    "hc_dropped_packets")
        return int(n.value), [int(pids[i]) for i in range(min(int(n.value), 16))]

    def _check(self, rc, what):
        if rc != 0:
            msg = self.lib.hc_last_error(self.h)
            raise HcError(f"{what} failed ({rc}): {msg.decode() if msg else ''}")

    def set_temperature(self, tdust: np.ndarray):
        self._check(self.lib.hc_set_temperature(self.h, _ptr(_f64(tdust))), "hc_set_temperature")

    def get_temperature(self) -> np.ndarray:
        m = self.model
        t = np.zeros((m.nrg, m.ntg, m.npg, 8), order="F")
        self._check(self.lib.hc_get_temperature(self.h, _ptr(t)), "hc_get_temperature")
        return t

    def run_scalars(self, np_: int, ipeel: int = 0, seed1: int = 12345678,
                    seed2: int = 87654321, gen: int = 1, seg0: int = 0):
        """A pass whose tallies stay on the device (the next Lucy step reads
        them there); returns the scalar tallies and the kernel time in ms."""
        t = HcTallies()
        rc = self.lib.hc_run_iteration(self.h, int(np_), int(ipeel), int(seed1), int(seed2),
                                       int(gen), int(seg0), C.byref(t), None)
        self._check(rc, "hc_run_iteration")
        return {k: getattr(t, k) for k in SCALARS}, t.kernel_ms

    def run_sed(self, np_: int, seed1: int = 12345678, seed2: int = 87654321, gen: int = 1,
                seg0: int = 0):
        """A Lucy pass that downloads only the SED tallies (per-cell tallies
        stay on the device); returns (scalars, kernel ms, sed[9, nfreq, nmu,
        nph, nap, 8])."""
        cfg = self.model.cfg
        sed = np.zeros((cfg.nfreq, cfg.nmu, cfg.nph, cfg.nap, NO, 9), order="F")
        t = HcTallies()
        t.sed = _ptr(sed)
        rc = self.lib.hc_run_iteration(self.h, int(np_), 0, int(seed1), int(seed2), int(gen),
                                       int(seg0), C.byref(t), None)
        self._check(rc, "hc_run_iteration")
        return {k: getattr(t, k) for k in SCALARS}, t.kernel_ms, np.moveaxis(sed, -1, 0)

    def run(self, np_: int, ipeel: int = 0, seed1: int = 12345678, seed2: int = 87654321,
            gen: int = 1, seg0: int = 0, trace: bool = False, images: bool = True) -> PassResult:
        m = self.model
        cfg = m.cfg
        shp4 = (m.nrg, m.ntg, m.npg, 8)
        shp3 = (m.nrg, m.ntg, m.npg)
        dtauabs = np.zeros(shp4, order="F")
        nabs = np.zeros(shp4, order="F")
        jmean = np.zeros(shp3, order="F")
        killed = np.zeros(shp3, np.int32, order="F")
        dbig = np.zeros(shp3, np.int32, order="F")
        sed = np.zeros((cfg.nfreq, cfg.nmu, cfg.nph, cfg.nap, NO, 9), order="F")
        peel = np.zeros((cfg.nfreq, max(cfg.npeel, 1), cfg.nap, NO, 9), order="F") if ipeel else None
        cube = bool(ipeel and self.imcube and images)
        nimg = (max(cfg.npeel, 1), cfg.nximg, cfg.nximg, cfg.nfreq)
        ct = np.float32 if self.cube_fp32 else np.float64
        img_i = np.zeros(nimg + (NO,), ct, order="F") if cube else None
        img_q = np.zeros(nimg, ct, order="F") if cube else None
        img_u = np.zeros(nimg, ct, order="F") if cube else None
        img_v = np.zeros(nimg, ct, order="F") if cube else None
        si_init = np.zeros((cfg.nfreq, 5), order="F")
        t = HcTallies()
        t.dtauabs, t.nabs, t.jmean = _ptr(dtauabs), _ptr(nabs), _ptr(jmean)
        t.killed, t.dbig = _ptr(killed, C.c_int32), _ptr(dbig, C.c_int32)
        t.sed = _ptr(sed)
        t.peel = _ptr(peel)
        t.img_i, t.img_q, t.img_u, t.img_v = (
            None if a is None else C.cast(a.ctypes.data, _dp) for a in (img_i, img_q, img_u, img_v))
        t.si_init = _ptr(si_init)
        tr = np.zeros(np_, TRACE_DTYPE) if trace else None
        rc = self.lib.hc_run_iteration(self.h, int(np_), int(ipeel), int(seed1), int(seed2),
                                       int(gen), int(seg0), C.byref(t),
                                       tr.ctypes.data if tr is not None else None)
        self._check(rc, "hc_run_iteration")
        scal = {k: getattr(t, k) for k in SCALARS}
        sed = np.moveaxis(sed, -1, 0)
        if peel is not None:
            peel = np.moveaxis(peel, -1, 0)
        filt = self.filter_images() if (ipeel and self._filters is not None) else None
        return PassResult(dtauabs, nabs, jmean, killed, dbig, sed, peel, img_i, img_q, img_u,
                          img_v, si_init, scal, tr, t.kernel_ms, filt)

    def temperature(self, np_: int, apply: bool = False):
        """Lucy temperature update from the last pass.  apply=False returns the
        new temperatures (reference layout); apply=True installs them on the
        device and returns None."""
        m = self.model
        t2 = None if apply else np.zeros((m.nrg, m.ntg, m.npg, 8), order="F")
        n = C.c_int64(0)
        mass = None
        if not getattr(self, "_mass_sent", False):
            mass = _f64(m.massarr)
            self._mass_sent = True
        self._check(self.lib.hc_temperature(self.h, _ptr(mass), float(m.ltot), float(np_),
                                            _ptr(t2), C.byref(n)), "hc_temperature")
        return t2, int(n.value)

    def run_lucy(self, niter: int, np_: int, seed1: int = 12345678, seed2: int = 87654321,
                 gen: int = 1, seg0: int = 0):
        """niter Lucy iterations (pass + device temperature update) replayed
        from one captured HIP graph; iteration i uses segments
        seg0 + i*np_ ...  Returns (scalars of the last pass, device ms of all
        iterations, changed cell-species of the last update)."""
        m = self.model
        mass = None
        if not getattr(self, "_mass_sent", False):
            mass = _f64(m.massarr)
            self._mass_sent = True
        t = HcTallies()
        n = C.c_int64(0)
        self._check(self.lib.hc_run_lucy(self.h, int(niter), int(np_), int(seed1), int(seed2),
                                         int(gen), int(seg0), _ptr(mass), float(m.ltot),
                                         C.byref(t), C.byref(n)), "hc_run_lucy")
        return {k: getattr(t, k) for k in SCALARS}, t.kernel_ms, int(n.value)

    def close(self):
        if getattr(self, "h", None):
            self.lib.hc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def run_batch_lucy(transports, niter: int, packets, seg0=None, seed1: int = 12345678,
                   seed2: int = 87654321, gen: int = 1, oversubscribe: int = 0,
                   sed: bool = False):
    """hc_run_batch: niter Lucy iterations of every Transport's model with all
    passes in one launch per kernel variant.  Returns (per-model list of
    (scalars, changed cell-species, SED or None), device ms of all
    iterations); sed downloads each model's last-pass SED
    [9, nfreq, nmu, nph, nap, 8]."""
    n = len(transports)
    lib = load()
    hs = (C.c_void_p * n)(*[t.h.value for t in transports])
    npk = (C.c_int64 * n)(*[int(p) for p in packets])
    sg = (C.c_uint64 * n)(*[int(x) for x in (seg0 if seg0 is not None else [0] * n)])
    mass_arrays = []
    mass = (_dp * n)()
    for i, t in enumerate(transports):
        if not getattr(t, "_mass_sent", False):
            a = _f64(t.model.massarr)
            mass_arrays.append(a)
            mass[i] = _ptr(a)
            t._mass_sent = True
    ltot = np.array([float(t.model.ltot) for t in transports])
    outs = (HcTallies * n)()
    seds = []
    for i, t in enumerate(transports):
        if sed:
            cfg = t.model.cfg
            a = np.zeros((cfg.nfreq, cfg.nmu, cfg.nph, cfg.nap, NO, 9), order="F")
            outs[i].sed = _ptr(a)
            seds.append(a)
        else:
            seds.append(None)
    nch = (C.c_int64 * n)()
    ms = C.c_double(0.0)
    rc = lib.hc_run_batch(n, hs, int(niter), npk, int(seed1), int(seed2), int(gen), sg, mass,
                          _ptr(ltot), int(oversubscribe), outs, nch, C.byref(ms))
    if rc != 0:
        msg = lib.hc_last_error(transports[0].h)
        raise HcError(f"hc_run_batch failed ({rc}): {msg.decode() if msg else ''}")
    res = []
    for i in range(n):
        scal = {k: getattr(outs[i], k) for k in SCALARS}
        res.append((scal, int(nch[i]), None if seds[i] is None else np.moveaxis(seds[i], -1, 0)))
    return res, ms.value


def build_density(fields: dict, nrg: int, ntg: int, npg: int, device: int = 0):
    """hc_build_density: (densarr, massarr, vcell) in the reference layout.
    `fields` holds the hc_density_model members (arrays as NumPy arrays)."""
    lib = load()
    keep = {}
    m = HcDensityModel()
    for k, v in fields.items():
        if k in ("rarr", "thetarr", "phiarr", "fractal"):
            if v is None:
                continue
            keep[k] = np.asfortranarray(v, dtype=np.float64)
            setattr(m, k, keep[k].ctypes.data_as(_dp))
        elif k in ("disk_id", "disk_fmass", "disk_z1", "disk_b"):
            arr = getattr(m, k)
            for i, x in enumerate(v):
                arr[i] = x
        else:
            setattr(m, k, v)
    dens = np.zeros((nrg, ntg, npg, 8), order="F")
    mass = np.zeros((nrg, ntg, npg, 8), order="F")
    vol = np.zeros((nrg, ntg, npg), order="F")
    rc = lib.hc_build_density(C.byref(m), int(device), _ptr(dens), _ptr(mass), _ptr(vol))
    if rc != 0:
        raise HcError(f"hc_build_density failed ({rc})")
    return dens, mass, vol


def density_model_struct(fields: dict, arrays: Optional[dict] = None) -> HcDensityModel:
    """An hc_density_model from `fields`; the array members (rarr, thetarr,
    phiarr, fractal) are taken from `arrays` as raw addresses (ints) when
    given, e.g. device addresses for hc_build_density_noalloc."""
    m = HcDensityModel()
    for k, v in fields.items():
        if k in ("rarr", "thetarr", "phiarr", "fractal"):
            continue
        if k in ("disk_id", "disk_fmass", "disk_z1", "disk_b"):
            arr = getattr(m, k)
            for i, x in enumerate(v):
                arr[i] = x
        else:
            setattr(m, k, v)
    for k, addr in (arrays or {}).items():
        if addr:
            setattr(m, k, C.cast(C.c_void_p(int(addr)), _dp))
    return m


def build_density_noalloc(fields: dict, arrays: dict, densarr: int, massarr: int, vcell: int,
                          device: int = 0, stream: int = 0) -> None:
    """hc_build_density_noalloc: the densities of `fields` written into the
    device arrays at addresses densarr, massarr, vcell; `arrays` maps rarr,
    thetarr, phiarr (and optionally fractal) to device addresses.  Queued on
    `stream` (0: the null stream); the caller synchronises."""
    lib = load()
    m = density_model_struct(fields, arrays)
    rc = lib.hc_build_density_noalloc(C.byref(m), int(device), C.c_void_p(int(stream) or None),
                                      C.c_void_p(int(densarr)), C.c_void_p(int(massarr)),
                                      C.c_void_p(int(vcell)))
    if rc != 0:
        raise HcError(f"hc_build_density_noalloc failed ({rc})")


def build_emissivity(nud, kd, ad, lTint, nuint, xq, wq, device: int = 0):
    """hc_build_emissivity: (kapBint (nT, nnu) F-order, kappap (nT,))."""
    lib = load()
    a = [np.ascontiguousarray(v, dtype=np.float64) for v in (nud, kd, ad, lTint, nuint, xq, wq)]
    nT, nnu = len(a[3]), len(a[4])
    kapB = np.zeros((nT, nnu), order="F")
    kp = np.zeros(nT)
    rc = lib.hc_build_emissivity(int(device), len(a[0]), _ptr(a[0]), _ptr(a[1]), _ptr(a[2]),
                                 nT, _ptr(a[3]), nnu, _ptr(a[4]), len(a[5]), _ptr(a[5]),
                                 _ptr(a[6]), _ptr(kapB), _ptr(kp))
    if rc != 0:
        raise HcError(f"hc_build_emissivity failed ({rc})")
    return kapB, kp


def fractal_counts_device(rarr, mcos_t, cph, sph, n: int, levels: int, scale, scale_r: float,
                          state: int, inc: int, device: int = 0) -> np.ndarray:
    """hc_fractal_counts: points per cell (nrg,ntg,npg) F-order, as float64;
    state/inc are the PCG64 state and increment (128-bit integers)."""
    lib = load()
    a = [np.ascontiguousarray(v, dtype=np.float64) for v in (rarr, mcos_t, cph, sph)]
    sc = np.ascontiguousarray(scale if len(scale) else [0.0], dtype=np.float64)
    m64 = (1 << 64) - 1
    pcg = np.array([state & m64, state >> 64, inc & m64, inc >> 64], dtype=np.uint64)
    nrg, ntg, npg = len(a[0]), len(a[1]), len(a[2])
    cnt = np.zeros((nrg, ntg, npg), dtype=np.uint32, order="F")
    rc = lib.hc_fractal_counts(int(device), nrg, ntg, npg, _ptr(a[0]), _ptr(a[1]), _ptr(a[2]),
                               _ptr(a[3]), int(n), int(levels), _ptr(sc), float(scale_r),
                               pcg.ctypes.data_as(C.POINTER(C.c_uint64)),
                               cnt.ctypes.data_as(C.POINTER(C.c_uint32)))
    if rc != 0:
        raise HcError(f"hc_fractal_counts failed ({rc})")
    return cnt.astype(np.float64, order="F")
