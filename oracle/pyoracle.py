"""ctypes binding of the FP64 CPU oracle (oracle/libhc_oracle.so).

TEST INFRASTRUCTURE: imported only by tests/, the smoke check and the
benchmark's CPU-baseline leg, as the checker.  The product never imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from typing import Optional

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIBS = {"libm": os.path.join(_HERE, "libhc_oracle.so"),
        "dm": os.path.join(_HERE, "libhc_oracle_dm.so")}
NDG = 8
NO = 8

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)


class Grid(C.Structure):
    _fields_ = [("nrg", C.c_int), ("ntg", C.c_int), ("npg", C.c_int),
                ("r2arr", _dp), ("tan2arr", _dp), ("thetarr", _dp), ("aarr", _dp),
                ("barr", _dp), ("phiarr", _dp), ("densarr", _dp), ("tdust", _dp),
                ("massarr", _dp), ("tsub", C.c_double), ("rarr", _dp),
                ("diffus", C.POINTER(C.c_int8)), ("diffdir", C.POINTER(C.c_int8))]


class Dust(C.Structure):
    _fields_ = [("nlammax", C.c_int), ("nlambda", C.c_int * NDG),
                ("nudust", _dp), ("kappad", _dp), ("adust", _dp), ("gdust", _dp),
                ("pldust", _dp), ("kappav", C.c_double * NDG), ("nT", C.c_int),
                ("nnu", C.c_int), ("lTint", _dp), ("nuint", _dp), ("kapBint", _dp),
                ("kappap", _dp), ("kapdBint", _dp), ("kappar", _dp)]


class Setup(C.Structure):
    _fields_ = [("rstar", C.c_double), ("tstar", C.c_double), ("tshock", C.c_double),
                ("limb", C.c_int), ("ips", C.c_int), ("iplanckst", C.c_int),
                ("ispot", C.c_int), ("nspot", C.c_int)] + [
        (n, C.c_double) for n in ("fspot", "csprad", "cspot1", "sspot1", "scspot1", "ssspot1",
                                  "cspot2", "sspot2", "scspot2", "ssspot2", "accfrac",
                                  "accfrac2")] + [
        ("nfreq", C.c_int), ("nmu", C.c_int), ("nph", C.c_int), ("nap", C.c_int),
        ("npeel", C.c_int), ("nxhst", C.c_int), ("numin", C.c_double), ("lnurat", C.c_double),
        ("u", _dp), ("aperture2", _dp), ("coste", _dp), ("sinte", _dp), ("cospe", _dp),
        ("sinpe", _dp), ("rmaxi", C.c_double), ("imcube", C.c_int), ("iintmax", C.c_int),
        ("ltot", C.c_double),
        ("rddustmin", C.c_double), ("rmaxdmax", C.c_double), ("rmax", C.c_double),
        ("massc", C.c_double), ("mdot", C.c_double),
        ("fmass", C.c_double * NDG), ("z1", C.c_double * NDG), ("bflare", C.c_double * NDG),
        ("nhnu", C.c_int), ("atmosnu", _dp), ("hnuint", _dp), ("lognu", _dp), ("loghnu", _dp),
        ("ntab", C.c_int), ("xmu", _dp), ("prob", _dp), ("delmu", C.c_double),
        ("bw", C.c_int), ("nfilt", C.c_int), ("nwfiltmax", C.c_int), ("nwfilt", _ip),
        ("filtwave", _dp), ("filtphi", _dp)]


class Tallies(C.Structure):
    _fields_ = [("dtauabs", _dp), ("nabs", _dp), ("jmean", _dp), ("killed", _ip),
                ("dbig", _ip), ("dsg", _ip), ("sed", _dp), ("peel", _dp), ("img_i", _dp),
                ("img_q", _dp), ("img_u", _dp), ("img_v", _dp), ("si_init", _dp)] + [
        (n, C.c_double) for n in ("tot", "flux", "sflux", "nscat", "scount", "abflux", "aflux",
                                  "sumsub", "sumall", "hit", "htot")] + [
        ("n_err_nowall", C.c_int64), ("n_err_dust", C.c_int64), ("n_warn", C.c_int64),
        ("image_b_i", _dp), ("image_b_q", _dp), ("image_b_u", _dp), ("image_b_v", _dp)]


class Rng(C.Structure):
    _fields_ = [("s1", C.c_int32), ("s2", C.c_int32)]


class Walls(C.Structure):
    _fields_ = [("tn1", C.c_double), ("tp1", C.c_double), ("tp2", C.c_double),
                ("in1", C.c_int), ("ip1", C.c_int), ("ip2", C.c_int)]


class Opac(C.Structure):
    _fields_ = [(n, C.c_double * NDG) for n in ("kappa", "albedo", "g", "pl", "g2", "g2p1",
                                                 "onemg2", "twog")]


class Scat(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("cost", "sint", "cosp", "sinp", "phi", "I", "Q", "U",
                                          "V", "hit", "htot")]


TRACE_DTYPE = np.dtype([
    ("x", "f8"), ("y", "f8"), ("z", "f8"), ("cost", "f8"), ("sint", "f8"), ("phi", "f8"),
    ("I", "f8"), ("Q", "f8"), ("U", "f8"), ("V", "f8"), ("nub", "f8"), ("path", "f8"),
    ("tau_peel", "f8"), ("ncross", "i8"), ("npeelcross", "i8"), ("nscat", "i8"), ("nabs", "i8"),
    ("nstar", "i8"), ("ntrial", "i8"), ("cellhash", "u8"), ("status", "i4"), ("ir", "i4"),
    ("it", "i4"), ("ip", "i4"), ("pad", "i4"),
], align=True)

SCALARS = ("tot", "flux", "sflux", "nscat", "scount", "abflux", "aflux", "sumsub", "sumall",
           "hit", "htot", "n_err_nowall", "n_err_dust", "n_warn")

_libs = {}


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def load(variant: str = "libm") -> C.CDLL:
    """variant "libm": elementary functions of the C library (as the
    reference's own build); "dm": the deterministic set the GPU uses; or the
    path of another build of hc_oracle.c (a .so file)."""
    if variant not in _libs:
        path = variant if variant.endswith(".so") else LIBS[variant]
        if not os.path.exists(path):
            build()
        lib = C.CDLL(path)
        lib.hco_run_pass.argtypes = [C.POINTER(Grid), C.POINTER(Dust), C.POINTER(Setup),
                                     C.POINTER(Tallies), C.c_int64, C.c_int64, C.c_int64,
                                     C.c_int, C.c_int32, C.c_int32, C.c_int, C.c_uint64,
                                     C.c_void_p]
        lib.hco_run_pass.restype = C.c_int
        lib.hco_tfinal.argtypes = [C.POINTER(Grid), C.POINTER(Dust), _dp, C.c_double, C.c_double,
                                   _dp]
        lib.hco_tfinal.restype = C.c_int64
        lib.hco_rng_packet.argtypes = [C.POINTER(Rng), C.c_int32, C.c_int32, C.c_int, C.c_uint64]
        lib.hco_ran.argtypes = [C.POINTER(Rng)]
        lib.hco_ran.restype = C.c_double
        lib.hco_reset_t.argtypes = [C.POINTER(Walls)]
        lib.hco_insert_t.argtypes = [C.POINTER(Walls), C.c_double, C.c_int]
        lib.hco_radface.argtypes = [C.POINTER(Walls)] + [C.c_double] * 8
        lib.hco_thetaface.argtypes = [C.POINTER(Walls)] + [C.c_double] * 8
        lib.hco_phiface.argtypes = [C.POINTER(Walls)] + [C.c_double] * 8
        lib.hco_find_next_wall.argtypes = [C.POINTER(Walls), C.c_int, _dp, C.POINTER(C.c_int)]
        lib.hco_find_next_wall.restype = C.c_int
        lib.hco_locate.argtypes = [_dp, C.c_int, C.c_double]
        lib.hco_locate.restype = C.c_int
        lib.hco_walk_ray.argtypes = ([C.POINTER(Grid)] + [C.c_double] * 6 +
                                     [C.POINTER(C.c_int)] * 3 + [C.c_int])
        lib.hco_walk_ray.restype = C.c_int
        lib.hco_acc_position.argtypes = [C.POINTER(Setup), C.POINTER(Rng), _dp]
        lib.hco_bbfreq.argtypes = [C.POINTER(Rng), C.c_double]
        lib.hco_bbfreq.restype = C.c_double
        lib.hco_lucydustfreq.argtypes = [C.POINTER(Dust), C.POINTER(Rng), C.c_double, C.c_int]
        lib.hco_lucydustfreq.restype = C.c_double
        lib.hco_dustfreq.argtypes = [C.POINTER(Dust), C.POINTER(Rng), C.c_double, C.c_int,
                                     C.c_double]
        lib.hco_dustfreq.restype = C.c_double
        lib.hco_kapp.argtypes = [C.POINTER(Dust), C.c_double, C.c_int]
        lib.hco_kapp.restype = C.c_double
        lib.hco_chir.argtypes = [C.POINTER(Dust), C.c_double, C.c_int]
        lib.hco_chir.restype = C.c_double
        lib.hco_isotrp.argtypes = [C.POINTER(Rng), C.c_double, C.c_double, C.c_double, _dp]
        lib.hco_emit_diffusion.argtypes = [C.POINTER(Grid), C.POINTER(Dust), C.POINTER(Setup),
                                           C.POINTER(Tallies), C.POINTER(Rng), C.c_double, _dp,
                                           _ip, _ip, _dp, _dp]
        lib.hco_emit_diffusion.restype = C.c_int
        lib.hco_dusttemp.argtypes = [C.POINTER(Dust)] + [C.c_double] * 5 + [C.c_int]
        lib.hco_dusttemp.restype = C.c_double
        lib.hco_opacset.argtypes = [C.POINTER(Dust), C.c_double, C.POINTER(Opac)]
        lib.hco_stokes.argtypes = [C.POINTER(Scat), C.POINTER(Opac), C.c_int, C.POINTER(Rng)]
        lib.hco_stokespeel.argtypes = [_dp, _dp, _dp, _dp, _dp, _dp, C.c_double, C.c_double,
                                       C.c_double, C.c_double, C.POINTER(Opac), C.c_int]
        lib.hco_findangle3.argtypes = [C.c_double] * 10 + [_dp] * 4
        lib.hco_initp_direction.argtypes = [C.POINTER(Rng), C.c_double, C.c_double, C.c_double,
                                            _dp, _dp, _dp]
        lib.hco_star_point.argtypes = [C.POINTER(Setup), C.POINTER(Rng), C.c_int, _dp, _dp, _dp]
        lib.hco_star_point.restype = C.c_int
        lib.hco_tauint_ray.argtypes = [C.POINTER(Grid), C.POINTER(Setup), C.POINTER(Tallies),
                                       _dp, _dp, _ip, C.c_int, _dp, _dp, C.c_double, _dp]
        lib.hco_tauint_ray.restype = C.c_int
        lib.hco_peel_ray.argtypes = [C.POINTER(Grid), C.POINTER(Setup), C.POINTER(Tallies), _dp,
                                     _ip, C.c_int, _dp, _dp, _dp, C.c_double, C.c_int, C.c_int,
                                     C.c_int]
        lib.hco_atmosfreq_xi.argtypes = [C.POINTER(Setup), C.c_double]
        lib.hco_atmosfreq_xi.restype = C.c_double
        lib.hco_darkening.argtypes = [C.POINTER(Setup), C.c_double]
        lib.hco_darkening.restype = C.c_double
        lib.hco_mulmod.argtypes = [C.c_int32] * 3
        lib.hco_mulmod.restype = C.c_int32
        lib.hco_math.argtypes = [C.c_int, _dp, _dp, _dp, C.c_int64]
        _libs[variant] = lib
    return _libs[variant]


def _p(a, ct=C.c_double):
    return None if a is None else a.ctypes.data_as(C.POINTER(ct))


class Oracle:
    """The oracle bound to a model (hochunk3d_amd.model.Model)."""

    def __init__(self, model, variant: str = "dm", bw: bool = False):
        """bw: Bjorkman-Wood re-emission (ilucy = 0): each absorption updates
        the oracle's copy of the cell temperature (self.tdust) in place and
        re-emits with dustfreq on the model's kapdBint table."""
        self.lib = load(variant)
        m = self.model = model
        f = lambda a: np.asfortranarray(np.asarray(a, np.float64))  # noqa: E731
        self._arrs = {k: f(getattr(m, k)) for k in (
            "r2arr", "tan2arr", "thetarr", "aarr", "barr", "phiarr", "densarr", "tdust",
            "massarr", "u", "aperture2", "coste", "sinte", "cospe", "sinpe")}
        # a Bjorkman-Wood pass writes the temperatures: the oracle's own copy
        self._arrs["tdust"] = np.array(self._arrs["tdust"], order="F", copy=True)
        d = m.dust
        self._darrs = {k: f(getattr(d, k)) for k in (
            "nudust", "kappad", "adust", "gdust", "pldust", "lTint", "nuint", "kapBint",
            "kappap")}
        A, D = self._arrs, self._darrs
        self.grid = Grid(m.nrg, m.ntg, m.npg, _p(A["r2arr"]), _p(A["tan2arr"]),
                         _p(A["thetarr"]), _p(A["aarr"]), _p(A["barr"]), _p(A["phiarr"]),
                         _p(A["densarr"]), _p(A["tdust"]), _p(A["massarr"]), m.cfg.tsub)
        dd = Dust()
        dd.nlammax = d.nudust.shape[0]
        for i in range(NDG):
            dd.nlambda[i] = int(d.nlambda[i])
            dd.kappav[i] = float(d.kappav[i])
        dd.nudust, dd.kappad, dd.adust = _p(D["nudust"]), _p(D["kappad"]), _p(D["adust"])
        dd.gdust, dd.pldust = _p(D["gdust"]), _p(D["pldust"])
        dd.nT, dd.nnu = d.lTint.size, d.nuint.size
        dd.lTint, dd.nuint, dd.kapBint, dd.kappap = (_p(D["lTint"]), _p(D["nuint"]),
                                                     _p(D["kapBint"]), _p(D["kappap"]))
        if bw:
            from hochunk3d_amd.dust import ensure_kapdB
            self._darrs["kapdBint"] = f(ensure_kapdB(d))
            dd.kapdBint = _p(self._darrs["kapdBint"])
        if getattr(m, "diffus", None) is not None:
            # the diffusion approximation: the model's mask, its walls and
            # the Rosseland table of chiR
            from hochunk3d_amd.dust import ensure_kappar
            A["rarr"] = f(m.rarr)
            A["diffus"] = np.asfortranarray(m.diffus, np.int8)
            A["diffdir"] = np.asfortranarray(m.diffdir, np.int8)
            self.grid.rarr = _p(A["rarr"])
            self.grid.diffus = _p(A["diffus"], C.c_int8)
            self.grid.diffdir = _p(A["diffdir"], C.c_int8)
            self._darrs["kappar"] = f(ensure_kappar(d))
            dd.kappar = _p(self._darrs["kappar"])
        self.dust = dd
        c = m.cfg
        s = Setup()
        s.rstar, s.tstar, s.tshock = m.rstar, m.tstar_ev, m.tshock_ev
        s.iplanckst = 1
        s.nspot = 1
        s.accfrac2 = c.accfrac2
        s.accfrac = m.accfrac
        s.rddustmin, s.rmaxdmax, s.rmax = m.rddustmin, m.rmaxdmax, float(m.rarr[-1])
        s.massc, s.mdot = c.massc, m.mdot
        for i in range(NDG):
            s.fmass[i], s.z1[i], s.bflare[i] = m.fmass8[i], m.z1_8[i], m.bflare8[i]
        s.nfreq, s.nmu, s.nph, s.nap, s.npeel, s.nxhst = (c.nfreq, c.nmu, c.nph, c.nap, c.npeel,
                                                          c.nximg)
        s.numin, s.lnurat = m.numin, m.lnurat
        s.u, s.aperture2 = _p(A["u"]), _p(A["aperture2"])
        s.coste, s.sinte, s.cospe, s.sinpe = (_p(A["coste"]), _p(A["sinte"]), _p(A["cospe"]),
                                              _p(A["sinpe"]))
        s.rmaxi = m.rmaxi
        s.imcube = int(c.imcube)
        s.iintmax = 10000
        s.ltot = m.ltot
        s.limb = int(c.limb)
        if getattr(m, "atmosphere", None) is not None:
            a = m.atmosphere
            for k in ("atmosnu", "hnuint", "lognu", "loghnu"):
                self._arrs[k] = f(getattr(a, k))
                setattr(s, k, _p(self._arrs[k]))
            s.nhnu = a.atmosnu.size
        if getattr(m, "limb_tab", None) is not None:
            self._arrs["xmu"], self._arrs["prob"] = (f(v) for v in m.limb_tab)
            s.ntab = self._arrs["xmu"].size
            s.xmu, s.prob = _p(self._arrs["xmu"]), _p(self._arrs["prob"])
            s.delmu = 1.0 / float(s.ntab)
        s.bw = int(bool(bw))
        self.setup = s

    def set_filters(self, filters):
        """Filter curves (hochunk3d_amd.filters.Filters) for the
        filter-convolved images of peeled passes; None removes them."""
        s = self.setup
        if filters is None:
            s.nfilt = 0
            self._filters = None
            return
        self._farrs = (np.ascontiguousarray(filters.nwfilt, np.int32),
                       np.asfortranarray(filters.filtwave, np.float64),
                       np.asfortranarray(filters.filtphi, np.float64))
        s.nfilt, s.nwfiltmax = filters.nfilt, filters.filtwave.shape[0]
        s.nwfilt = _p(self._farrs[0], C.c_int32)
        s.filtwave, s.filtphi = _p(self._farrs[1]), _p(self._farrs[2])
        self._filters = filters

    @property
    def tdust(self) -> np.ndarray:
        """The oracle's cell temperatures (updated by Bjorkman-Wood passes)."""
        return self._arrs["tdust"]

    def set_temperature(self, tdust):
        self._arrs["tdust"][...] = tdust

    def walk_ray(self, x, y, z, ux, uy, uz, ir, it, ip, on_wall=False):
        """Crossings of a straight ray through the empty grid (hco_walk_ray)
        and the last cell: (n, ir, it, ip); n = -1 for a lost ray."""
        cir, cit, cip = C.c_int(ir), C.c_int(it), C.c_int(ip)
        n = self.lib.hco_walk_ray(C.byref(self.grid), float(x), float(y), float(z), float(ux),
                                  float(uy), float(uz), C.byref(cir), C.byref(cit),
                                  C.byref(cip), int(on_wall))
        return n, cir.value, cit.value, cip.value

    def _leaf_tallies(self):
        m = self.model
        c = m.cfg
        out = {
            "dtauabs": np.zeros((m.nrg, m.ntg, m.npg, NDG), order="F"),
            "jmean": np.zeros((m.nrg, m.ntg, m.npg), order="F"),
            "peel": np.zeros((c.nfreq, max(c.npeel, 1), c.nap, NO, 9), order="F"),
            "img_i": np.zeros((max(c.npeel, 1), c.nximg, c.nximg, c.nfreq, NO), order="F"),
        }
        for k in ("img_q", "img_u", "img_v"):
            out[k] = np.zeros((max(c.npeel, 1), c.nximg, c.nximg, c.nfreq), order="F")
        t = Tallies()
        t.dtauabs, t.jmean, t.peel = _p(out["dtauabs"]), _p(out["jmean"]), _p(out["peel"])
        t.img_i, t.img_q, t.img_u, t.img_v = (_p(out[k]) for k in ("img_i", "img_q", "img_u",
                                                                     "img_v"))
        if getattr(self, "_filters", None) is not None:
            nb = (max(c.npeel, 1), c.nximg, c.nximg, self._filters.nfilt)
            for k in ("image_b_i", "image_b_q", "image_b_u", "image_b_v"):
                out[k] = np.zeros(nb, order="F")
                setattr(t, k, _p(out[k]))
        out["_struct"] = t
        return out

    def tauint_ray(self, xyz, u, cell, tau, kapd, albedo, on_wall=False, tallies=None):
        """hco_tauint_ray: (status, end point, end cell, tallies)."""
        out = tallies if tallies is not None else self._leaf_tallies()
        x = np.ascontiguousarray(xyz, dtype=np.float64)
        uu = np.ascontiguousarray(u, dtype=np.float64)
        cl = np.ascontiguousarray(cell, dtype=np.int32)
        kd = np.ascontiguousarray(kapd, dtype=np.float64)
        al = np.ascontiguousarray(albedo, dtype=np.float64)
        end = np.zeros(3)
        st = self.lib.hco_tauint_ray(C.byref(self.grid), C.byref(self.setup),
                                     C.byref(out["_struct"]), _p(x), _p(uu), _p(cl, C.c_int32),
                                     int(on_wall), _p(kd), _p(al), float(tau), _p(end))
        return st, end, tuple(int(v) for v in cl), out

    def peel_ray(self, xyz, cell, kapd, stokes, peelid, nub, iflag=0, i_orig=1, direction=None,
                 on_wall=False, tallies=None):
        """hco_peel_ray towards observer peelid; returns the tallies."""
        out = tallies if tallies is not None else self._leaf_tallies()
        x = np.ascontiguousarray(xyz, dtype=np.float64)
        cl = np.ascontiguousarray(cell, dtype=np.int32)
        kd = np.ascontiguousarray(kapd, dtype=np.float64)
        sv = np.ascontiguousarray(stokes, dtype=np.float64)
        d = np.ascontiguousarray(direction if direction is not None else (1.0, 0.0, 0.0),
                                 dtype=np.float64)
        self.lib.hco_peel_ray(C.byref(self.grid), C.byref(self.setup), C.byref(out["_struct"]),
                              _p(x), _p(cl, C.c_int32), int(on_wall), _p(kd), _p(sv), _p(d),
                              float(nub), int(iflag), int(i_orig), int(peelid))
        return out

    def emit_diffusion(self, xyz, cell, kapd, nstar: float, rng, nabs=None, on_wall=False):
        """hco_emit_diffusion: the packet at xyz in cell (1-based) leaves the
        diffusion layer; updates self.tdust in place.  Returns (species
        1-based, x, y, z, ux, uy, uz, nub, cell, on_wall)."""
        m = self.model
        nab = (np.zeros((m.nrg, m.ntg, m.npg, NDG), order="F") if nabs is None
               else np.asfortranarray(nabs, np.float64))
        t = Tallies()
        t.nabs = _p(nab)
        x = np.ascontiguousarray(xyz, dtype=np.float64)
        cl = np.ascontiguousarray(cell, dtype=np.int32)
        kd = np.ascontiguousarray(kapd, dtype=np.float64)
        ow = C.c_int(int(on_wall))
        out = np.zeros(7)
        id2 = self.lib.hco_emit_diffusion(C.byref(self.grid), C.byref(self.dust),
                                          C.byref(self.setup), C.byref(t), C.byref(rng),
                                          float(nstar), _p(x), _p(cl, C.c_int32), C.byref(ow),
                                          _p(kd), _p(out))
        return (id2, *out[:7], tuple(int(v) for v in cl), bool(ow.value))

    def run(self, np_: int, p_begin: int = 0, p_end: Optional[int] = None, ipeel: int = 0,
            seed1: int = 12345678, seed2: int = 87654321, gen: int = 1, seg0: int = 0,
            trace: bool = False, images: bool = True, lucy: bool = True) -> dict:
        m = self.model
        c = m.cfg
        p_end = np_ if p_end is None else p_end
        shp4 = (m.nrg, m.ntg, m.npg, NDG)
        shp3 = (m.nrg, m.ntg, m.npg)
        out = {
            "dtauabs": np.zeros(shp4, order="F") if lucy else None,
            "nabs": np.zeros(shp4, order="F"),
            "jmean": np.zeros(shp3, order="F") if lucy else None,
            "killed": np.zeros(shp3, np.int32, order="F"),
            "dbig": np.zeros(shp3, np.int32, order="F"),
            "sed": np.zeros((c.nfreq, c.nmu, c.nph, c.nap, NO, 9), order="F"),
            "peel": np.zeros((c.nfreq, max(c.npeel, 1), c.nap, NO, 9), order="F") if ipeel else None,
            "si_init": np.zeros((c.nfreq, 5), order="F"),
        }
        cube = bool(ipeel and c.imcube and images)
        nimg = (max(c.npeel, 1), c.nximg, c.nximg, c.nfreq)
        out["img_i"] = np.zeros(nimg + (NO,), order="F") if cube else None
        for k in ("img_q", "img_u", "img_v"):
            out[k] = np.zeros(nimg, order="F") if cube else None
        t = Tallies()
        t.dtauabs, t.nabs, t.jmean = _p(out["dtauabs"]), _p(out["nabs"]), _p(out["jmean"])
        t.killed, t.dbig = _p(out["killed"], C.c_int32), _p(out["dbig"], C.c_int32)
        t.sed, t.peel, t.si_init = _p(out["sed"]), _p(out["peel"]), _p(out["si_init"])
        t.img_i, t.img_q, t.img_u, t.img_v = (_p(out[k]) for k in ("img_i", "img_q", "img_u",
                                                                     "img_v"))
        if ipeel and getattr(self, "_filters", None) is not None:
            nb = (max(c.npeel, 1), c.nximg, c.nximg, self._filters.nfilt)
            for k in ("image_b_i", "image_b_q", "image_b_u", "image_b_v"):
                out[k] = np.zeros(nb, order="F")
                setattr(t, k, _p(out[k]))
        tr = np.zeros(p_end - p_begin, TRACE_DTYPE) if trace else None
        rc = self.lib.hco_run_pass(C.byref(self.grid), C.byref(self.dust), C.byref(self.setup),
                                   C.byref(t), int(np_), int(p_begin), int(p_end), int(ipeel),
                                   int(seed1), int(seed2), int(gen), int(seg0),
                                   tr.ctypes.data if tr is not None else None)
        if rc != 0:
            raise RuntimeError(f"hco_run_pass failed ({rc})")
        out["sed"] = np.moveaxis(out["sed"], -1, 0)
        if out["peel"] is not None:
            out["peel"] = np.moveaxis(out["peel"], -1, 0)
        out["scalars"] = {k: getattr(t, k) for k in SCALARS}
        out["trace"] = tr
        return out

    def run_into(self, out: dict, np_: int, p_begin: int, p_end: int, seg0: int = 0,
                 seed1: int = 12345678, seed2: int = 87654321, gen: int = 1) -> None:
        """A Lucy pass over packets [p_begin, p_end) accumulated into the
        tallies of `out` (from new_tallies); for repeated bounded samples."""
        t = out["_struct"]
        rc = self.lib.hco_run_pass(C.byref(self.grid), C.byref(self.dust), C.byref(self.setup),
                                   C.byref(t), int(np_), int(p_begin), int(p_end), 0,
                                   int(seed1), int(seed2), int(gen), int(seg0), None)
        if rc != 0:
            raise RuntimeError(f"hco_run_pass failed ({rc})")

    def new_tallies(self) -> dict:
        m = self.model
        c = m.cfg
        out = {
            "dtauabs": np.zeros((m.nrg, m.ntg, m.npg, NDG), order="F"),
            "nabs": np.zeros((m.nrg, m.ntg, m.npg, NDG), order="F"),
            "jmean": np.zeros((m.nrg, m.ntg, m.npg), order="F"),
            "killed": np.zeros((m.nrg, m.ntg, m.npg), np.int32, order="F"),
            "dbig": np.zeros((m.nrg, m.ntg, m.npg), np.int32, order="F"),
            "sed": np.zeros((c.nfreq, c.nmu, c.nph, c.nap, NO, 9), order="F"),
            "si_init": np.zeros((c.nfreq, 5), order="F"),
        }
        t = Tallies()
        t.dtauabs, t.nabs, t.jmean = _p(out["dtauabs"]), _p(out["nabs"]), _p(out["jmean"])
        t.killed, t.dbig = _p(out["killed"], C.c_int32), _p(out["dbig"], C.c_int32)
        t.sed, t.si_init = _p(out["sed"]), _p(out["si_init"])
        out["_struct"] = t
        return out

    def tfinal(self, dtauabs, nphot):
        m = self.model
        t2 = np.zeros((m.nrg, m.ntg, m.npg, NDG), order="F")
        d = np.asfortranarray(dtauabs, dtype=np.float64)
        n = self.lib.hco_tfinal(C.byref(self.grid), C.byref(self.dust), _p(d), float(m.ltot),
                                float(nphot), _p(t2))
        return t2, int(n)
