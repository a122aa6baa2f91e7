"""ctypes front-end of the CPU oracle -- TEST INFRASTRUCTURE ONLY.

Loads oracle/liboracle.so (the C restatement, advdiff_oracle.c) and, when it
has been built, oracle/_ref/libmohid_mf_ref.so (the reference's own
ModuleFunctions compiled from source).  Only tests/, __graft_entry__.smoke()
and bench.py's cpu_baseline leg import this module; the product package
never does.
"""
from __future__ import annotations

import ctypes as ct
import os
from typing import Optional

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libmohid_mf_ref.so")

_dp = ct.POINTER(ct.c_double)
_ip = ct.POINTER(ct.c_int32)

# status codes of advdiff_oracle.h
OK, ERR_OPTION, ERR_METHOD, ERR_UNSUPPORTED, ERR_BADARG, ERR_PIVOT = 0, 200, 201, 202, 203, 204


class Size3D(ct.Structure):
    _fields_ = [(n, ct.c_int32) for n in ("ILB", "IUB", "JLB", "JUB", "KLB", "KUB")]


class Geometry(ct.Structure):
    _fields_ = [("DUX", _dp), ("DVY", _dp), ("DZX", _dp), ("DZY", _dp),
                ("KFloorZ", _ip), ("AreaU", _dp), ("AreaV", _dp),
                ("DZZ", _dp), ("DWZ", _dp)]


class Args(ct.Structure):
    _fields_ = [
        ("schmidt_H", ct.c_double), ("SchmidtCoef_V", ct.c_double),
        ("SchmidtBackground_V", ct.c_double),
        ("AdvMethodH", ct.c_int32), ("TVDLimitationH", ct.c_int32),
        ("AdvMethodV", ct.c_int32), ("TVDLimitationV", ct.c_int32),
        ("Upwind2H", ct.c_int32), ("Upwind2V", ct.c_int32),
        ("VolumeRelMax", ct.c_double), ("DTProp", ct.c_double),
        ("ImpExp_AdvV", ct.c_double), ("ImpExp_DifV", ct.c_double),
        ("ImpExp_AdvXX", ct.c_double), ("ImpExp_AdvYY", ct.c_double),
        ("ImpExp_DifH", ct.c_double), ("NullDif", ct.c_int32),
        ("Wflux_X", _dp), ("Wflux_Y", _dp), ("Wflux_Z", _dp),
        ("VolumeZOld", _dp), ("VolumeZ", _dp),
        ("OpenPoints3D", _ip), ("LandPoints3D", _ip),
        ("ComputeFacesU3D", _ip), ("ComputeFacesV3D", _ip), ("ComputeFacesW3D", _ip),
        ("WaterPoints3D", _ip), ("Visc_H", _dp), ("Diff_V", _dp),
        ("Optimize", ct.c_int32), ("FirstProperty", ct.c_int32), ("Now", ct.c_double),
        ("NoDifFlux", ct.c_int32), ("NoFluxU", _ip), ("NoFluxV", _ip), ("NoFluxW", _ip),
        ("SmallDepths", _ip),
        ("ReferenceProp", _dp), ("BoundaryCondition", ct.c_int32), ("DecayTime", ct.c_double),
        ("BoundaryPoints2D", _ip), ("PROPOld", _dp), ("CellFluxes", ct.c_int32),
        ("NoAdvFlux", ct.c_int32),
        ("AdvectionNudging", ct.c_int32),
        ("AdvectionNudgingCells", ct.c_int32),
    ]


def _ptr(a: np.ndarray, kind):
    return a.ctypes.data_as(kind)


_lib = None
_ref = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle library missing: {LIB_PATH} (run make -C oracle)")
        L = ct.CDLL(LIB_PATH)
        L.ora_start.argtypes = [ct.POINTER(Size3D), ct.POINTER(Size3D), ct.c_int, ct.c_int,
                                ct.c_int, ct.c_int, ct.POINTER(Geometry), ct.POINTER(ct.c_void_p)]
        L.ora_start.restype = ct.c_int
        L.ora_set_geometry.argtypes = [ct.c_void_p, ct.POINTER(Geometry)]
        L.ora_set_geometry.restype = ct.c_int
        L.ora_advection_diffusion.argtypes = [ct.c_void_p, ct.POINTER(Args), _dp]
        L.ora_advection_diffusion.restype = ct.c_int
        L.ora_kill.argtypes = [ct.c_void_p]
        L.ora_internal.argtypes = [ct.c_void_p, ct.c_char_p]
        L.ora_internal.restype = _dp
        L.ora_thomasz.argtypes = [ct.POINTER(Size3D), ct.POINTER(Size3D), _dp, _dp, _dp, _dp,
                                  _ip, _dp, _dp, _dp]
        L.ora_set_discharges.argtypes = [ct.c_void_p, ct.c_int, _dp, _dp, _dp, _ip, _ip, _ip,
                                         _ip, _ip, _ip, _ip, _ip, _ip]
        L.ora_set_discharges.restype = ct.c_int
        L.ora_unset_discharges.argtypes = [ct.c_void_p]
        L.ora_unset_discharges.restype = ct.c_int
        L.ora_thomas3d.argtypes = [ct.POINTER(Size3D), ct.c_int, ct.c_int, ct.c_int, ct.c_int,
                                   ct.c_int, ct.c_int, ct.c_int, ct.c_int, _dp, _dp, _dp, _dp,
                                   _dp, _dp, _dp, _ip]
        L.ora_thomas3d.restype = ct.c_int
        L.ora_advection1d_v2.argtypes = [ct.c_int, ct.c_int, ct.c_int, ct.c_double, _dp, _dp,
                                         _dp, _dp, _ip, _dp, _dp, _dp, _dp, ct.c_int, ct.c_int,
                                         ct.c_double, ct.c_int]
        L.ora_advection1d_v2.restype = ct.c_int
        L.ora_advection_face.argtypes = [_dp, _dp, _dp, ct.c_double, ct.c_double, ct.c_double,
                                         ct.c_int, ct.c_int, ct.c_int, ct.c_int, _dp]
        L.ora_advection_face.restype = ct.c_int
        L.ora_orlanski2d.argtypes = [ct.POINTER(Size3D), _dp, _dp, _dp, _ip] + [ct.c_int] * 10 + \
            [ct.c_double] * 3
        L.ora_orlanski2d.restype = ct.c_double
        sz = ct.POINTER(Size3D)
        L.ora_waterflux_z.argtypes = [sz, sz, _dp, _dp, _dp, _dp, _dp, _ip, _ip, _ip, ct.c_double,
                                      ct.c_int]
        L.ora_waterflux_z.restype = None
        L.ora_compute_volumes.argtypes = [sz, sz] + [_dp] * 7
        L.ora_compute_volumes.restype = None
        L.ora_compute_areas.argtypes = [sz, sz] + [_dp] * 11
        L.ora_compute_areas.restype = None
        L.ora_advection1d_tvd_superbee_2.argtypes = [ct.c_int, ct.c_int, ct.c_int, ct.c_double,
                                                     _dp, _dp, _dp, _dp, _ip, _dp, _dp]
        _lib = L
    return _lib


def ref_available() -> bool:
    return os.path.exists(REF_PATH)


def ref():
    """The reference's own ModuleFunctions (compiled from source), or raise."""
    global _ref
    if _ref is None:
        if not ref_available():
            raise RuntimeError(f"reference build missing: {REF_PATH} (run make -C oracle ref)")
        R = ct.CDLL(REF_PATH)
        R.ref_thomasz.argtypes = [_ip, _ip, ct.c_int, ct.c_int, _dp, _dp, _dp, _dp, _ip, _dp,
                                  _dp, _dp]
        R.ref_thomas3d.argtypes = [_ip, _ip, ct.c_int, ct.c_int, ct.c_int, ct.c_int, _dp, _dp,
                                   _dp, _dp, _dp, _dp, _dp]
        R.ref_advection1d_v2.argtypes = [ct.c_int, ct.c_int, ct.c_int, ct.c_double, _dp, _dp,
                                         _dp, _dp, _ip, _dp, _dp, _dp, _dp, ct.c_int, ct.c_int,
                                         ct.c_double, ct.c_int]
        R.ref_advection1d_tvd_superbee_2.argtypes = [ct.c_int, ct.c_int, ct.c_int, ct.c_double,
                                                     _dp, _dp, _dp, _dp, _ip, _dp, _dp]
        R.ref_advection_face.argtypes = [_dp, _dp, _dp, ct.c_double, ct.c_double, ct.c_double,
                                         ct.c_int, ct.c_int, ct.c_int, ct.c_int, _dp]
        R.ref_orlanski2d.argtypes = [_ip, _dp, _dp, _dp, _ip] + [ct.c_int] * 10 + \
            [ct.c_double] * 3 + [_dp]
        _ref = R
    return _ref


class Instance:
    """One ModuleAdvectionDiffusion instance (StartAdvectionDiffusion .. Kill)."""

    def __init__(self, case, docycle_method: int = 1, vertical1d: bool = False,
                 xzflow: bool = False, nthreads: Optional[int] = None, geom=None):
        self.case = case
        L = lib()
        self._keep = []
        self.geom = self._geometry(case.geom if geom is None else geom)
        self.size = Size3D(*case.size)
        self.ws = Size3D(*case.worksize)
        if nthreads is None:
            nthreads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        h = ct.c_void_p()
        rc = L.ora_start(ct.byref(self.size), ct.byref(self.ws), docycle_method,
                         int(vertical1d), int(xzflow), nthreads, ct.byref(self.geom),
                         ct.byref(h))
        if rc != 0:
            raise RuntimeError(f"ora_start failed: {rc}")
        self.h = h
        self.now = 0.0

    def _k(self, a):
        self._keep.append(a)
        return a

    def _geometry(self, g):
        """A Geometry struct over copies of the arrays of dict g (kept alive)."""
        return Geometry(*[_ptr(self._k(np.array(g[n], dtype=np.float64, order="F")), _dp)
                          for n in ("DUX", "DVY", "DZX", "DZY")],
                        _ptr(self._k(np.array(g["KFloorZ"], dtype=np.int32, order="F")), _ip),
                        *[_ptr(self._k(np.array(g[n], dtype=np.float64, order="F")), _dp)
                          for n in ("AreaU", "AreaV", "DZZ", "DWZ")])

    def set_geometry(self, geom):
        """The geometry of the following steps (the reference fetches areas and
        distances on every call, AD:1353-1401)."""
        self.geom = self._geometry(geom)
        rc = lib().ora_set_geometry(self.h, ct.byref(self.geom))
        if rc != 0:
            raise RuntimeError(f"ora_set_geometry failed: {rc}")

    def step(self, prop: np.ndarray, first_property: bool = True, now: Optional[float] = None,
             hydro=None, **overrides) -> int:
        """One AdvectionDiffusion call on `prop` (modified in place)."""
        case = self.case
        p = dict(case.params)
        p.update(overrides)
        hy = case.hydro if hydro is None else hydro
        keep = []

        def d(n):
            a = np.asfortranarray(hy[n], dtype=np.float64)
            keep.append(a)
            return _ptr(a, _dp)

        def i(n):
            a = np.asfortranarray(hy[n], dtype=np.int32)
            keep.append(a)
            return _ptr(a, _ip)

        def opt_i(n):
            return i(n) if hy.get(n) is not None else None

        ref = p.get("ReferenceProp")
        if ref is not None:
            ref = np.asfortranarray(ref, dtype=np.float64)
            keep.append(ref)
            ref = _ptr(ref, _dp)
        # PROPOld is written in place (the caller's array, AD:5428)
        old = p.get("PROPOld")
        if old is not None:
            assert old.flags.f_contiguous and old.dtype == np.float64
            old = _ptr(old, _dp)

        args = Args(p["schmidt_H"], p["SchmidtCoef_V"], p["SchmidtBackground_V"],
                    p["AdvMethodH"], p["TVDLimitationH"], p["AdvMethodV"], p["TVDLimitationV"],
                    int(p["Upwind2H"]), int(p["Upwind2V"]), p["VolumeRelMax"], p["DTProp"],
                    p["ImpExp_AdvV"], p["ImpExp_DifV"], p["ImpExp_AdvXX"], p["ImpExp_AdvYY"],
                    p["ImpExp_DifH"], int(p["NullDif"]),
                    d("Wflux_X"), d("Wflux_Y"), d("Wflux_Z"), d("VolumeZOld"), d("VolumeZ"),
                    i("OpenPoints3D"), i("LandPoints3D"), i("ComputeFacesU3D"),
                    i("ComputeFacesV3D"), i("ComputeFacesW3D"), i("WaterPoints3D"),
                    d("Visc_H"), d("Diff_V"), int(p["Optimize"]), int(first_property),
                    self.now if now is None else now, int(bool(p.get("NoDifFlux", False))),
                    opt_i("NoFluxU"), opt_i("NoFluxV"), opt_i("NoFluxW"), opt_i("SmallDepths"),
                    ref, int(p.get("BoundaryCondition", 0)),
                    float(p.get("DecayTime", -9.9e15)), opt_i("BoundaryPoints2D"), old,
                    int(bool(p.get("CellFluxes", False))), int(bool(p.get("NoAdvFlux", False))),
                    int(bool(p.get("AdvectionNudging", False))),
                    int(p.get("AdvectionNudgingCells", 0)))
        assert prop.flags.f_contiguous and prop.dtype == np.float64
        return lib().ora_advection_diffusion(self.h, ct.byref(args), _ptr(prop, _dp))

    def set_discharges(self, d) -> int:
        """SetDischarges (AD:978-1034) from a dict of arrays: per discharge
        cell flow, conc, concMF, I, J, K, Kmin, Kmax; per discharge vert,
        ignore, ncells, bypass (see ora_set_discharges)."""
        f = lambda n: np.ascontiguousarray(d[n], dtype=np.float64)
        i = lambda n: np.ascontiguousarray(d[n], dtype=np.int32)
        arrs = [f("flow"), f("conc"), f("concMF")] + [i(n) for n in ("I", "J", "K", "Kmin",
                                                                      "Kmax", "vert", "ignore",
                                                                      "ncells", "bypass")]
        self._disch = arrs                      # kept alive while set
        ptrs = [_ptr(a, _dp) for a in arrs[:3]] + [_ptr(a, _ip) for a in arrs[3:]]
        return lib().ora_set_discharges(self.h, len(arrs[8]), *ptrs)

    def unset_discharges(self) -> int:
        rc = lib().ora_unset_discharges(self.h)
        self._disch = None
        return rc

    def internal(self, name: str) -> np.ndarray:
        ptr = lib().ora_internal(self.h, name.encode())
        n = int(np.prod(self.case.shape))
        return np.ctypeslib.as_array(ptr, shape=(n,)).reshape(self.case.shape, order="F").copy()

    def close(self):
        if self.h:
            lib().ora_kill(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def host_threads() -> int:
    """Threads the oracle may use here: the cores this process may run on,
    capped by OMP_NUM_THREADS when the host sets it (its CPU share)."""
    n = len(os.sched_getaffinity(0))
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        n = min(n, int(env))
    return max(n, 1)


def run_case(case, steps: int, nthreads: Optional[int] = None, tracers=None):
    """Advance all tracers of `case` by `steps` steps with the oracle (the
    reference's call pattern: one AdvectionDiffusion call per tracer per step,
    FirstProperty on the first, WP:14776-14822).  Returns the new tracers."""
    inst = Instance(case, nthreads=nthreads)
    props = [np.array(t, order="F", copy=True) for t in (case.tracers if tracers is None else tracers)]
    dt = case.params["DTProp"]
    for s in range(steps):
        now = (s + 1) * dt
        for t, prop in enumerate(props):
            rc = inst.step(prop, first_property=(t == 0), now=now)
            if rc != 0:
                raise RuntimeError(f"oracle step failed rc={rc}")
    inst.close()
    return props
