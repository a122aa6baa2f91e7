"""Python mirror of ModuleAdvectionDiffusion's public interface on the GPU.

StartAdvectionDiffusion / AdvectionDiffusion / KillAdvectionDiffusion
(Software/MOHIDBase2/ModuleAdvectionDiffusion.F90:400, 1108, 5849) and
THOMASZ_NewType2 (Software/MOHIDBase1/ModuleFunctions.F90:4026) with the
reference's argument names and meaning, calling the HIP library through its
C ABI (include/mohid_gpu.h).  Arrays are numpy arrays in Fortran order
spanning Size (host mode) or device tensors with the same layout (device
mode: torch tensors on the GPU, passed by data pointer).

Errors are raised as MgpuError carrying the library's status code, where the
reference would `stop` or return a non-zero STAT.
"""
from __future__ import annotations

import ctypes as ct
from typing import Dict, Optional, Sequence

import numpy as np

from . import _lib as L
from . import constants as C

__all__ = ["AdvectionDiffusion", "thomasz", "Transport"]

_GEOM_D = ("DUX", "DVY", "DZX", "DZY")
_GEOM_3 = ("AreaU", "AreaV", "DZZ", "DWZ")
_HYDRO_D = ("Wflux_X", "Wflux_Y", "Wflux_Z", "VolumeZOld", "VolumeZ")
_MASKS = ("OpenPoints3D", "LandPoints3D", "ComputeFacesU3D", "ComputeFacesV3D",
          "ComputeFacesW3D")


def _is_device(a) -> bool:
    return hasattr(a, "data_ptr") and getattr(a, "is_cuda", False)


def _addr(a, keep: list, dtype):
    """Address of `a` (numpy: a Fortran-ordered copy of dtype; torch: data_ptr)."""
    if _is_device(a):
        return ct.c_void_p(a.data_ptr())
    arr = np.asfortranarray(a, dtype=dtype)
    keep.append(arr)
    return ct.c_void_p(arr.ctypes.data)


def _geometry(geom: Dict[str, np.ndarray], keep: list) -> L.Geometry:
    """mgpu_geometry from host arrays, or from device tensors when the 3-D
    geometry is on the GPU (2-D metrics are then uploaded as well, so that
    every pointer of the struct lives on the same side)."""
    device = _is_device(geom["DWZ"])
    if device:
        import torch
        dev = geom["DWZ"].device

        def up(n, dt):
            a = geom[n]
            if not _is_device(a):
                a = np.asfortranarray(a, dtype=dt)
                # Fortran (nI, nJ) layout == C layout of the transposed array
                a = torch.from_numpy(np.ascontiguousarray(a.T)).to(dev)
            keep.append(a)
            return a.data_ptr()
        ptr = {n: up(n, np.float64) for n in ("DUX", "DVY", "DZX", "DZY", "AreaU", "AreaV",
                                               "DZZ", "DWZ")}
        ptr["KFloorZ"] = up("KFloorZ", np.int32)
        dp = lambda n: ct.cast(ct.c_void_p(ptr[n]), ct.POINTER(ct.c_double))
        return L.Geometry(dp("DUX"), dp("DVY"), dp("DZX"), dp("DZY"),
                          ct.cast(ct.c_void_p(ptr["KFloorZ"]), ct.POINTER(ct.c_int32)),
                          dp("AreaU"), dp("AreaV"), dp("DZZ"), dp("DWZ"), 1)

    def d(n):
        a = np.asfortranarray(geom[n], dtype=np.float64)
        keep.append(a)
        return a.ctypes.data_as(ct.POINTER(ct.c_double))
    kf = np.asfortranarray(geom["KFloorZ"], dtype=np.int32)
    keep.append(kf)
    return L.Geometry(d("DUX"), d("DVY"), d("DZX"), d("DZY"),
                      kf.ctypes.data_as(ct.POINTER(ct.c_int32)),
                      d("AreaU"), d("AreaV"), d("DZZ"), d("DWZ"), 0)


def _args(hydro: Dict, params: Dict, keep: list, device: bool, optimize: bool,
          first_property: bool) -> L.ADArgs:
    def dp(n):
        v = _addr(hydro[n], keep, np.float64)
        return ct.cast(v, ct.POINTER(ct.c_double))

    def ip(n):
        v = _addr(hydro[n], keep, np.int32)
        return ct.cast(v, ct.POINTER(ct.c_int32))

    def opt_ip(n):
        # optional arrays (NoFluxU/V/W, SmallDepths, BoundaryPoints2D): null when absent
        return ip(n) if hydro.get(n) is not None else ct.POINTER(ct.c_int32)()

    p = params
    # ReferenceProp is per property (WP passes its assimilation field)
    ref = ct.POINTER(ct.c_double)()
    if p.get("ReferenceProp") is not None:
        ref = ct.cast(_addr(p["ReferenceProp"], keep, np.float64), ct.POINTER(ct.c_double))
    # PROPOld (Orlanski_) is written: the caller's own array
    old = ct.POINTER(ct.c_double)()
    if p.get("PROPOld") is not None:
        a = p["PROPOld"]
        if not _is_device(a) and not (isinstance(a, np.ndarray) and a.dtype == np.float64 and
                                      a.flags.f_contiguous):
            raise ValueError("PROPOld must be a Fortran-ordered float64 array (written in place)")
        old = ct.cast(_addr(a, keep, np.float64), ct.POINTER(ct.c_double))
    return L.ADArgs(
        float(p["schmidt_H"]), float(p["SchmidtCoef_V"]), float(p["SchmidtBackground_V"]),
        int(p["AdvMethodH"]), int(p["TVDLimitationH"]), int(p["AdvMethodV"]),
        int(p["TVDLimitationV"]), int(bool(p["Upwind2H"])), int(bool(p["Upwind2V"])),
        float(p["VolumeRelMax"]), int(bool(p.get("AdvectionNudging", False))),
        int(p.get("AdvectionNudgingCells", 0)), float(p["DTProp"]),
        float(p["ImpExp_AdvV"]), float(p["ImpExp_DifV"]), float(p["ImpExp_AdvXX"]),
        float(p["ImpExp_AdvYY"]), float(p["ImpExp_DifH"]), int(bool(p["NullDif"])),
        dp("Wflux_X"), dp("Wflux_Y"), dp("Wflux_Z"), dp("VolumeZOld"), dp("VolumeZ"),
        ip("OpenPoints3D"), ip("LandPoints3D"), ip("ComputeFacesU3D"), ip("ComputeFacesV3D"),
        ip("ComputeFacesW3D"), dp("Visc_H"), dp("Diff_V"), ip("WaterPoints3D"),
        int(bool(optimize)), int(bool(first_property)),
        int(bool(p.get("CellFluxes", False))), int(bool(p.get("NoAdvFlux", False))),
        ref, int(p.get("BoundaryCondition", 0)), float(p.get("DecayTime", C.NULL_REAL)),
        opt_ip("BoundaryPoints2D"), old,
        int(bool(p.get("NoDifFlux", False))),
        opt_ip("NoFluxU"), opt_ip("NoFluxV"), opt_ip("NoFluxW"), opt_ip("SmallDepths"),
        int(bool(device)))


class AdvectionDiffusion:
    """One GPU instance of the AdvectionDiffusion module.

    ``StartAdvectionDiffusion`` <-> constructor, ``AdvectionDiffusion`` <->
    ``__call__`` (PROP advanced in place), ``KillAdvectionDiffusion`` <->
    ``kill``.
    """

    def __init__(self, size: Sequence[int], worksize: Sequence[int], geometry: Dict,
                 docycle_method: int = 1, vertical1d: bool = False, xzflow: bool = False):
        keep: list = []
        g = _geometry(geometry, keep)
        h = ct.c_void_p()
        self._size = L.Size3D(*size)
        self._ws = L.Size3D(*worksize)
        L.check("mgpu_ad_create",
                L.lib().mgpu_ad_create(ct.byref(self._size), ct.byref(self._ws),
                                       int(docycle_method), int(vertical1d), int(xzflow),
                                       ct.byref(g), ct.byref(h)))
        self._h = h

    def __call__(self, prop, hydro: Dict, params: Dict, optimize: Optional[bool] = None,
                 first_property: bool = True) -> None:
        if self._h is None:
            raise L.MgpuError("AdvectionDiffusion", L.MGPU_OFF_ERR)
        keep: list = []
        device = _is_device(prop)
        if optimize is None:
            optimize = bool(params.get("Optimize", False))
        args = _args(hydro, params, keep, device, optimize, first_property)
        if device:
            addr = ct.c_void_p(prop.data_ptr())
        else:
            if not (isinstance(prop, np.ndarray) and prop.flags.f_contiguous
                    and prop.dtype == np.float64):
                raise TypeError("PROP must be a Fortran-ordered float64 array")
            addr = ct.c_void_p(prop.ctypes.data)
        L.check("AdvectionDiffusion", L.lib().mgpu_ad_step(self._h, ct.byref(args), addr))

    def set_geometry(self, geometry: Dict) -> None:
        """The geometry of the following calls (the reference fetches DZZ, DWZ,
        AreaU and AreaV on every call, AD:1353-1401; ModuleGeometry updates
        them every step as the free surface moves)."""
        keep: list = []
        g = _geometry(geometry, keep)
        L.check("mgpu_ad_set_geometry", L.lib().mgpu_ad_set_geometry(self._h, ct.byref(g)))

    def set_discharges(self, d: Dict) -> None:
        """SetDischarges (AD:978-1034) from a dict of arrays: per discharge
        cell flow, conc, concMF, I, J, K, Kmin, Kmax; per discharge vert,
        ignore, ncells, bypass (see mgpu_ad_set_discharges)."""
        f = lambda n: np.ascontiguousarray(d[n], dtype=np.float64)
        i = lambda n: np.ascontiguousarray(d[n], dtype=np.int32)
        a = {n: f(n) for n in ("flow", "conc", "concMF")}
        a.update({n: i(n) for n in ("I", "J", "K", "Kmin", "Kmax", "vert", "ignore", "ncells",
                                    "bypass")})
        p = {n: ct.c_void_p(v.ctypes.data) for n, v in a.items()}
        L.check("SetDischarges", L.lib().mgpu_ad_set_discharges(
            self._h, p["flow"], p["conc"], p["I"], p["J"], p["K"], p["Kmin"], p["Kmax"],
            p["vert"], len(a["vert"]), p["ignore"], p["ncells"], p["bypass"], p["concMF"]))

    def unset_discharges(self) -> None:
        L.check("UnSetDischarges", L.lib().mgpu_ad_unset_discharges(self._h))

    def fluxes(self, like=None) -> Dict:
        """GetAdvFlux / GetDifFlux (AD:697-860): the six cell-flux arrays of the
        last call made with CellFluxes, as Fortran-ordered arrays over Size
        (or device tensors shaped like `like` when it is one)."""
        names = ("AdvFluxX", "AdvFluxY", "AdvFluxZ", "DifFluxX", "DifFluxY", "DifFluxZ")
        if like is not None and _is_device(like):
            import torch
            out = {n: torch.empty_like(like) for n in names}
            ptrs = [ct.c_void_p(out[n].data_ptr()) for n in names]
            dev = 1
        else:
            z = self._size
            shape = (z.IUB - z.ILB + 1, z.JUB - z.JLB + 1, z.KUB - z.KLB + 1)
            out = {n: np.zeros(shape, order="F") for n in names}
            ptrs = [ct.c_void_p(out[n].ctypes.data) for n in names]
            dev = 0
        L.check("GetAdvFlux", L.lib().mgpu_ad_get_fluxes(self._h, *ptrs, dev))
        return out

    def zero_pivots(self) -> int:
        c = ct.c_uint64()
        L.check("mgpu_ad_zero_pivots", L.lib().mgpu_ad_zero_pivots(self._h, ct.byref(c)))
        return int(c.value)

    def kill(self) -> None:
        if self._h is not None:
            L.check("KillAdvectionDiffusion", L.lib().mgpu_ad_destroy(self._h))
            self._h = None

    def __del__(self):
        try:
            self.kill()
        except Exception:
            pass


def thomasz(size, worksize, D, E, F, TI, waterpoints, res) -> None:
    """THOMASZ_NewType2: solve the water columns into ``res`` (in place)."""
    keep: list = []
    dev = _is_device(res)
    addrs = [_addr(a, keep, np.float64) for a in (D, E, F, TI)]
    wp = _addr(waterpoints, keep, np.int32)
    if not dev and not (isinstance(res, np.ndarray) and res.flags.f_contiguous
                        and res.dtype == np.float64):
        raise TypeError("RES must be a Fortran-ordered float64 array")
    r = ct.c_void_p(res.data_ptr() if dev else res.ctypes.data)
    s = L.Size3D(*size)
    w = L.Size3D(*worksize)
    L.check("THOMASZ_NewType2", L.lib().mgpu_thomasz(ct.byref(s), ct.byref(w), *addrs, wp, r,
                                                     int(dev), None))


def thomas3d(size, IJmin, IJmax, JImin, JImax, Kmin, Kmax, di, dj, D, E, F, TI, res) -> None:
    """THOMAS_3D (MF:3667-3870): the tridiagonal lines along j (di=0, dj=1)
    or i (di=1, dj=0), rows JImin..JImax+1, into ``res`` (in place).  Raises
    MgpuError with MGPU_ERR_PIVOT at a zero pivot, where the reference stops."""
    keep: list = []
    dev = _is_device(res)
    addrs = [_addr(a, keep, np.float64) for a in (D, E, F, TI)]
    if not dev and not (isinstance(res, np.ndarray) and res.flags.f_contiguous
                        and res.dtype == np.float64):
        raise TypeError("RES must be a Fortran-ordered float64 array")
    r = ct.c_void_p(res.data_ptr() if dev else res.ctypes.data)
    s = L.Size3D(*size)
    L.check("THOMAS_3D", L.lib().mgpu_thomas3d(ct.byref(s), int(IJmin), int(IJmax), int(JImin),
                                               int(JImax), int(Kmin), int(Kmax), int(di),
                                               int(dj), *addrs, r, int(dev), None))


class Transport:
    """Batched, device-resident transport of several tracers that share one
    hydrodynamic state (the tracer loop of
    ModuleWaterProperties::Advection_Diffusion_Processes, WP:14520-15169)."""

    def __init__(self, size, worksize, geometry: Dict, ntracers: int, precision: int = 0,
                 docycle_method: int = 1, kernel: int = 0, tracers_per_launch: int = 0,
                 explicit_tracers_per_lane: int = 0, column_tracers_per_lane: int = 0,
                 refinements: int = 0):
        keep: list = []
        g = _geometry(geometry, keep)
        opts = L.TransportOpts(int(ntracers), int(precision), int(docycle_method),
                               int(kernel), int(tracers_per_launch),
                               int(explicit_tracers_per_lane), int(column_tracers_per_lane),
                               int(refinements))
        h = ct.c_void_p()
        self._size = L.Size3D(*size)
        self._ws = L.Size3D(*worksize)
        L.check("mgpu_transport_create",
                L.lib().mgpu_transport_create(ct.byref(self._size), ct.byref(self._ws),
                                              ct.byref(opts), ct.byref(g), ct.byref(h)))
        self._h = h
        self.ntracers = ntracers
        self.column_order = L.MGPU_COLUMNS_AUTO

    def set_hydro(self, hydro: Dict, params: Dict) -> None:
        keep: list = []
        dev = _is_device(hydro["Wflux_X"])
        args = _args(hydro, params, keep, dev, params.get("Optimize", False), True)
        L.check("mgpu_transport_set_hydro", L.lib().mgpu_transport_set_hydro(self._h,
                                                                             ct.byref(args)))

    def set_geometry(self, geometry: Dict) -> None:
        """The geometry of the following steps (as AdvectionDiffusion.set_geometry)."""
        keep: list = []
        g = _geometry(geometry, keep)
        L.check("mgpu_transport_set_geometry",
                L.lib().mgpu_transport_set_geometry(self._h, ct.byref(g)))

    def set_tracer(self, index: int, prop) -> None:
        keep: list = []
        dev = _is_device(prop)
        L.check("mgpu_transport_set_tracer",
                L.lib().mgpu_transport_set_tracer(self._h, index, _addr(prop, keep, np.float64),
                                                  int(dev)))

    def get_tracer(self, index: int, out=None):
        shape = (self._size.IUB - self._size.ILB + 1, self._size.JUB - self._size.JLB + 1,
                 self._size.KUB - self._size.KLB + 1)
        if out is None:
            out = np.empty(shape, dtype=np.float64, order="F")
        dev = _is_device(out)
        addr = ct.c_void_p(out.data_ptr() if dev else out.ctypes.data)
        L.check("mgpu_transport_get_tracer",
                L.lib().mgpu_transport_get_tracer(self._h, index, addr, int(dev)))
        return out

    def run(self, nsteps: int, use_graph: bool = True) -> None:
        L.check("mgpu_transport_run", L.lib().mgpu_transport_run(self._h, int(nsteps),
                                                                 int(bool(use_graph))))

    def status(self) -> None:
        """Waits for the handle's work; raises where the reference would have
        stopped (a rejected face method, a zero pivot of a line solve)."""
        L.check("mgpu_transport_status", L.lib().mgpu_transport_status(self._h))

    def timing(self, enable: bool, max_steps: int = 64) -> None:
        L.check("mgpu_transport_timing", L.lib().mgpu_transport_timing(self._h, int(enable),
                                                                       int(max_steps)))

    def kernel_times(self):
        """Mean ms per step of (face-conductance kernel, explicit kernel(s),
        vertical kernel(s)) and the number of steps timed."""
        a, b, c, n = ct.c_double(), ct.c_double(), ct.c_double(), ct.c_int32()
        L.check("mgpu_transport_kernel_times",
                L.lib().mgpu_transport_kernel_times(self._h, ct.byref(a), ct.byref(b),
                                                    ct.byref(c), ct.byref(n)))
        return a.value, b.value, c.value, n.value

    def set_lanes(self, tracers_per_launch: int = 0, explicit_tracers_per_lane: int = 0,
                  column_tracers_per_lane: int = 0) -> None:
        """Launch shape (tracers per launch, tracers per lane of the lane
        variant's explicit and column kernels); results do not depend on it."""
        L.check("mgpu_transport_set_lanes",
                L.lib().mgpu_transport_set_lanes(self._h, int(tracers_per_launch),
                                                 int(explicit_tracers_per_lane),
                                                 int(column_tracers_per_lane)))

    def set_kernel(self, kernel: int) -> None:
        """Kernel variant (MGPU_KERNEL_LANES or _FUSED); results do not depend on it."""
        L.check("mgpu_transport_set_kernel", L.lib().mgpu_transport_set_kernel(self._h,
                                                                               int(kernel)))

    def set_column_order(self, order: int) -> None:
        """Column walk order (MGPU_COLUMNS_MEMORY, _BY_DEPTH or _AUTO, the
        default); results do not depend on it."""
        L.check("mgpu_transport_set_column_order",
                L.lib().mgpu_transport_set_column_order(self._h, int(order)))
        self.column_order = int(order)

    def stream_handle(self) -> int:
        s = ct.c_void_p()
        L.check("mgpu_transport_stream", L.lib().mgpu_transport_stream(self._h, ct.byref(s)))
        return s.value or 0

    def close(self) -> None:
        if self._h is not None:
            L.check("mgpu_transport_destroy", L.lib().mgpu_transport_destroy(self._h))
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
