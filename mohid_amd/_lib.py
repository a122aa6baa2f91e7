"""ctypes binding of the C ABI declared in include/mohid_gpu.h.

The shared library is built in-tree (mohid_amd/libmohid_gpu.so, see
mohid_amd/build.py).  Loading fails loudly when it is missing: there is no
CPU fallback in the product path.
"""
from __future__ import annotations

import ctypes as ct
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmohid_gpu.so")

_dp = ct.POINTER(ct.c_double)
_ip = ct.POINTER(ct.c_int32)

# status codes (include/mohid_gpu.h)
MGPU_SUCCESS = 0
MGPU_OFF_ERR = 1
MGPU_UNKNOWN = -99
MGPU_ERR_HIP = 100
MGPU_ERR_BADARG = 101
MGPU_ERR_OPTION = 102
MGPU_ERR_UNSUPPORTED = 103
MGPU_ERR_MASKS = 104
MGPU_ERR_STATE = 105
MGPU_ERR_NOMEM = 106
MGPU_ERR_PIVOT = 107
MGPU_PREC_FP64 = 0
MGPU_PREC_FAST = 1
MGPU_PREC_MIXED = 2
MGPU_KERNEL_AUTO = 0
MGPU_KERNEL_FUSED = 1
MGPU_KERNEL_SPLIT = 2
MGPU_KERNEL_LANES = 3
MGPU_COLUMNS_MEMORY = 0
MGPU_COLUMNS_BY_DEPTH = 1
MGPU_COLUMNS_AUTO = 2
MGPU_SLOT_MIN_STEPS = 4
MGPU_NOALLOC_MAX_CELLS = 16777216

STATUS_NAMES = {
    0: "SUCCESS", 1: "OFF_ERR", -99: "UNKNOWN", 100: "ERR_HIP", 101: "ERR_BADARG",
    102: "ERR_OPTION", 103: "ERR_UNSUPPORTED", 104: "ERR_MASKS", 105: "ERR_STATE",
    106: "ERR_NOMEM", 107: "ERR_PIVOT",
}


class Size3D(ct.Structure):
    _fields_ = [(n, ct.c_int32) for n in ("ILB", "IUB", "JLB", "JUB", "KLB", "KUB")]


class Geometry(ct.Structure):
    _fields_ = [("DUX", _dp), ("DVY", _dp), ("DZX", _dp), ("DZY", _dp),
                ("KFloorZ", _ip), ("AreaU", _dp), ("AreaV", _dp), ("DZZ", _dp), ("DWZ", _dp),
                ("device_pointers", ct.c_int32)]


class ADArgs(ct.Structure):
    _fields_ = [
        ("schmidt_H", ct.c_double), ("SchmidtCoef_V", ct.c_double),
        ("SchmidtBackground_V", ct.c_double),
        ("AdvMethodH", ct.c_int32), ("TVDLimitationH", ct.c_int32),
        ("AdvMethodV", ct.c_int32), ("TVDLimitationV", ct.c_int32),
        ("Upwind2H", ct.c_int32), ("Upwind2V", ct.c_int32),
        ("VolumeRelMax", ct.c_double),
        ("AdvectionNudging", ct.c_int32), ("AdvectionNudgingCells", ct.c_int32),
        ("DTProp", ct.c_double),
        ("ImpExp_AdvV", ct.c_double), ("ImpExp_DifV", ct.c_double),
        ("ImpExp_AdvXX", ct.c_double), ("ImpExp_AdvYY", ct.c_double),
        ("ImpExp_DifH", ct.c_double), ("NullDif", ct.c_int32),
        ("Wflux_X", _dp), ("Wflux_Y", _dp), ("Wflux_Z", _dp),
        ("VolumeZOld", _dp), ("VolumeZ", _dp),
        ("OpenPoints3D", _ip), ("LandPoints3D", _ip),
        ("ComputeFacesU3D", _ip), ("ComputeFacesV3D", _ip), ("ComputeFacesW3D", _ip),
        ("Visc_H", _dp), ("Diff_V", _dp), ("WaterPoints3D", _ip),
        ("Optimize", ct.c_int32), ("FirstProperty", ct.c_int32),
        ("CellFluxes", ct.c_int32), ("NoAdvFlux", ct.c_int32),
        ("ReferenceProp", _dp), ("BoundaryCondition", ct.c_int32), ("DecayTime", ct.c_double),
        ("BoundaryPoints2D", _ip), ("PROPOld", _dp),
        ("NoDifFlux", ct.c_int32), ("NoFluxU", _ip), ("NoFluxV", _ip), ("NoFluxW", _ip),
        ("SmallDepths", _ip),
        ("device_pointers", ct.c_int32),
    ]


class TransportOpts(ct.Structure):
    _fields_ = [("ntracers", ct.c_int32), ("precision", ct.c_int32),
                ("docycle_method", ct.c_int32), ("kernel", ct.c_int32),
                ("tracers_per_launch", ct.c_int32), ("explicit_tracers_per_lane", ct.c_int32),
                ("column_tracers_per_lane", ct.c_int32), ("refinements", ct.c_int32)]


# every symbol include/mohid_gpu.h declares
EXPORTS = (
    "mgpu_ad_create", "mgpu_ad_set_geometry", "mgpu_ad_step", "mgpu_ad_destroy",
    "mgpu_ad_zero_pivots", "mgpu_ad_set_discharges", "mgpu_ad_unset_discharges",
    "mgpu_ad_get_fluxes", "mgpu_waterflux_z", "mgpu_compute_volumes", "mgpu_compute_areas",
    "mgpu_thomasz", "mgpu_thomasz_noalloc", "mgpu_thomas3d", "mgpu_transport_status",
    "mgpu_transport_create", "mgpu_transport_set_hydro", "mgpu_transport_set_geometry",
    "mgpu_transport_set_tracer",
    "mgpu_transport_get_tracer", "mgpu_transport_run", "mgpu_transport_tracer_ptr",
    "mgpu_transport_set_lanes", "mgpu_transport_set_kernel", "mgpu_transport_set_column_order",
    "mgpu_transport_stream", "mgpu_transport_destroy", "mgpu_transport_timing",
    "mgpu_transport_kernel_times",
    "mgpu_selfcheck_xdiv", "mgpu_version", "mgpu_last_error",
)

_lib = None


def lib():
    """The loaded library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"HIP library not built: {LIB_PATH} (python -m mohid_amd.build)")
    # One HIP runtime per process: torch ships its own libamdhip64 (same
    # SONAME, different file).  Loading torch first makes this library's
    # libamdhip64.so.7 dependency resolve to that already-loaded runtime, so
    # device tensors and library handles share one runtime.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = ct.CDLL(LIB_PATH)
    vp = ct.c_void_p
    L.mgpu_ad_create.argtypes = [ct.POINTER(Size3D), ct.POINTER(Size3D), ct.c_int32, ct.c_int32,
                                 ct.c_int32, ct.POINTER(Geometry), ct.POINTER(vp)]
    L.mgpu_ad_set_geometry.argtypes = [vp, ct.POINTER(Geometry)]
    L.mgpu_ad_step.argtypes = [vp, ct.POINTER(ADArgs), vp]
    L.mgpu_ad_destroy.argtypes = [vp]
    L.mgpu_ad_zero_pivots.argtypes = [vp, ct.POINTER(ct.c_uint64)]
    L.mgpu_thomasz.argtypes = [ct.POINTER(Size3D), ct.POINTER(Size3D), vp, vp, vp, vp, vp, vp,
                               ct.c_int32, vp]
    L.mgpu_thomasz_noalloc.argtypes = [ct.POINTER(Size3D), ct.POINTER(Size3D), vp, vp, vp, vp,
                                       vp, vp, vp]
    i32 = ct.c_int32
    L.mgpu_ad_set_discharges.argtypes = [vp, vp, vp, vp, vp, vp, vp, vp, vp, i32, vp, vp, vp, vp]
    L.mgpu_ad_unset_discharges.argtypes = [vp]
    L.mgpu_ad_get_fluxes.argtypes = [vp] + [vp] * 6 + [i32]
    sz = ct.POINTER(Size3D)
    L.mgpu_waterflux_z.argtypes = [sz, sz] + [vp] * 8 + [ct.c_double, i32, i32]
    L.mgpu_compute_volumes.argtypes = [sz, sz] + [vp] * 7 + [i32]
    L.mgpu_compute_areas.argtypes = [sz, sz] + [vp] * 11 + [i32]
    L.mgpu_thomas3d.argtypes = [ct.POINTER(Size3D), i32, i32, i32, i32, i32, i32, i32, i32, vp,
                                vp, vp, vp, vp, i32, vp]
    L.mgpu_transport_status.argtypes = [vp]
    L.mgpu_transport_create.argtypes = [ct.POINTER(Size3D), ct.POINTER(Size3D),
                                        ct.POINTER(TransportOpts), ct.POINTER(Geometry),
                                        ct.POINTER(vp)]
    L.mgpu_transport_set_hydro.argtypes = [vp, ct.POINTER(ADArgs)]
    L.mgpu_transport_set_geometry.argtypes = [vp, ct.POINTER(Geometry)]
    L.mgpu_transport_set_tracer.argtypes = [vp, ct.c_int32, vp, ct.c_int32]
    L.mgpu_transport_get_tracer.argtypes = [vp, ct.c_int32, vp, ct.c_int32]
    L.mgpu_transport_run.argtypes = [vp, ct.c_int32, ct.c_int32]
    L.mgpu_transport_tracer_ptr.argtypes = [vp, ct.c_int32, ct.POINTER(vp),
                                            ct.POINTER(ct.c_int64)]
    L.mgpu_transport_stream.argtypes = [vp, ct.POINTER(vp)]
    L.mgpu_transport_set_lanes.argtypes = [vp, ct.c_int32, ct.c_int32, ct.c_int32]
    L.mgpu_transport_set_kernel.argtypes = [vp, ct.c_int32]
    L.mgpu_transport_set_column_order.argtypes = [vp, ct.c_int32]
    L.mgpu_transport_destroy.argtypes = [vp]
    L.mgpu_transport_timing.argtypes = [vp, ct.c_int32, ct.c_int32]
    L.mgpu_transport_kernel_times.argtypes = [vp, ct.POINTER(ct.c_double),
                                              ct.POINTER(ct.c_double), ct.POINTER(ct.c_double),
                                              ct.POINTER(ct.c_int32)]
    L.mgpu_selfcheck_xdiv.argtypes = [ct.c_int64, ct.c_uint64, ct.POINTER(ct.c_uint64),
                                      ct.POINTER(ct.c_uint64)]
    L.mgpu_version.restype = ct.c_char_p
    L.mgpu_last_error.restype = ct.c_char_p
    for name in EXPORTS:
        if name not in ("mgpu_version", "mgpu_last_error"):
            getattr(L, name).restype = ct.c_int
    _lib = L
    return L


class MgpuError(RuntimeError):
    def __init__(self, where: str, code: int):
        msg = f"{where} failed: {STATUS_NAMES.get(code, code)}"
        try:
            detail = lib().mgpu_last_error().decode()
            if detail and code == MGPU_ERR_HIP:
                msg += f" ({detail})"
        except Exception:
            pass
        super().__init__(msg)
        self.code = code


def check(where: str, rc: int) -> None:
    if rc != MGPU_SUCCESS:
        raise MgpuError(where, rc)
