"""Upstream producers of the transport's inputs on the GPU.

Mirrors of ModuleHydrodynamic's continuity (WaterFlux_Z_VarGrid /
WaterFlux_Z_FixGrid, ModuleHydrodynamic.F90:43270-43343) and ModuleGeometry's
ComputeVolumes / ComputeAreas (ModuleGeometry.F90:4108-4327), over arrays that
span Size in Fortran order.  Outputs are updated in place.  Host arrays
(numpy, Fortran-ordered) are staged per call; device tensors (the layout of
``fixtures.make_case(..., device="cuda")``) stay on the device, so geometry,
continuity and the batched transport chain without host copies.
"""
from __future__ import annotations

import ctypes as ct
from typing import Optional, Sequence

import numpy as np

from . import _lib as L
from .advection_diffusion import _is_device


def _p(a, dtype, out: bool = False):
    if a is None:
        return None, None
    if _is_device(a):
        return ct.c_void_p(a.data_ptr()), None
    if out:
        if not (isinstance(a, np.ndarray) and a.dtype == dtype and a.flags.f_contiguous):
            raise ValueError("outputs must be Fortran-ordered arrays of the right dtype")
        return ct.c_void_p(a.ctypes.data), a
    arr = np.asfortranarray(a, dtype=dtype)
    return ct.c_void_p(arr.ctypes.data), arr


def _call(name: str, size: Sequence[int], worksize: Sequence[int], arrays, tail):
    keep = []
    ptrs = []
    device = None
    for a, dtype, out in arrays:
        p, k = _p(a, dtype, out)
        keep.append(k)
        ptrs.append(p)
        if a is not None:
            d = _is_device(a)
            if device is None:
                device = d
            elif device != d:
                raise ValueError("mix of host and device arrays")
    L.check(name, getattr(L.lib(), name)(ct.byref(L.Size3D(*size)), ct.byref(L.Size3D(*worksize)),
                                         *ptrs, *tail, int(bool(device))))


def waterflux_z(size, worksize, WaterFlux_X, WaterFlux_Y, WaterFlux_Z, ComputeFaces3D_U,
                ComputeFaces3D_V, WaterPoints3D, DT: float,
                Volume_Z_New=None, Volume_Z_Old=None) -> None:
    """Modify_VerticalWaterFlow's continuity: WaterFlux_Z in place; with the
    volumes (variable grid, HYD:43226-43229) their change is taken out."""
    var = Volume_Z_New is not None
    _call("mgpu_waterflux_z", size, worksize,
          [(Volume_Z_New, np.float64, False), (Volume_Z_Old, np.float64, False),
           (WaterFlux_X, np.float64, False), (WaterFlux_Y, np.float64, False),
           (WaterFlux_Z, np.float64, True), (ComputeFaces3D_U, np.int32, False),
           (ComputeFaces3D_V, np.int32, False), (WaterPoints3D, np.int32, False)],
          [float(DT), int(var)])


def compute_volumes(size, worksize, DWZ, DUX, DVY, VolumeZ, VolumeU, VolumeV,
                    VolumeW=None) -> None:
    """ComputeVolumes (GEO:4108-4231), outputs in place."""
    _call("mgpu_compute_volumes", size, worksize,
          [(DWZ, np.float64, False), (DUX, np.float64, False), (DVY, np.float64, False),
           (VolumeZ, np.float64, True), (VolumeU, np.float64, True),
           (VolumeV, np.float64, True), (VolumeW, np.float64, True)], [])


def compute_areas(size, worksize, DUZ, DVZ, DXX, DYY, AreaU, AreaV, SZZ=None, Coef_U=None,
                  CoefX_U=None, Coef_V=None, CoefX_V=None) -> None:
    """ComputeAreas (GEO:4238-4327), outputs in place; SZZ and the four
    coefficient arrays (indexed by k over Size) switch the impermeability on."""
    _call("mgpu_compute_areas", size, worksize,
          [(DUZ, np.float64, False), (DVZ, np.float64, False), (DXX, np.float64, False),
           (DYY, np.float64, False), (AreaU, np.float64, True), (AreaV, np.float64, True),
           (SZZ, np.float64, False), (Coef_U, np.float64, False), (CoefX_U, np.float64, False),
           (Coef_V, np.float64, False), (CoefX_V, np.float64, False)], [])
