"""Pins the oracle's upstream producers without the reference.

ModuleHydrodynamic and ModuleGeometry do not build here, so their continuity
(WaterFlux_Z_VarGrid / _FixGrid, ModuleHydrodynamic.F90:43270-43343) and
ComputeVolumes / ComputeAreas (ModuleGeometry.F90:4108-4327) are pinned by
what the mathematics fixes:

* a horizontal transport that is the discrete curl of a stream function
  (integers, so every sum is exact) makes no vertical flux at any level on a
  fixed grid;
* on a variable grid the flux out of a column's top equals the flux through
  its bottom plus the column's summed horizontal convergence minus its summed
  volume change (the recurrence telescoped, summed in another order);
* a uniform grid of power-of-two spacings has VolumeZ = dx dy dz and every
  U/V/W volume equal to it exactly; Coef = 1, CoefX = 0 leaves the areas as
  DUZ DYY and DVZ DXX, and a water level lifts them by CoefX * level.
"""
import ctypes as ct

import numpy as np

from oracle import oracle as ora


def P(a, kind=ora._dp):
    return a.ctypes.data_as(kind)


def _grid(ni=7, nj=6, nk=5):
    size = ora.Size3D(0, ni + 1, 0, nj + 1, 0, nk + 1)
    ws = ora.Size3D(1, ni, 1, nj, 1, nk)
    return size, ws, (ni + 2, nj + 2, nk + 2)


def _stream_flow(shape, rng):
    """Qx(i, j) = psi(i+1, j) - psi(i, j), Qy(i, j) = psi(i, j) - psi(i, j+1)
    on corner points: zero divergence in every cell, all faces computed."""
    psi = rng.integers(-50, 50, size=(shape[0] + 1, shape[1] + 1, shape[2])).astype(np.float64)
    qx = np.asfortranarray(psi[1:, :-1, :] - psi[:-1, :-1, :])
    qy = np.asfortranarray(psi[:-1, :-1, :] - psi[:-1, 1:, :])
    return qx, qy


def _continuity_zero_err():
    size, ws, shape = _grid()
    rng = np.random.default_rng(1)
    qx, qy = _stream_flow(shape, rng)
    one = np.asfortranarray(np.ones(shape, dtype=np.int32))
    qz = np.asfortranarray(np.zeros(shape))
    ora.lib().ora_waterflux_z(ct.byref(size), ct.byref(ws), None, None, P(qx), P(qy), P(qz),
                              P(one, ora._ip), P(one, ora._ip), P(one, ora._ip), 60.0, 0)
    return float(np.max(np.abs(qz)))


def _continuity_telescoped_err():
    size, ws, shape = _grid()
    rng = np.random.default_rng(2)
    qx = np.asfortranarray(rng.normal(size=shape))
    qy = np.asfortranarray(rng.normal(size=shape))
    cfu = np.asfortranarray((rng.uniform(size=shape) < 0.8).astype(np.int32))
    cfv = np.asfortranarray((rng.uniform(size=shape) < 0.8).astype(np.int32))
    wp = np.asfortranarray(np.ones(shape, dtype=np.int32))
    vn = np.asfortranarray(rng.uniform(1e3, 2e3, size=shape))
    vo = np.asfortranarray(vn * (1 + 0.01 * rng.normal(size=shape)))
    qz = np.asfortranarray(np.zeros(shape))
    qz[:, :, 1] = rng.normal(size=shape[:2])          # a bottom flux
    q0 = qz[:, :, 1].copy()
    dt = 30.0
    ora.lib().ora_waterflux_z(ct.byref(size), ct.byref(ws), P(vn), P(vo), P(qx), P(qy), P(qz),
                              P(cfu, ora._ip), P(cfv, ora._ip), P(wp, ora._ip), dt, 1)
    ni, nj, nk = shape[0] - 2, shape[1] - 2, shape[2] - 2
    I, J, K = slice(1, ni + 1), slice(1, nj + 1), slice(1, nk + 1)
    conv = (qx[I, J, K] * cfu[I, J, K] - qx[I, 2:nj + 2, K] * cfu[I, 2:nj + 2, K] +
            qy[I, J, K] * cfv[I, J, K] - qy[2:ni + 2, J, K] * cfv[2:ni + 2, J, K] -
            (vn[I, J, K] - vo[I, J, K]) / dt)
    want = q0[I, J] + conv.sum(axis=2)
    return float(np.max(np.abs(qz[I, J, nk + 1] - want)) / np.max(np.abs(want)))


def _volumes_areas_err():
    size, ws, shape = _grid()
    dz, dx, dy = 0.5, 64.0, 32.0
    dwz = np.asfortranarray(np.full(shape, dz))
    dux = np.asfortranarray(np.full(shape[:2], dx))
    dvy = np.asfortranarray(np.full(shape[:2], dy))
    out = [np.asfortranarray(np.full(shape, -1.0)) for _ in range(4)]
    ora.lib().ora_compute_volumes(ct.byref(size), ct.byref(ws), P(dwz), P(dux), P(dvy),
                                  *[P(o) for o in out])
    vz, vu, vv, vw = out
    I, J, K = slice(1, shape[0] - 1), slice(1, shape[1] - 1), slice(1, shape[2] - 1)
    err = float(np.max(np.abs(vz[I, J, K] - dx * dy * dz)))
    err = max(err, float(np.max(np.abs(vu[I, 2:shape[1] - 1, K] - dx * dy * dz))))
    err = max(err, float(np.max(np.abs(vv[2:shape[0] - 1, J, K] - dx * dy * dz))))
    err = max(err, float(np.max(np.abs(vw[I, J, 2:shape[2] - 1] - dx * dy * dz))))
    # cells outside the loops keep their values
    err = max(err, float(np.max(np.abs(vu[:, 1, :] + 1.0))))
    duz = np.asfortranarray(np.full(shape, 2.0))
    dvz = np.asfortranarray(np.full(shape, 4.0))
    au = np.asfortranarray(np.zeros(shape))
    av = np.asfortranarray(np.zeros(shape))
    nk = shape[2]
    ones, zeros = np.ones(nk), np.zeros(nk)
    szz = np.asfortranarray(np.zeros(shape))
    ora.lib().ora_compute_areas(ct.byref(size), ct.byref(ws), P(duz), P(dvz), P(dux), P(dvy),
                                P(au), P(av), P(szz), P(ones), P(zeros), P(ones), P(zeros))
    err = max(err, float(np.max(np.abs(au[I, J, K] - 2.0 * dy))),
              float(np.max(np.abs(av[I, J, K] - 4.0 * dx))))
    szz[:, :, shape[2] - 2] = -3.0                       # water level 3 at KUB
    cx = np.full(nk, 0.25)
    ora.lib().ora_compute_areas(ct.byref(size), ct.byref(ws), P(duz), P(dvz), P(dux), P(dvy),
                                P(au), P(av), P(szz), P(ones), P(cx), P(ones), P(cx))
    return max(err, float(np.max(np.abs(au[I, J, K] - 2.0 * dy * 1.75))))


def test_continuity_of_a_stream_function_flow_is_zero():
    assert _continuity_zero_err() == 0.0


def test_continuity_telescoped_column_sum():
    assert _continuity_telescoped_err() <= 1e-13


def test_volumes_and_areas_closed_forms():
    assert _volumes_areas_err() == 0.0
