"""The upstream producers on the GPU against the oracle: continuity
(WaterFlux_Z_VarGrid / _FixGrid, ModuleHydrodynamic.F90:43270-43343),
ComputeVolumes and ComputeAreas (ModuleGeometry.F90:4108-4327), on random
inputs with masks, in host and device mode; bit-identical arrays required
(the same operations in the same order)."""
import ctypes as ct

import numpy as np
import pytest

from oracle import oracle as ora

pytestmark = pytest.mark.gpu


def _grid(ni=37, nj=29, nk=11):
    return (0, ni + 1, 0, nj + 1, 0, nk + 1), (1, ni, 1, nj, 1, nk), (ni + 2, nj + 2, nk + 2)


def _dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a.transpose(2, 1, 0) if a.ndim == 3
                                                 else a.T)).cuda()


def _host(t, ndim):
    a = t.cpu().numpy()
    return np.asfortranarray(a.transpose(2, 1, 0) if ndim == 3 else a.T)


@pytest.mark.parametrize("vargrid", [False, True])
@pytest.mark.parametrize("device", [False, True])
def test_waterflux_z(vargrid, device):
    from mohid_amd import producers as pr
    size, ws, shape = _grid()
    rng = np.random.default_rng(3 + vargrid)
    qx = np.asfortranarray(rng.normal(size=shape))
    qy = np.asfortranarray(rng.normal(size=shape))
    cfu = np.asfortranarray((rng.uniform(size=shape) < 0.8).astype(np.int32))
    cfv = np.asfortranarray((rng.uniform(size=shape) < 0.8).astype(np.int32))
    wp = np.asfortranarray((rng.uniform(size=shape) < 0.9).astype(np.int32))
    vn = np.asfortranarray(rng.uniform(1e3, 2e3, size=shape))
    vo = np.asfortranarray(vn * (1 + 0.01 * rng.normal(size=shape)))
    qz0 = np.asfortranarray(rng.normal(size=shape))
    want = qz0.copy(order="F")
    S, W = ora.Size3D(*size), ora.Size3D(*ws)
    ora.lib().ora_waterflux_z(ct.byref(S), ct.byref(W),
                              vn.ctypes.data_as(ora._dp) if vargrid else None,
                              vo.ctypes.data_as(ora._dp) if vargrid else None,
                              qx.ctypes.data_as(ora._dp), qy.ctypes.data_as(ora._dp),
                              want.ctypes.data_as(ora._dp), cfu.ctypes.data_as(ora._ip),
                              cfv.ctypes.data_as(ora._ip), wp.ctypes.data_as(ora._ip), 30.0,
                              int(vargrid))
    vols = dict(Volume_Z_New=vn, Volume_Z_Old=vo) if vargrid else {}
    if device:
        q = _dev(qz0)
        pr.waterflux_z(size, ws, _dev(qx), _dev(qy), q, _dev(cfu), _dev(cfv), _dev(wp), 30.0,
                       **{k: _dev(v) for k, v in vols.items()})
        got = _host(q, 3)
    else:
        got = qz0.copy(order="F")
        pr.waterflux_z(size, ws, qx, qy, got, cfu, cfv, wp, 30.0, **vols)
    assert np.array_equal(got, want)


@pytest.mark.parametrize("device", [False, True])
def test_volumes_and_areas(device):
    from mohid_amd import producers as pr
    size, ws, shape = _grid()
    rng = np.random.default_rng(9)
    dwz = np.asfortranarray(rng.uniform(0.2, 5.0, size=shape))
    dux = np.asfortranarray(rng.uniform(50.0, 200.0, size=shape[:2]))
    dvy = np.asfortranarray(rng.uniform(50.0, 200.0, size=shape[:2]))
    init = [np.asfortranarray(rng.normal(size=shape)) for _ in range(4)]
    want = [a.copy(order="F") for a in init]
    S, W = ora.Size3D(*size), ora.Size3D(*ws)
    ora.lib().ora_compute_volumes(ct.byref(S), ct.byref(W), *[a.ctypes.data_as(ora._dp)
                                                              for a in [dwz, dux, dvy] + want])
    duz = np.asfortranarray(rng.uniform(0.2, 5.0, size=shape))
    dvz = np.asfortranarray(rng.uniform(0.2, 5.0, size=shape))
    szz = np.asfortranarray(rng.normal(0, 2, size=shape))
    cf = [np.ascontiguousarray(rng.uniform(0.5, 1.5, size=shape[2])) for _ in range(4)]
    ainit = [np.asfortranarray(rng.normal(size=shape)) for _ in range(2)]
    awant = [a.copy(order="F") for a in ainit]
    ora.lib().ora_compute_areas(ct.byref(S), ct.byref(W), *[a.ctypes.data_as(ora._dp) for a in
                                                            [duz, dvz, dux, dvy] + awant + [szz] + cf])
    if device:
        import torch
        vol = [_dev(a) for a in init]
        pr.compute_volumes(size, ws, _dev(dwz), _dev(dux), _dev(dvy), *vol)
        got = [_host(v, 3) for v in vol]
        ar = [_dev(a) for a in ainit]
        pr.compute_areas(size, ws, _dev(duz), _dev(dvz), _dev(dux), _dev(dvy), *ar, _dev(szz),
                         *[torch.from_numpy(c).cuda() for c in cf])
        agot = [_host(a, 3) for a in ar]
    else:
        got = [a.copy(order="F") for a in init]
        pr.compute_volumes(size, ws, dwz, dux, dvy, *got)
        agot = [a.copy(order="F") for a in ainit]
        pr.compute_areas(size, ws, duz, dvz, dux, dvy, *agot, szz, *cf)
    for g, w in zip(got + agot, want + awant):
        assert np.array_equal(g, w)
