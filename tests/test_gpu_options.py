"""NoDifFlux faces and small-depth columns on the GPU against the oracle.

NoDifFlux zeroes the horizontal diffusion coefficient of the flagged U/V
faces (ModuleAdvectionDiffusion.F90:2496-2531, non-optimised path) and the
theta-scheme vertical coefficient of the flagged W faces (AD:2741-2745).
SmallDepths switches off vertical diffusion in flagged columns (AD:2680-2699,
2733, 2813-2907).  The reference keeps the horizontal coefficients of the
previous call of the same step (Set_Internal_State, AD:5746-5835), so a
property without NoDifFlux that follows one with it in the same step sees
that property's zeros; the GPU keeps them the same way.  Bar: |gpu - oracle|
<= 1e-12 |oracle| on every cell.
"""
import numpy as np
import pytest

from mohid_amd import constants as C
from oracle import oracle as ora

from .cases import make

pytestmark = pytest.mark.gpu

RTOL = 1e-12


def _mods():
    from mohid_amd import _lib as L
    from mohid_amd.advection_diffusion import AdvectionDiffusion, Transport
    return L, AdvectionDiffusion, Transport


def _rel_err(a, b):
    den = np.maximum(np.abs(b), 1e-300)
    return float(np.max(np.abs(a - b) / den))


def _with_options(case, seed, noflux=True, small=True):
    hy = {k: np.array(v, order="F", copy=True) for k, v in case.hydro.items()}
    rng = np.random.default_rng(seed)
    shape = hy["WaterPoints3D"].shape
    if noflux:
        for n in ("NoFluxU", "NoFluxV", "NoFluxW"):
            hy[n] = np.asfortranarray((rng.uniform(size=shape) < 0.3).astype(np.int32))
    if small:
        hy["SmallDepths"] = np.asfortranarray(
            (rng.uniform(size=shape[:2]) < 0.4).astype(np.int32))
    return hy


def _drop_in(case, hy, params_of, steps=4, docycle=1):
    _, AD, _ = _mods()
    gpu = AD(case.size, case.worksize, case.geom, docycle_method=docycle)
    inst = ora.Instance(case, docycle_method=docycle, nthreads=1)
    pg = [np.array(p, order="F", copy=True) for p in case.tracers]
    po = [np.array(p, order="F", copy=True) for p in case.tracers]
    worst, bitwise = 0.0, True
    for s in range(steps):
        for t in range(len(pg)):
            prm = params_of(s, t)
            assert inst.step(po[t], first_property=t == 0, now=(s + 1) * prm["DTProp"],
                             hydro=hy, **prm) == 0
            gpu(pg[t], hy, prm, first_property=t == 0)
            err = _rel_err(pg[t], po[t])
            worst = max(worst, err)
            assert err <= RTOL, (s, t, err)
            bitwise &= bool(np.array_equal(pg[t], po[t]))
    gpu.kill()
    inst.close()
    return worst, bitwise


@pytest.mark.parametrize("name", ["tvd_land_sigma", "upwind_land_ragged", "tvd_ragged_opt"])
def test_nodifflux_drop_in(name):
    """NoDifFlux on random faces, non-optimised path, every tracer of a step."""
    case = make(name)
    hy = _with_options(case, 3, small=False)
    params = dict(case.params, Optimize=False, NoDifFlux=True)
    worst, bitwise = _drop_in(case, hy, lambda s, t: params)
    print(f"{name} NoDifFlux: {worst:.3e} bit-identical={bitwise}")


def test_nodifflux_leaks_to_the_next_property():
    """Tracer 0 with NoDifFlux, tracers 1.. without, same Schmidt number: the
    later calls of the step keep the first call's zeroed coefficients."""
    case = make("tvd_ragged_opt")
    hy = _with_options(case, 5, small=False)
    base = dict(case.params, Optimize=False)
    worst, bitwise = _drop_in(case, hy, lambda s, t: dict(base, NoDifFlux=(t == 0)))
    print(f"NoDifFlux on the first property only: {worst:.3e} bit-identical={bitwise}")


@pytest.mark.parametrize("optimize,docycle", [(False, 1), (True, 1), (True, 2), (False, 2)])
def test_small_depths_drop_in(optimize, docycle):
    case = make("tvd_land_sigma")
    hy = _with_options(case, 7, noflux=False)
    params = dict(case.params, Optimize=optimize)
    worst, bitwise = _drop_in(case, hy, lambda s, t: params, docycle=docycle)
    print(f"SmallDepths optimize={optimize} docycle={docycle}: {worst:.3e} "
          f"bit-identical={bitwise}")


@pytest.mark.parametrize("optimize", [False, True])
def test_options_batched(optimize):
    """Batched transport with small depths (and NoDifFlux on the
    non-optimised path) against the oracle's per-tracer calls."""
    L, _, Transport = _mods()
    case = make("tvd_ragged_opt")
    hy = _with_options(case, 11, noflux=not optimize)
    params = dict(case.params, Optimize=optimize, NoDifFlux=not optimize)
    T = len(case.tracers)
    tr = Transport(case.size, case.worksize, case.geom, T)
    tr.set_hydro(hy, params)
    for t, p in enumerate(case.tracers):
        tr.set_tracer(t, p)
    tr.run(3, use_graph=True)
    tr.run(1, use_graph=False)
    tr.status()
    got = [tr.get_tracer(t) for t in range(T)]
    tr.close()
    inst = ora.Instance(case, nthreads=1)
    ref = [np.array(p, order="F", copy=True) for p in case.tracers]
    for s in range(4):
        for t, p in enumerate(ref):
            assert inst.step(p, first_property=t == 0, now=(s + 1) * params["DTProp"],
                             hydro=hy, **params) == 0
    inst.close()
    for t in range(T):
        assert _rel_err(got[t], ref[t]) <= RTOL, t


def test_noadvflux_stays_on_the_cpu():
    L, AD, _ = _mods()
    case = make("upwind_flat")
    gpu = AD(case.size, case.worksize, case.geom)
    p = np.array(case.tracers[0], order="F", copy=True)
    with pytest.raises(L.MgpuError) as ei:
        gpu(p, case.hydro, dict(case.params, NoAdvFlux=True))
    assert ei.value.code == L.MGPU_ERR_UNSUPPORTED
    with pytest.raises(L.MgpuError) as ei:         # NoDifFlux without the face flags
        gpu(p, case.hydro, dict(case.params, NoDifFlux=True))
    assert ei.value.code == L.MGPU_ERR_BADARG
    assert np.array_equal(p, case.tracers[0])
    gpu.kill()
