"""Advection nudging on the GPU against the oracle.

With AdvectionNudging the reference gathers the advective terms in their own
accumulators (AdvectionTi/D/E/F: HorizontalAdvectionXX/YY_Nudge and
VerticalAdvection_Nudge, ModuleAdvectionDiffusion.F90:3211-3350, 4586-4733,
4957-5118) and adds them to the rows with a weight that rises from
1/AdvectionNudgingCells at the domain's edges to 1 inside (AdvectionNudging,
AD:1989-2068).  Every call of a few steps and properties, nudged and not,
explicit and implicit vertical advection, with NoAdvFlux, cell fluxes and
the open boundary beside it; |gpu - oracle| <= 1e-12 |oracle| on every cell.
"""
import numpy as np
import pytest

from mohid_amd import constants as C
from oracle import oracle as ora

from .cases import GRID, make

pytestmark = pytest.mark.gpu

RTOL = 1e-12
FLUXES = ("AdvFluxX", "AdvFluxY", "AdvFluxZ", "DifFluxX", "DifFluxY", "DifFluxZ")


def _rel_err(a, b):
    den = np.maximum(np.abs(b), 1e-300)
    return float(np.max(np.abs(a - b) / den))


def _mods():
    from mohid_amd import _lib as L
    from mohid_amd.advection_diffusion import AdvectionDiffusion, Transport
    return L, AdvectionDiffusion, Transport


def _flags(case, seed):
    hy = {k: np.array(v, order="F", copy=True) for k, v in case.hydro.items()}
    rng = np.random.default_rng(seed)
    for n in ("NoFluxU", "NoFluxV", "NoFluxW"):
        hy[n] = np.asfortranarray((rng.uniform(size=case.shape) < 0.3).astype(np.int32))
    return hy


@pytest.mark.parametrize("name,cells,over", [
    ("upwind_land_ragged", 3, {}),
    ("upwind_land_ragged", 3, dict(ImpExp_AdvV=0.0)),
    ("tvd_land_sigma", 4, dict(Optimize=False)),
    ("tvd_stretched_stiff", 2, dict(Optimize=False)),
    ("tvd_stretched_stiff", 0, dict(Optimize=False, ImpExp_AdvV=0.0)),
    ("upwind_flat", 5, dict(ImpExp_DifV=0.5, CellFluxes=True)),
    ("tvd_sigma", 3, dict(Optimize=False, ImpExp_AdvV=0.0, CellFluxes=True)),
    # edge bands wider than half the domain: the i < nCells + 1 branch wins
    # where both hold (AD:2021-2035)
    ("upwind_land_ragged", 12, {})])
def test_nudging_drop_in(name, cells, over):
    """Property 0 nudged, property 1 not and with NoAdvFlux (an upwind
    method keeps the coefficients the nudged call made, uncleared), the
    rest nudged; the first step's properties swap roles on the next."""
    _, AD, _ = _mods()
    case = make(name)
    hy = _flags(case, 17)
    T = max(len(case.tracers), 3)
    tracers = [np.asfortranarray(case.tracers[t % len(case.tracers)] * (1.0 + 0.1 * t))
               for t in range(T)]
    gpu = AD(case.size, case.worksize, case.geom)
    inst = ora.Instance(case, nthreads=1)
    pg = [t.copy(order="F") for t in tracers]
    po = [t.copy(order="F") for t in tracers]
    bitwise = True
    for s in range(3):
        for t in range(T):
            nudged = (t != 1) if s % 2 == 0 else (t == 1)
            prm = dict(case.params, **over)
            prm.update(AdvectionNudging=nudged, AdvectionNudgingCells=cells,
                       NoAdvFlux=not nudged and not prm.get("Optimize", False))
            assert inst.step(po[t], first_property=t == 0, now=(s + 1) * prm["DTProp"],
                             hydro=hy, **prm) == 0
            gpu(pg[t], hy, prm, first_property=t == 0)
            err = _rel_err(pg[t], po[t])
            assert err <= RTOL, (s, t, nudged, err)
            bitwise &= bool(np.array_equal(pg[t], po[t]))
            if prm.get("CellFluxes"):
                got = gpu.fluxes()
                for n in FLUXES:
                    e = _rel_err(got[n], inst.internal(n))
                    assert e <= RTOL, (s, t, n, e)
    gpu.kill()
    inst.close()
    print(f"{name} cells={cells} {over} nudging: bit-identical={bitwise}")


def test_nudging_weights_matter():
    """The nudged step differs from the plain one (the weights are not all 1)
    and both match the oracle."""
    _, AD, _ = _mods()
    case = make("upwind_land_ragged")
    out = []
    for nudge in (False, True):
        gpu = AD(case.size, case.worksize, case.geom)
        p = np.array(case.tracers[0], order="F", copy=True)
        prm = dict(case.params, AdvectionNudging=nudge, AdvectionNudgingCells=4)
        gpu(p, case.hydro, prm)
        gpu.kill()
        q = np.array(case.tracers[0], order="F", copy=True)
        inst = ora.Instance(case, nthreads=1)
        assert inst.step(q, first_property=True, now=prm["DTProp"], **prm) == 0
        inst.close()
        assert _rel_err(p, q) <= RTOL
        out.append(p)
    wp = case.hydro["WaterPoints3D"] == 1
    assert np.max(np.abs(out[0] - out[1])[wp]) > 1e-3


def test_nudging_unsupported_branches():
    """Optimize, implicit horizontal advection: MGPU_ERR_UNSUPPORTED, PROP
    untouched; a weight other than 0 or 1 on the vertical advection:
    MGPU_ERR_OPTION (VerticalAdvection_Nudge ERR01); the batched transport
    leaves nudging, NoAdvFlux and cell fluxes to the drop-in."""
    L, AD, Transport = _mods()
    case = make("upwind_land_ragged")
    gpu = AD(case.size, case.worksize, case.geom)
    p = np.array(case.tracers[0], order="F", copy=True)
    base = dict(case.params, AdvectionNudging=True, AdvectionNudgingCells=3)
    for extra, code in ((dict(Optimize=True), L.MGPU_ERR_UNSUPPORTED),
                        (dict(ImpExp_AdvXX=1.0), L.MGPU_ERR_UNSUPPORTED),
                        (dict(ImpExp_AdvV=0.5), L.MGPU_ERR_OPTION)):
        with pytest.raises(L.MgpuError) as ei:
            gpu(p, case.hydro, dict(base, **extra))
        assert ei.value.code == code, extra
    assert np.array_equal(p, case.tracers[0])
    gpu.kill()
    hy = _flags(case, 3)
    for extra in (dict(AdvectionNudging=True, AdvectionNudgingCells=2),
                  dict(NoAdvFlux=True), dict(CellFluxes=True)):
        tr = Transport(case.size, case.worksize, case.geom, 2)
        with pytest.raises(L.MgpuError) as ei:
            tr.set_hydro(hy, dict(case.params, **extra))
        assert ei.value.code == L.MGPU_ERR_UNSUPPORTED, extra
        tr.close()


@pytest.mark.parametrize("bc", [C.MASS_CONSERVATION, C.IMPOSED_VALUE, C.NULL_GRADIENT])
def test_nudging_with_open_boundary(bc):
    """Nudging joins the advective terms before the open boundary rewrites
    the boundary rows (AD:1739-1746), and the NullGradient_ pass follows the
    solve."""
    from .test_gpu_boundaries import _boundary_points, _compare, _hydro
    case = make("tvd_land_sigma")
    bp = _boundary_points(case, interior=2, seed=bc)
    hy = _hydro(case, bp)
    rng = np.random.default_rng(40 + bc)
    refs = [np.asfortranarray(rng.uniform(0.5, 30.0, size=case.shape)) for _ in case.tracers]
    base = dict(case.params, Optimize=False, BoundaryCondition=bc, DecayTime=1800.0,
                AdvectionNudging=True, AdvectionNudgingCells=3)
    worst, bitwise = _compare(case, hy, lambda s, t: base, refs)
    print(f"nudging with boundary condition {bc}: {worst:.3e} bit-identical={bitwise}")


def test_nudging_one_layer():
    """One layer (WorkSize KUB = 1): no vertical terms (AD:1725), the
    horizontal advective terms alone join the rows with the weights."""
    from mohid_amd import fixtures as fx
    _, AD, _ = _mods()
    case = fx.make_case(fx.config("c3", ni=21, nj=17, nk=1, land_frac=0.3,
                                  tracers=("temperature", "lognormal", "salinity"), **GRID))
    gpu = AD(case.size, case.worksize, case.geom)
    inst = ora.Instance(case, nthreads=1)
    pg = [np.array(p, order="F", copy=True) for p in case.tracers]
    po = [np.array(p, order="F", copy=True) for p in case.tracers]
    bitwise = True
    for s in range(3):
        for t in range(len(pg)):
            prm = dict(case.params, Optimize=False, AdvectionNudging=t != 1,
                       AdvectionNudgingCells=4)
            assert inst.step(po[t], first_property=t == 0, now=(s + 1) * prm["DTProp"],
                             **prm) == 0
            gpu(pg[t], case.hydro, prm, first_property=t == 0)
            err = _rel_err(pg[t], po[t])
            assert err <= RTOL, (s, t, err)
            bitwise &= bool(np.array_equal(pg[t], po[t]))
    gpu.kill()
    inst.close()
    print(f"one layer nudging: bit-identical={bitwise}")
