"""NoDifFlux faces, small-depth columns and discharges on the GPU against the
oracle.

NoDifFlux zeroes the horizontal diffusion coefficient of the flagged U/V
faces (ModuleAdvectionDiffusion.F90:2496-2531, non-optimised path) and the
theta-scheme vertical coefficient of the flagged W faces (AD:2741-2745).
SmallDepths switches off vertical diffusion in flagged columns (AD:2680-2699,
2733, 2813-2907).  The reference keeps the horizontal coefficients of the
previous call of the same step (Set_Internal_State, AD:5746-5835), so a
property without NoDifFlux that follows one with it in the same step sees
that property's zeros; the GPU keeps them the same way.  Discharges
(SetDischarges/UnSetDischarges, AD:978-1034, 4032-4120) are set before each
property's call with that property's concentrations and unset after it, as
WP:14763-14830 does.  Bar: |gpu - oracle| <= 1e-12 |oracle| on every cell.
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


def test_flag_arguments_required():
    L, AD, _ = _mods()
    case = make("upwind_flat")
    gpu = AD(case.size, case.worksize, case.geom)
    p = np.array(case.tracers[0], order="F", copy=True)
    for opt in ("NoAdvFlux", "NoDifFlux"):           # without the face flags
        with pytest.raises(L.MgpuError) as ei:
            gpu(p, case.hydro, dict(case.params, **{opt: True}))
        assert ei.value.code == L.MGPU_ERR_BADARG
    assert np.array_equal(p, case.tracers[0])
    gpu.kill()


@pytest.mark.parametrize("name,method,optimize,implicit", [
    ("upwind_land_ragged", C.UPWIND_ORDER1, False, False),
    ("upwind_land_ragged", C.CENTRAL_DIF, False, False),
    ("upwind_land_ragged", C.UPWIND_ORDER1, False, True),
    ("tvd_land_sigma", C.P2_TVD, False, False),
    ("tvd_ragged_opt", C.P2_TVD, True, False),
    ("upwind_land_ragged", C.UPWIND_ORDER1, True, False)])
def test_noadvflux_drop_in(name, method, optimize, implicit):
    """NoAdvFlux (AD:4434-4443, 4804-4813) on every other property: with an
    upwind method the XX coefficients the first property clears (its NoFluxU
    faces, and its NoFluxV positions through the YY branch) stay cleared for
    the rest of the step; TVD recomputes them; the optimised path ignores
    the option."""
    case = make(name)
    hy = _with_options(case, 13, small=False)
    T = max(len(case.tracers), 3)
    tracers = [np.asfortranarray(case.tracers[t % len(case.tracers)] * (1.0 + 0.1 * t))
               for t in range(T)]

    def params_of(s, t):
        p = dict(case.params, AdvMethodH=method, Optimize=optimize,
                 NoAdvFlux=(t + s) % 2 == 0)
        if implicit:
            p.update(ImpExp_AdvXX=float(s % 2 == 0), ImpExp_AdvYY=float(s % 2 == 1))
        return p

    _, AD, _ = _mods()
    gpu = AD(case.size, case.worksize, case.geom)
    inst = ora.Instance(case, nthreads=1)
    pg = [t.copy(order="F") for t in tracers]
    po = [t.copy(order="F") for t in tracers]
    bitwise = True
    for s in range(3):
        for t in range(T):
            prm = params_of(s, t)
            assert inst.step(po[t], first_property=t == 0, now=(s + 1) * prm["DTProp"],
                             hydro=hy, **prm) == 0
            gpu(pg[t], hy, prm, first_property=t == 0)
            err = _rel_err(pg[t], po[t])
            assert err <= RTOL, (s, t, err)
            bitwise &= bool(np.array_equal(pg[t], po[t]))
    gpu.kill()
    inst.close()
    print(f"{name} method={method} optimize={optimize} implicit={implicit} NoAdvFlux: "
          f"bit-identical={bitwise}")


def _discharges(case, rng, T, non_open=None):
    """Five discharges: point, uniform over the column, uniform over an
    explicit range (bypass), a withdrawal, and an ignored one; two of them
    share a cell.  Concentrations per tracer (DischConc is per property)."""
    KUB = case.worksize[5]
    wp = case.hydro["WaterPoints3D"][:, :, KUB] == 1
    ii, jj = np.nonzero(wp)
    pick = rng.choice(len(ii), size=4, replace=False)
    cells = [(int(ii[p]), int(jj[p])) for p in pick]
    cells.insert(2, cells[0])                       # the third shares the first's column
    kinds = ["point", "uniform", "range", "withdraw", "point"]
    ignore = [0, 0, 0, 0, 1]
    d = {n: [] for n in ("flow", "I", "J", "K", "Kmin", "Kmax")}
    conc = [[] for _ in range(T)]
    mf = [[] for _ in range(T)]
    for q, (kind, (i, j)) in enumerate(zip(kinds, cells)):
        if ignore[q]:
            continue
        kf = int(case.geom["KFloorZ"][i, j])
        d["I"].append(i); d["J"].append(j)
        d["K"].append(KUB if kind == "point" else int(rng.integers(kf, KUB + 1)))
        d["Kmin"].append(max(kf, KUB - 2) if kind == "range" else -9999999)
        d["Kmax"].append(KUB if kind == "range" else -9999999)
        d["flow"].append(-0.6 if kind == "withdraw" else float(rng.uniform(0.5, 3.0)))
        for t in range(T):
            conc[t].append(float(rng.uniform(1.0, 20.0)))
            mf[t].append(float(rng.uniform(0.5, 1.5)))
    if non_open is not None:
        i, j, k = non_open
        d["I"].append(i); d["J"].append(j); d["K"].append(k)
        d["Kmin"].append(-9999999); d["Kmax"].append(-9999999)
        d["flow"].append(1.25)
        for t in range(T):
            conc[t].append(9.0)
            mf[t].append(1.0)
        kinds.append("point"); ignore.append(0)
    d["vert"] = [5 if k in ("uniform", "range") else 1 for k in kinds]
    d["ignore"] = ignore
    d["ncells"] = [1] * len(kinds)
    d["bypass"] = [1 if k == "range" else 0 for k in kinds]
    return [dict(d, conc=conc[t], concMF=mf[t]) for t in range(T)]


def _dry_column(case, i, j):
    """The case's hydro with column (i, j) dry: water points that are not open,
    with no computed face on any side (AD:4104-4117 treats a discharge into
    such a cell without dilution)."""
    hy = {n: np.array(v, order="F", copy=True) for n, v in case.hydro.items()}
    hy["OpenPoints3D"][i, j, :] = 0
    for cf, q, cells in (("ComputeFacesU3D", "Wflux_X", ((i, j), (i, j + 1))),
                         ("ComputeFacesV3D", "Wflux_Y", ((i, j), (i + 1, j))),
                         ("ComputeFacesW3D", "Wflux_Z", ((i, j),))):
        for c in cells:
            hy[cf][c[0], c[1], :] = 0
            hy[q][c[0], c[1], :] = 0.0
    return hy


@pytest.mark.parametrize("name,optimize,implicit", [
    ("tvd_land_sigma", False, None), ("tvd_ragged_opt", True, None),
    ("upwind_land_ragged", False, None), ("tvd_ragged_opt", True, "x"),
    ("tvd_land_sigma", False, "y")])
def test_discharges_drop_in(name, optimize, implicit):
    L, AD, _ = _mods()
    case = make(name)
    T = len(case.tracers)
    rng = np.random.default_rng(17)
    KUB = case.worksize[5]
    hy = case.hydro
    non_open = None
    if implicit is None:
        wp = case.hydro["WaterPoints3D"][:, :, KUB] == 1
        ii, jj = np.nonzero(wp)
        i, j = int(ii[len(ii) // 2]), int(jj[len(jj) // 2])
        non_open = (i, j, KUB)
        hy = _dry_column(case, i, j)
    sets = _discharges(case, rng, T, non_open=non_open)
    params = dict(case.params, Optimize=optimize)
    if params["AdvMethodH"] in (C.UPWIND_ORDER2, C.UPWIND_ORDER3) and implicit:
        params["AdvMethodH"] = C.UPWIND_ORDER1
    gpu = AD(case.size, case.worksize, case.geom)
    inst = ora.Instance(case, nthreads=1)
    pg = [np.array(p, order="F", copy=True) for p in case.tracers]
    po = [np.array(p, order="F", copy=True) for p in case.tracers]
    worst, bitwise = 0.0, True
    for s in range(4):
        prm = dict(params)
        if implicit:
            xx = (s % 2 == 0) == (implicit == "x")
            prm.update(ImpExp_AdvXX=float(xx), ImpExp_AdvYY=float(not xx))
        for t in range(T):
            # WP's pattern: set before each property's call, unset after (WP:14763, 14830)
            assert inst.set_discharges(sets[t]) == 0
            assert inst.step(po[t], first_property=t == 0, now=(s + 1) * prm["DTProp"],
                             hydro=hy, **prm) == 0
            assert inst.unset_discharges() == 0
            gpu.set_discharges(sets[t])
            gpu(pg[t], hy, prm, first_property=t == 0)
            gpu.unset_discharges()
            err = _rel_err(pg[t], po[t])
            worst = max(worst, err)
            assert err <= RTOL, (s, t, err)
            bitwise &= bool(np.array_equal(pg[t], po[t]))
    # after unset, a call has no discharges
    q = np.array(po[0], order="F", copy=True)
    assert inst.step(q, first_property=True, now=9 * params["DTProp"], hydro=hy, **params) == 0
    gpu(pg[0], hy, params, first_property=True)
    assert _rel_err(pg[0], q) <= RTOL
    gpu.kill()
    inst.close()
    print(f"{name} optimize={optimize} implicit={implicit} discharges: {worst:.3e} "
          f"bit-identical={bitwise}")


def test_discharges_reject_bad_cells():
    L, AD, _ = _mods()
    case = make("upwind_flat")
    gpu = AD(case.size, case.worksize, case.geom)
    d = dict(flow=[1.0], conc=[1.0], concMF=[1.0], I=[0], J=[1], K=[1], Kmin=[-9999999],
             Kmax=[-9999999], vert=[1], ignore=[0], ncells=[1], bypass=[0])
    with pytest.raises(L.MgpuError) as ei:          # I outside the WorkSize
        gpu.set_discharges(d)
    assert ei.value.code == L.MGPU_ERR_BADARG
    gpu.kill()


@pytest.mark.parametrize("bc,nudge", [(C.MASS_CONSERVATION, False), (C.NULL_GRADIENT, True)])
def test_discharges_with_open_boundary_and_fluxes(bc, nudge):
    """Discharges (in the explicit right-hand side, AD:1702), the open
    boundary rows or the null-gradient pass, cell fluxes and, in one case,
    nudging in the same calls, including discharges into boundary columns."""
    from .test_gpu_boundaries import _boundary_points, _hydro
    _, AD, _ = _mods()
    case = make("tvd_land_sigma")
    T = len(case.tracers)
    rng = np.random.default_rng(23)
    bp = _boundary_points(case, interior=2, seed=bc)
    hy = _hydro(case, bp)
    sets = _discharges(case, rng, T)
    # one more point discharge into a boundary column's top cell
    KUB = case.worksize[5]
    bi, bj = (int(v[0]) for v in np.nonzero((bp == 1) & (hy["WaterPoints3D"][:, :, KUB] == 1)))
    for t, d in enumerate(sets):
        sets[t] = dict(d, flow=d["flow"] + [0.8], I=d["I"] + [bi], J=d["J"] + [bj],
                       K=d["K"] + [KUB], Kmin=d["Kmin"] + [-9999999],
                       Kmax=d["Kmax"] + [-9999999], conc=d["conc"] + [4.0],
                       concMF=d["concMF"] + [1.0], vert=d["vert"] + [1],
                       ignore=d["ignore"] + [0], ncells=d["ncells"] + [1],
                       bypass=d["bypass"] + [0])
    refs = [np.asfortranarray(rng.uniform(0.5, 30.0, size=case.shape)) for _ in range(T)]
    base = dict(case.params, Optimize=False, BoundaryCondition=bc, DecayTime=1800.0,
                CellFluxes=True, AdvectionNudging=nudge, AdvectionNudgingCells=3)
    gpu = AD(case.size, case.worksize, case.geom)
    inst = ora.Instance(case, nthreads=1)
    pg = [np.array(p, order="F", copy=True) for p in case.tracers]
    po = [np.array(p, order="F", copy=True) for p in case.tracers]
    bitwise = True
    names = ("AdvFluxX", "AdvFluxY", "AdvFluxZ", "DifFluxX", "DifFluxY", "DifFluxZ")
    for s in range(3):
        for t in range(T):
            prm = dict(base) if bc == C.NULL_GRADIENT else dict(base, ReferenceProp=refs[t])
            assert inst.set_discharges(sets[t]) == 0
            assert inst.step(po[t], first_property=t == 0, now=(s + 1) * prm["DTProp"],
                             hydro=hy, **prm) == 0
            assert inst.unset_discharges() == 0
            gpu.set_discharges(sets[t])
            gpu(pg[t], hy, prm, first_property=t == 0)
            gpu.unset_discharges()
            err = _rel_err(pg[t], po[t])
            assert err <= RTOL, (s, t, err)
            bitwise &= bool(np.array_equal(pg[t], po[t]))
            got = gpu.fluxes()
            for n in names:
                e = _rel_err(got[n], inst.internal(n))
                assert e <= RTOL, (s, t, n, e)
    gpu.kill()
    inst.close()
    print(f"discharges + boundary {bc} + fluxes, nudging={nudge}: bit-identical={bitwise}")
