"""Parity of the HIP path with the oracle, through the C ABI, on the GPU.

Every comparison runs the GPU library and the oracle (the reference's
algorithm restated on the CPU) on identical seeded inputs.  The bar is the
north-star tolerance for the FP64 path: |gpu - oracle| <= 1e-12 * |oracle|
on every water cell after every step.  The FP64 kernels keep the reference's
operation order and are compiled without contraction, so the tests also
record whether the results are bit-identical.
"""
import numpy as np
import pytest

from mohid_amd import constants as C
from oracle import oracle as ora

from .cases import GRID, SMALL, make

pytestmark = pytest.mark.gpu

RTOL = 1e-12


def _ad():
    from mohid_amd.advection_diffusion import AdvectionDiffusion, Transport, thomasz
    return AdvectionDiffusion, Transport, thomasz


def _rel_err(a, b, mask):
    a = a[mask]
    b = b[mask]
    den = np.maximum(np.abs(b), 1e-300)
    return float(np.max(np.abs(a - b) / den)) if a.size else 0.0


def _solved_mask(case):
    """Cells the reference writes: rows 1..KUB+1 of columns with
    WaterPoints3D(i,j,KUB) == 1 (ModuleFunctions.F90:4086-4107)."""
    wp = case.hydro["WaterPoints3D"]
    kub = case.worksize[5]
    m = np.zeros(wp.shape, dtype=bool)
    col = wp[:, :, kub] == 1
    m[:, :, 1:kub + 2] = col[:, :, None]
    return m


def test_thomasz_gpu_matches_oracle():
    _, _, thomasz = _ad()
    import ctypes as ct
    rng = np.random.default_rng(5)
    ni, nj, nk = 33, 21, 17
    size = (0, ni + 1, 0, nj + 1, 0, nk + 1)
    ws = (1, ni, 1, nj, 1, nk)
    shape = (ni + 2, nj + 2, nk + 2)
    D = np.asfortranarray(-rng.uniform(0, 1, shape))
    F = np.asfortranarray(-rng.uniform(0, 1, shape))
    E = np.asfortranarray(1.0 + rng.uniform(0, 1, shape) - D - F)
    TI = np.asfortranarray(rng.normal(0, 3, shape))
    WP = np.asfortranarray((rng.uniform(size=shape) < 0.8).astype(np.int32))
    res_gpu = np.asfortranarray(rng.normal(size=shape))
    res_ora = res_gpu.copy(order="F")
    thomasz(size, ws, D, E, F, TI, WP, res_gpu)
    w = np.zeros(nk + 3)
    g = np.zeros(nk + 3)
    dp = ct.POINTER(ct.c_double)
    ora.lib().ora_thomasz(ct.byref(ora.Size3D(*size)), ct.byref(ora.Size3D(*ws)),
                          D.ctypes.data_as(dp), E.ctypes.data_as(dp), F.ctypes.data_as(dp),
                          TI.ctypes.data_as(dp), WP.ctypes.data_as(ct.POINTER(ct.c_int32)),
                          res_ora.ctypes.data_as(dp), w.ctypes.data_as(dp), g.ctypes.data_as(dp))
    assert np.array_equal(res_gpu, res_ora)


@pytest.mark.parametrize("name", sorted(SMALL))
def test_advection_diffusion_matches_oracle(name):
    """Per-tracer AdvectionDiffusion calls (the WaterProperties call pattern,
    WP:14776-14822) for 6 steps, GPU vs oracle after every call."""
    AD, _, _ = _ad()
    case = make(name)
    steps = 6
    gpu = AD(case.size, case.worksize, case.geom)
    inst = ora.Instance(case, nthreads=1)
    p_gpu = [np.array(t, order="F", copy=True) for t in case.tracers]
    p_ora = [np.array(t, order="F", copy=True) for t in case.tracers]
    solved = _solved_mask(case)
    bitwise = True
    for s in range(steps):
        for t in range(len(p_gpu)):
            first = t == 0
            rc = inst.step(p_ora[t], first_property=first, now=(s + 1) * case.params["DTProp"])
            assert rc == 0
            gpu(p_gpu[t], case.hydro, case.params, first_property=first)
            err = _rel_err(p_gpu[t], p_ora[t], solved)
            assert err <= RTOL, (name, s, t, err)
            # cells outside the solved columns are never written
            assert np.array_equal(p_gpu[t][~solved], p_ora[t][~solved])
            bitwise &= bool(np.array_equal(p_gpu[t], p_ora[t]))
    gpu.kill()
    inst.close()
    print(f"{name}: bit-identical={bitwise}")


@pytest.mark.parametrize("variant", [
    dict(ImpExp_AdvV=C.EXPLICIT_SCHEME),
    dict(ImpExp_DifV=C.EXPLICIT_SCHEME),
    dict(ImpExp_DifV=0.5),
    dict(AdvMethodH=C.UPWIND_ORDER2, AdvMethodV=C.UPWIND_ORDER2, ImpExp_AdvV=C.EXPLICIT_SCHEME),
    dict(AdvMethodH=C.UPWIND_ORDER3, AdvMethodV=C.P2_TVD, TVDLimitationV=C.MINMOD),
    dict(AdvMethodH=C.CENTRAL_DIF, AdvMethodV=C.CENTRAL_DIF),
    dict(AdvMethodH=C.P2_TVD, TVDLimitationH=C.VANLEER, TVDLimitationV=C.MUSCL),
    dict(AdvMethodH=C.P2_TVD, TVDLimitationH=C.PDM, AdvMethodV=C.UPWIND_ORDER1),
    dict(NullDif=True),
])
def test_advection_diffusion_options(variant):
    """Method/limiter/theta variants of the non-optimised path."""
    AD, _, _ = _ad()
    case = make("tvd_land_sigma")
    params = dict(case.params)
    params.update(variant)
    params["Optimize"] = False
    params.setdefault("Upwind2H", True)
    params["Upwind2H"] = True
    params["Upwind2V"] = True
    gpu = AD(case.size, case.worksize, case.geom)
    inst = ora.Instance(case, nthreads=1)
    p_gpu = np.array(case.tracers[0], order="F", copy=True)
    p_ora = p_gpu.copy(order="F")
    solved = _solved_mask(case)
    for s in range(4):
        rc = inst.step(p_ora, first_property=True, now=(s + 1) * 30.0, **params)
        assert rc == 0
        gpu(p_gpu, case.hydro, params, first_property=True)
        err = _rel_err(p_gpu, p_ora, solved)
        assert err <= RTOL, (variant, s, err)
    gpu.kill()


def _batched(case, steps_graph, steps_eager, **kw):
    _, Transport, _ = _ad()
    T = len(case.tracers)
    tr = Transport(case.size, case.worksize, case.geom, T, **kw)
    tr.set_hydro(case.hydro, case.params)
    for t, p in enumerate(case.tracers):
        tr.set_tracer(t, p)
    if steps_graph:
        tr.run(steps_graph, use_graph=True)
    if steps_eager:
        tr.run(steps_eager, use_graph=False)
    out = [tr.get_tracer(t) for t in range(T)]
    tr.close()
    return out


@pytest.mark.parametrize("per_launch", [0, 2])
def test_batched_transport_chunks_match_oracle(per_launch):
    """Device-resident multi-tracer transport (hipGraph of 5 steps + 2 eager
    steps) against the oracle's per-tracer Optimize call pattern; with 2
    tracers per launch the 3 tracers run in two chunks."""
    case = make("tvd_ragged_opt")
    got = _batched(case, 5, 2, tracers_per_launch=per_launch)
    ref = ora.run_case(case, 7, nthreads=1)
    solved = _solved_mask(case)
    for t in range(len(got)):
        err = _rel_err(got[t], ref[t], solved)
        assert err <= RTOL, (t, err)
        assert np.array_equal(got[t][~solved], case.tracers[t][~solved])


@pytest.mark.parametrize("kernel", ["lanes", "split", "fused"])
@pytest.mark.parametrize("name", sorted(SMALL))
def test_batched_transport_matches_oracle(name, kernel):
    """Batched FP64 transport on every small case (optimised and plain
    operation orders, flat, sigma, ragged, land, stiff) for 6 steps, with the
    lane-per-(cell, tracer) kernels, the tile kernels and the fused
    column-tile kernel."""
    from mohid_amd import _lib as L
    case = make(name)
    kid = {"lanes": L.MGPU_KERNEL_LANES, "split": L.MGPU_KERNEL_SPLIT,
           "fused": L.MGPU_KERNEL_FUSED}[kernel]
    got = _batched(case, 4, 2, kernel=kid)
    ref = ora.run_case(case, 6, nthreads=1)
    solved = _solved_mask(case)
    bitwise = True
    for t in range(len(got)):
        err = _rel_err(got[t], ref[t], solved)
        assert err <= RTOL, (name, t, err)
        bitwise &= bool(np.array_equal(got[t], ref[t]))
    print(f"{name} ({kernel}): batched bit-identical={bitwise}")


# Fast mode: FP64 with reciprocal-multiply division.  Each replaced division
# differs from the IEEE quotient by at most a few ulp, so after a handful of
# steps the results stay within 1e-11 of the oracle (relative, every water
# cell); the observed error is printed.
RTOL_FAST = 1e-11


@pytest.mark.parametrize("name", sorted(SMALL))
def test_fast_transport_matches_oracle(name):
    from mohid_amd import _lib as L
    case = make(name)
    got = _batched(case, 4, 2, precision=L.MGPU_PREC_FAST)
    ref = ora.run_case(case, 6, nthreads=1)
    solved = _solved_mask(case)
    worst = 0.0
    for t in range(len(got)):
        err = _rel_err(got[t], ref[t], solved)
        worst = max(worst, err)
        assert err <= RTOL_FAST, (name, t, err)
        assert np.array_equal(got[t][~solved], case.tracers[t][~solved])
    print(f"{name}: fast max rel err {worst:.3e}")


def test_batched_graph_replays_agree():
    """A captured step loop replayed several times gives the eager result."""
    _, Transport, _ = _ad()
    case = make("tvd_land_sigma")
    T = len(case.tracers)
    a = Transport(case.size, case.worksize, case.geom, T)
    b = Transport(case.size, case.worksize, case.geom, T)
    for tr in (a, b):
        tr.set_hydro(case.hydro, case.params)
        for t, p in enumerate(case.tracers):
            tr.set_tracer(t, p)
    for rep in range(3):
        a.run(3, use_graph=True)
        b.run(3, use_graph=False)
        for t in range(T):
            assert np.array_equal(a.get_tracer(t), b.get_tracer(t)), (rep, t)
    a.close()
    b.close()


def test_rejects_like_the_reference():
    """Option sets the reference stops on, and branches not offloaded."""
    from mohid_amd import _lib as L
    AD, _, _ = _ad()
    case = make("upwind_flat")
    gpu = AD(case.size, case.worksize, case.geom)
    p = np.array(case.tracers[0], order="F", copy=True)
    bad = [
        (dict(ImpExp_DifH=1.0), L.MGPU_ERR_OPTION),               # AD:1340 ERR02
        (dict(ImpExp_AdvXX=1.0, ImpExp_AdvYY=1.0), L.MGPU_ERR_OPTION),  # ERR03
        (dict(ImpExp_AdvXX=0.5), L.MGPU_ERR_OPTION),               # AD:4527 ERR01
        (dict(CellFluxes=True, ImpExp_AdvXX=1.0), L.MGPU_ERR_UNSUPPORTED),
    ]
    for ov, code in bad:
        params = dict(case.params)
        params.update(ov)
        with pytest.raises(L.MgpuError) as ei:
            gpu(p, case.hydro, params)
        assert ei.value.code == code, ov
    # Optimize call without a FirstProperty call
    params = dict(case.params, Optimize=True)
    g2 = AD(case.size, case.worksize, case.geom)
    with pytest.raises(L.MgpuError) as ei:
        g2(p, case.hydro, params, first_property=False)
    assert ei.value.code == L.MGPU_ERR_STATE
    # the input is left untouched by rejected calls
    assert np.array_equal(p, case.tracers[0])
    gpu.kill()
    g2.kill()


def test_device_resident_inputs_match_host_inputs():
    """Fixtures built directly on the device (device-pointer geometry, hydro
    and tracers) give the same transport result as host inputs, and the
    drop-in's device mode updates PROP in place like the host mode."""
    import torch
    from mohid_amd import fixtures as fx
    AD, Transport, _ = _ad()
    cfg = fx.config("c3", ni=30, nj=26, nk=8, kfloor=(2, 8),
                    tracers=("temperature", "lognormal"), **GRID)
    host = fx.make_case(cfg)
    dev = fx.make_case(cfg, device="cuda")
    for t in range(2):
        got = dev.tracers[t].cpu().numpy()                       # (nK, nJ, nI)
        assert np.array_equal(got.transpose(2, 1, 0), host.tracers[t])
    outs = []
    for case in (host, dev):
        tr = Transport(case.size, case.worksize, case.geom, 2)
        tr.set_hydro(case.hydro, case.params)
        for t, p in enumerate(case.tracers):
            tr.set_tracer(t, p)
        tr.run(3, use_graph=True)
        outs.append([tr.get_tracer(t) for t in range(2)])
        tr.close()
    for a, b in zip(*outs):
        assert np.array_equal(a, b)
    # drop-in, device mode: PROP is a device tensor updated in place
    gpu = AD(dev.size, dev.worksize, dev.geom)
    prop = dev.tracers[1].clone()
    gpu(prop, dev.hydro, dev.params, first_property=True)
    ref = np.array(host.tracers[1], order="F", copy=True)
    gpu_h = AD(host.size, host.worksize, host.geom)
    gpu_h(ref, host.hydro, host.params, first_property=True)
    assert np.array_equal(prop.cpu().numpy().transpose(2, 1, 0), ref)
    gpu.kill()
    gpu_h.kill()


def test_every_launch_shape_gives_identical_results():
    """The launch shapes (lane pair or fused walk, tracers per launch, tracers
    per lane in the explicit and column kernels, and the column walk's order:
    by first row or in memory order) change speed, never results: every
    candidate the autotuner can pick reproduces the oracle bit for bit."""
    from mohid_amd import _lib as L
    from mohid_amd.autotune import apply, autotune, candidates
    _, Transport, _ = _ad()
    case = make("tvd_ragged_opt", tracers=("temperature", "salinity") + ("lognormal",) * 6)
    T = len(case.tracers)
    ref = ora.run_case(case, 2, nthreads=1)
    tr = Transport(case.size, case.worksize, case.geom, T)
    tr.set_hydro(case.hydro, case.params)
    # the automatic order switches to the depth-ordered walk part-way through
    for order in (L.MGPU_COLUMNS_BY_DEPTH, L.MGPU_COLUMNS_MEMORY, L.MGPU_COLUMNS_AUTO):
        tr.set_column_order(order)
        for shape in candidates(T):
            apply(tr, shape)
            for t, p in enumerate(case.tracers):
                tr.set_tracer(t, p)
            tr.run(1, use_graph=True)
            tr.run(1, use_graph=False)
            for t in range(T):
                assert np.array_equal(tr.get_tracer(t), ref[t]), (order, shape, t)
    res = autotune(tr, steps=1)
    assert tuple(res["shape"]) in candidates(T)
    tr.close()


@pytest.mark.parametrize("shape", [
    dict(base="c3", ni=9, nj=7, nk=5, land_frac=1.0),                 # no water at all
    dict(base="c1", ni=1, nj=2, nk=6),                                # one pair of columns
    dict(base="c2", ni=12, nj=10, nk=1, depth=(5.0, 30.0)),           # one layer (2-D)
    dict(base="c3", ni=70, nj=9, nk=11, kfloor=(1, 11), land_frac=0.5,
         tracers=("temperature", "salinity", "lognormal")),           # ragged, two i-tiles
])
def test_edge_shapes_match_oracle(shape):
    """Degenerate and edge domains through the drop-in and the batched path."""
    from mohid_amd import fixtures as fx
    AD, Transport, _ = _ad()
    shape = dict(shape)
    case = fx.make_case(fx.config(shape.pop("base"), **dict(GRID, **shape)))
    T = len(case.tracers)
    ref = ora.run_case(case, 3, nthreads=1)
    got = _batched(case, 2, 1)
    gpu = AD(case.size, case.worksize, case.geom)
    drop = [np.array(t, order="F", copy=True) for t in case.tracers]
    for s in range(3):
        for t in range(T):
            gpu(drop[t], case.hydro, case.params, first_property=(t == 0))
    gpu.kill()
    for t in range(T):
        assert np.array_equal(got[t], ref[t]), t
        assert np.array_equal(drop[t], ref[t]), t


@pytest.mark.parametrize("name", ["tvd_ragged_opt", "tvd_land_sigma"])
def test_moving_surface_matches_oracle(name):
    """A moving free surface: DWZ, DZZ, AreaU, AreaV, VolumeZ and VolumeZOld
    change every step (fixtures.moving_surface), set before each step with
    mgpu_ad_set_geometry (drop-in) and mgpu_transport_set_geometry (batched
    handle, one graph-captured step per geometry) as the reference fetches
    them inside every call (AD:1353-1401); bit-identical to the oracle.  A
    second handle walks the columns by depth and takes each step's geometry
    after its hydrodynamic state, so that set_geometry refreshes the walk's
    level-major copies in place."""
    from mohid_amd import _lib as L
    from mohid_amd import fixtures as fx
    AD, Transport, _ = _ad()
    case = make(name)
    T = len(case.tracers)
    steps = 4
    geoms = fx.moving_surface(case, steps)
    dt = case.params["DTProp"]
    inst = ora.Instance(case, nthreads=1)
    ref = [np.array(t, order="F", copy=True) for t in case.tracers]
    gpu = AD(case.size, case.worksize, case.geom)
    drop = [np.array(t, order="F", copy=True) for t in case.tracers]
    tr = Transport(case.size, case.worksize, case.geom, T)
    td = Transport(case.size, case.worksize, case.geom, T)
    td.set_column_order(L.MGPU_COLUMNS_BY_DEPTH)
    for t, p in enumerate(case.tracers):
        tr.set_tracer(t, p)
        td.set_tracer(t, p)
    for s, (geom, hy) in enumerate(geoms):
        inst.set_geometry(geom)
        gpu.set_geometry(geom)
        tr.set_geometry(geom)
        tr.set_hydro(hy, case.params)
        tr.run(1, use_graph=True)
        td.set_hydro(hy, case.params)
        td.set_geometry(geom)
        td.run(1, use_graph=True)
        for t in range(T):
            assert inst.step(ref[t], first_property=t == 0, now=(s + 1) * dt, hydro=hy) == 0
            gpu(drop[t], hy, case.params, first_property=t == 0)
            assert np.array_equal(drop[t], ref[t]), (s, t)
    for t in range(T):
        assert np.array_equal(tr.get_tracer(t), ref[t]), t
        assert np.array_equal(td.get_tracer(t), ref[t]), t
    inst.close()
    gpu.kill()
    tr.close()
    td.close()
    fixed = ora.run_case(case, steps, nthreads=1)
    assert not np.array_equal(fixed[0], ref[0])
