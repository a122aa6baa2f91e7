"""CellFluxes on the GPU against the oracle.

With CellFluxes the reference keeps the advective and diffusive flux through
every computed face (ModuleAdvectionDiffusion.F90:3356-3954): the explicit
parts with PROP before the solve, weighted 1 - ImpExp (AD:2776, 2930, 3129,
4531, 4900, 5210-5361), the implicit parts with PROP after it (AD:1885-1912),
read back with GetAdvFlux / GetDifFlux (AD:697-860).  Every call of a few
steps and tracers, both Optimize paths, implicit, explicit and theta-weighted
vertical terms; the six arrays and PROP against the oracle, |gpu - oracle| <=
1e-12 |oracle| on every cell.
"""
import numpy as np
import pytest

from mohid_amd import constants as C
from oracle import oracle as ora

from .cases import make

pytestmark = pytest.mark.gpu

RTOL = 1e-12
NAMES = ("AdvFluxX", "AdvFluxY", "AdvFluxZ", "DifFluxX", "DifFluxY", "DifFluxZ")


def _rel_err(a, b):
    den = np.maximum(np.abs(b), 1e-300)
    return float(np.max(np.abs(a - b) / den))


@pytest.mark.parametrize("name,over", [
    ("tvd_land_sigma", {}), ("upwind_land_ragged", {}), ("tvd_ragged_opt", {}),
    ("upwind_flat", dict(ImpExp_DifV=0.5)), ("tvd_sigma", dict(ImpExp_DifV=0.0)),
    ("tvd_land_sigma", dict(ImpExp_AdvV=0.0)),
    ("upwind_land_ragged", dict(ImpExp_AdvV=0.0, ImpExp_DifV=0.3, NullDif=True))])
def test_cell_fluxes_drop_in(name, over):
    from mohid_amd.advection_diffusion import AdvectionDiffusion as AD
    case = make(name)
    gpu = AD(case.size, case.worksize, case.geom)
    inst = ora.Instance(case, nthreads=1)
    pg = [np.array(p, order="F", copy=True) for p in case.tracers]
    po = [np.array(p, order="F", copy=True) for p in case.tracers]
    prm = dict(case.params, CellFluxes=True, **over)
    worst, bitwise = 0.0, True
    for s in range(3):
        for t in range(len(pg)):
            assert inst.step(po[t], first_property=t == 0, now=(s + 1) * prm["DTProp"],
                             **prm) == 0
            gpu(pg[t], case.hydro, prm, first_property=t == 0)
            got = gpu.fluxes()
            err = _rel_err(pg[t], po[t])
            bitwise &= bool(np.array_equal(pg[t], po[t]))
            for n in NAMES:
                want = inst.internal(n)
                e = _rel_err(got[n], want)
                assert e <= RTOL, (s, t, n, e)
                err = max(err, e)
                bitwise &= bool(np.array_equal(got[n], want))
            assert err <= RTOL, (s, t, err)
            worst = max(worst, err)
    gpu.kill()
    inst.close()
    print(f"{name} {over} cell fluxes: {worst:.3e} bit-identical={bitwise}")


def test_cell_fluxes_state():
    """No fluxes before a call with CellFluxes (GetAdvFlux ERR01), none after
    a call without; implicit horizontal advection stays on the CPU."""
    from mohid_amd import _lib as L
    from mohid_amd.advection_diffusion import AdvectionDiffusion as AD
    case = make("upwind_flat")
    gpu = AD(case.size, case.worksize, case.geom)
    with pytest.raises(L.MgpuError) as ei:
        gpu.fluxes()
    assert ei.value.code == L.MGPU_ERR_STATE
    p = np.array(case.tracers[0], order="F", copy=True)
    gpu(p, case.hydro, dict(case.params, CellFluxes=True))
    assert np.max(np.abs(gpu.fluxes()["AdvFluxX"])) > 0
    gpu(p, case.hydro, dict(case.params))
    with pytest.raises(L.MgpuError) as ei:
        gpu.fluxes()
    assert ei.value.code == L.MGPU_ERR_STATE
    with pytest.raises(L.MgpuError) as ei:
        gpu(p, case.hydro, dict(case.params, CellFluxes=True, ImpExp_AdvXX=1.0,
                                AdvMethodH=C.UPWIND_ORDER1))
    assert ei.value.code == L.MGPU_ERR_UNSUPPORTED
    gpu.kill()
