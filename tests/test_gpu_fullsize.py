"""Full-size runs on the GPU, checked by properties that hold at any size.

The oracle checks every kernel cell by cell on small grids; at the full
workloads (1024 x 1024 x 40 with 8 tracers through the batched transport, and
512 x 512 x 50 sigma with one tracer through the per-call drop-in) the
closed-basin invariants stand in for it: the synthetic basins are closed
(land halo, no flux through the surface face), the flow is divergence-free
and VolumeZOld = VolumeZ, so the step keeps sum(V * C) of every tracer
(each face flux leaves one cell and enters its neighbour) and keeps a
constant field constant.  Sums are taken in FP64 on the device.
"""
import numpy as np
import pytest

from mohid_amd import fixtures as fx

pytestmark = pytest.mark.gpu

MASS_RTOL = 1e-12       # relative drift of sum(V * C) over a few steps
CONST_RTOL = 1e-12      # max |C - c| / c on water points


def _mods():
    import torch
    from mohid_amd import _lib as L
    from mohid_amd.advection_diffusion import AdvectionDiffusion, Transport
    return torch, L, AdvectionDiffusion, Transport


def _mass(torch, V, wp, C):
    return float(torch.sum(torch.where(wp, V * C, torch.zeros_like(C))).item())


def test_full_size_batched_c3():
    """1024 x 1024 x 40, 8 tracers, 30 % land, ragged KFloor, TVD-SuperBee,
    FP64, 5 steps captured as one hipGraph: each tracer's mass is kept; then
    tracer 0 set to 7 on every water point stays 7."""
    torch, L, _, Transport = _mods()
    torch.cuda.set_device(0)
    case = fx.make_case(fx.config("c3"), device="cuda")
    V = case.hydro["VolumeZ"]
    assert torch.equal(case.hydro["VolumeZOld"], V)
    wp = case.hydro["WaterPoints3D"] == 1
    T = len(case.tracers)
    tr = Transport(case.size, case.worksize, case.geom, T)
    tr.set_hydro(case.hydro, case.params)
    m0 = []
    for t, p in enumerate(case.tracers):
        tr.set_tracer(t, p)
        m0.append(_mass(torch, V, wp, p))
    tr.run(5, use_graph=True)
    tr.status()
    out = torch.empty_like(case.tracers[0])
    drift = []
    for t in range(T):
        tr.get_tracer(t, out)
        torch.cuda.synchronize()
        assert bool(torch.isfinite(torch.where(wp, out, torch.zeros_like(out))).all())
        moved = torch.max(torch.abs(torch.where(wp, out - case.tracers[t],
                                                torch.zeros_like(out)))).item()
        assert moved > 1e-6, (t, moved)          # the steps did transport the tracer
        drift.append(abs(_mass(torch, V, wp, out) - m0[t]) / abs(m0[t]))
    print(f"c3 batched: mass drift per tracer {['%.1e' % d for d in drift]}")
    assert max(drift) <= MASS_RTOL, drift
    const = torch.where(wp, torch.full_like(out, 7.0), case.tracers[0])
    tr.set_tracer(0, const)
    tr.run(3, use_graph=False)
    tr.status()
    tr.get_tracer(0, out)
    torch.cuda.synchronize()
    err = float(torch.max(torch.abs(torch.where(wp, out - 7.0, torch.zeros_like(out)))).item())
    tr.close()
    print(f"c3 batched: constant 7 kept to {err / 7.0:.1e}")
    assert err / 7.0 <= CONST_RTOL


def test_full_size_drop_in_c2():
    """512 x 512 x 50 sigma, one tracer, TVD-SuperBee, per-call drop-in on
    device arrays, 3 calls: mass kept, and a constant field kept."""
    torch, L, AD, _ = _mods()
    torch.cuda.set_device(0)
    case = fx.make_case(fx.config("c2"), device="cuda")
    V = case.hydro["VolumeZ"]
    assert torch.equal(case.hydro["VolumeZOld"], V)
    wp = case.hydro["WaterPoints3D"] == 1
    gpu = AD(case.size, case.worksize, case.geom)
    p = case.tracers[0].clone()
    m0 = _mass(torch, V, wp, p)
    p0 = p.clone()
    const = torch.where(wp, torch.full_like(p, 7.0), p)
    for s in range(3):
        gpu(p, case.hydro, case.params, first_property=True)
        gpu(const, case.hydro, case.params, first_property=False)
    torch.cuda.synchronize()
    gpu.kill()
    moved = torch.max(torch.abs(torch.where(wp, p - p0, torch.zeros_like(p)))).item()
    assert moved > 1e-6, moved
    drift = abs(_mass(torch, V, wp, p) - m0) / abs(m0)
    err = float(torch.max(torch.abs(torch.where(wp, const - 7.0,
                                                torch.zeros_like(const)))).item())
    print(f"c2 drop-in: mass drift {drift:.1e}, constant kept to {err / 7.0:.1e}")
    assert drift <= MASS_RTOL
    assert err / 7.0 <= CONST_RTOL
