"""Launch-shape autotuner for the batched transport.

A transport step's launches take shape parameters that change speed but not
results:

* the kernel variant: the lane pair (an explicit kernel over the solved
  cells, then a column kernel over the solved columns) or the fused walk
  (one lane per column and tracer group builds each level's explicit
  right-hand side inside its column solve); both share the handle's scratch;
* tracers per launch (all, 4 or 2 of the handle's tracers in one launch: the
  register and cache footprint of a launch against the number of launches);
* tracers per lane of the explicit kernel and of the column kernel (or walk):
  1, 2, 4 or 8 interleaved values read and written as one run by a lane, i.e.
  how many waves the launch has and how many registers each lane holds, which
  sets the waves per SIMD the kernel runs at.

`autotune` times every candidate on the caller's own handle (its
hydrodynamic state and tracers) with HIP events on the handle's stream and
leaves the handle on the fastest shape.  Each candidate runs `steps` timed
steps after one untimed step; the tracer state advances while tuning, so tune
before the run that matters (or reload the tracers afterwards).  A JSON cache
keyed by the GPU's architecture and the problem's shape skips the search on
later calls.
"""
from __future__ import annotations

import itertools
import json
import os
from typing import Dict, Iterable, List, Optional, Tuple

# (kernel variant, tracers per launch, explicit tracers per lane, column
# tracers per lane); the variant ids are include/mohid_gpu.h's MGPU_KERNEL_*
Shape = Tuple[int, int, int, int]

LANES, FUSED = 3, 1
TPL_CHOICES = (1, 2, 4, 8)


def candidates(ntracers: int, per_launch: Iterable[int] = (0, 4, 2),
               kernels: Iterable[int] = (LANES, FUSED)) -> List[Shape]:
    """All valid shapes for `ntracers` tracers: tracers per launch from
    `per_launch` (0 = all, at most 8), tracers per lane dividing the launch's
    chunk (the fused walk takes at most 4; its explicit work has no separate
    shape, so its explicit entry is 0)."""
    out: List[Shape] = []
    for kern in kernels:
        for pl in per_launch:
            chunk = min(pl or 8, ntracers)
            if pl and pl >= ntracers:
                continue               # same as launching all at once
            if kern == FUSED:
                for c in TPL_CHOICES:
                    if c <= min(chunk, 4) and chunk % c == 0:
                        out.append((kern, pl, 0, c))
                continue
            for e, c in itertools.product(TPL_CHOICES, TPL_CHOICES):
                if e <= chunk and c <= chunk and chunk % e == 0 and chunk % c == 0:
                    out.append((kern, pl, e, c))
    return sorted(set(out))


def apply(tr, shape: Shape) -> None:
    """Put handle `tr` on `shape`."""
    kern, pl, e, c = shape
    tr.set_kernel(kern)
    tr.set_lanes(pl, e, c)


def device_arch() -> str:
    import torch
    p = torch.cuda.get_device_properties(torch.cuda.current_device())
    return getattr(p, "gcnArchName", p.name)


def cache_key(size, worksize, ntracers: int, precision: int, optimize: bool,
              device: str = "") -> str:
    return json.dumps([list(size), list(worksize), ntracers, precision, bool(optimize),
                       device or device_arch()])


def autotune(tr, steps: int = 3, shapes: Optional[List[Shape]] = None,
             cache_path: Optional[str] = None, key: Optional[str] = None,
             kernels: Iterable[int] = (LANES, FUSED)) -> Dict:
    """Time the candidate shapes on transport handle `tr` and set the best.
    Shapes the handle refuses (the fused walk with mixed precision or
    implicit horizontal advection) are skipped.  Returns {"shape": best,
    "ms": {shape: ms per step}}."""
    import torch
    from . import _lib as L

    cache: Dict[str, List[int]] = {}
    if cache_path and key and os.path.exists(cache_path):
        with open(cache_path) as fh:
            cache = json.load(fh)
        if key in cache:
            best = tuple(cache[key])
            apply(tr, best)
            return {"shape": best, "ms": {}, "cached": True}
    shapes = shapes or candidates(tr.ntracers, kernels=kernels)
    stream = torch.cuda.ExternalStream(tr.stream_handle())
    # with the automatic column order, a hydrodynamic state that has run
    # this many steps walks by depth: every candidate is then timed as the
    # handle runs it from now on
    tr.run(L.MGPU_SLOT_MIN_STEPS, use_graph=False)
    timings: Dict[Shape, float] = {}
    for shp in shapes:
        try:
            apply(tr, shp)
        except L.MgpuError as e:
            if e.code in (L.MGPU_ERR_UNSUPPORTED, L.MGPU_ERR_STATE):
                continue
            raise
        tr.run(1, use_graph=False)
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        tr.run(steps, use_graph=False)
        e1.record(stream)
        e1.synchronize()
        timings[shp] = e0.elapsed_time(e1) / steps
    best = min(timings, key=timings.get)
    apply(tr, best)
    if cache_path and key:
        cache[key] = list(best)
        with open(cache_path, "w") as fh:
            json.dump(cache, fh, indent=1)
    return {"shape": best, "ms": {str(k): v for k, v in timings.items()}, "cached": False}
