"""Launch-shape autotuner for the batched transport.

The lane variant's kernels take three shape parameters that change speed but
not results: tracers per launch, tracers per lane in the explicit kernel and
tracers per lane in the column kernel.  `autotune` times every candidate on
the caller's own handle (its hydrodynamic state and tracers), with HIP
events on the handle's stream, and leaves the handle on the fastest shape.
Each candidate runs `steps` timed steps after one untimed step; the tracer
state advances while tuning, so tune before the run that matters (or reload
the tracers afterwards).  A JSON cache keyed by the problem's shape skips the
search on later calls.
"""
from __future__ import annotations

import itertools
import json
import os
from typing import Dict, Iterable, List, Optional, Tuple

Shape = Tuple[int, int, int]   # (tracers_per_launch, explicit tpl, column tpl)

TPL_CHOICES = (1, 2, 4, 8)


def candidates(ntracers: int, per_launch: Iterable[int] = (0, 4)) -> List[Shape]:
    """All valid shapes for `ntracers` tracers: tracers per launch from
    `per_launch` (0 = all, at most 8), tracers per lane dividing the launch's
    chunk."""
    out: List[Shape] = []
    for pl in per_launch:
        chunk = min(pl or 8, ntracers)
        if pl and pl >= ntracers:
            continue               # same as launching all at once
        for e, c in itertools.product(TPL_CHOICES, TPL_CHOICES):
            if e <= chunk and c <= chunk and chunk % e == 0 and chunk % c == 0:
                out.append((pl, e, c))
    return sorted(set(out))


def cache_key(size, worksize, ntracers: int, precision: int, optimize: bool,
              device: str = "") -> str:
    return json.dumps([list(size), list(worksize), ntracers, precision, bool(optimize), device])


def autotune(tr, steps: int = 3, shapes: Optional[List[Shape]] = None,
             cache_path: Optional[str] = None, key: Optional[str] = None) -> Dict:
    """Time the candidate shapes on transport handle `tr` and set the best.
    Returns {"shape": best, "ms": {shape: ms per step}}."""
    import torch

    cache: Dict[str, List[int]] = {}
    if cache_path and key and os.path.exists(cache_path):
        with open(cache_path) as fh:
            cache = json.load(fh)
        if key in cache:
            best = tuple(cache[key])
            tr.set_lanes(*best)
            return {"shape": best, "ms": {}, "cached": True}
    shapes = shapes or candidates(tr.ntracers)
    stream = torch.cuda.ExternalStream(tr.stream_handle())
    timings: Dict[Shape, float] = {}
    for shp in shapes:
        tr.set_lanes(*shp)
        tr.run(1, use_graph=False)
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        tr.run(steps, use_graph=False)
        e1.record(stream)
        e1.synchronize()
        timings[shp] = e0.elapsed_time(e1) / steps
    best = min(timings, key=timings.get)
    tr.set_lanes(*best)
    if cache_path and key:
        cache[key] = list(best)
        with open(cache_path, "w") as fh:
            json.dump(cache, fh, indent=1)
    return {"shape": best, "ms": {str(k): v for k, v in timings.items()}, "cached": False}
