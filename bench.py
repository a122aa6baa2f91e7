"""Benchmark of the MI355X tracer-transport step.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--config c3]
                    [--precision fp64|fast] [--no-graph] [--no-cpu]

Workload (default): 8 tracers on a 1024 x 1024 x 40 Cartesian grid with ~30 %
land and per-column KFloor uniform in 5..40, TVD-SuperBee advection in all
directions, implicit vertical diffusion/advection, dt = 60 s -- the batched
multi-tracer step the headline metric is quoted on.  Synthetic, seeded inputs
(mohid_amd/fixtures.py); all inputs are resident in HBM before timing.

One step = one advection-diffusion pass of every tracer (face conductances
once per step, then one transport launch per tracer), as the reference's
WaterProperties loop does per time step.  The timed region replays K such
steps as one captured hipGraph.

Prints ONE JSON line.  `value` is tracer cell-updates/s over water cells
(T x N_wet x K / wall).  `roofline` is for the per-tracer transport kernel:
algorithmic bytes per step = (16 + 64/T) B x T x N_wet (FP64 byte model: one
read and one write of each tracer plus eight shared FP64 fields amortised over
T tracers) divided by the mean duration per step of the transport kernels
(explicit + vertical, all tracers) from HIP events recorded on the library's
stream around each launch.  `cpu_baseline` times the CPU oracle
(a restatement of the reference algorithm, OpenMP over the host cores) on a
bounded sample of the same workload.
"""
from __future__ import annotations

import argparse
import json
import os
import re
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, published


def _log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def _copy_ceiling():
    """A measured HBM ceiling beside the published peak: a 2 GiB device-to-
    device copy, HIP events around 5 copies after 2 warm-ups, counting the
    bytes read and written."""
    import torch
    n = 256 * 1024 * 1024                      # doubles: 2 GiB per buffer
    a = torch.ones(n, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    for _ in range(2):
        b.copy_(a)
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(5):
        b.copy_(a)
    e1.record()
    e1.synchronize()
    ms = e0.elapsed_time(e1) / 5
    del a, b
    torch.cuda.empty_cache()
    return {"GBps": 2 * n * 8 / (ms * 1e-3) / 1e9, "method": "2 GiB device copy, read + write"}


def _traffic_from_profiles(kernel_key: str, workload: str):
    """HBM bytes per step of the transport kernels measured with rocprofv3
    --pmc (FETCH_SIZE, WRITE_SIZE), as recorded in profiles/*.md (a JSON block
    per workload)."""
    prof = os.path.join(ROOT, "profiles")
    if not os.path.isdir(prof):
        return None
    for fn in sorted(os.listdir(prof)):
        if not fn.endswith(".md"):
            continue
        text = open(os.path.join(prof, fn)).read()
        for m in re.finditer(r"```json\n(.*?)\n```", text, re.S):
            try:
                rec = json.loads(m.group(1))
            except ValueError:
                continue
            if rec.get("kernel") == kernel_key and rec.get("workload") == workload:
                return rec.get("traffic_bytes_per_launch")
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--config", default="c3")
    ap.add_argument("--precision", default="fp64", choices=("fp64", "fast", "mixed"))
    ap.add_argument("--refinements", type=int, default=2,
                    help="mixed precision: residual-and-correction passes per step")
    ap.add_argument("--no-mixed", action="store_true",
                    help="skip the mixed-precision measurement beside the headline")
    ap.add_argument("--no-graph", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-window", type=int, default=384)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--kernel", default="auto", choices=("auto", "lanes", "split", "fused"),
                    help="auto: the autotuner picks the lane pair or the fused walk")
    ap.add_argument("--tracers-per-launch", type=int, default=0)
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--implicit", default="none", choices=("none", "x", "y"),
                    help="implicit horizontal advection (THOMAS_3D line solves), first "
                         "direction; it alternates every step")
    ap.add_argument("--suite", action="store_true",
                    help="instead of the headline line, run the comparison suite: plain FP64 "
                         "drop-in vs batched on 512x512x50, FAST vs FP64 per-tracer error and "
                         "mass drift on the stiff 512x512x50 case, the autotuned "
                         "2048x2048x60 run; one JSON line per measurement")
    ap.add_argument("--boundaries", action="store_true",
                    help="instead of the headline line, time the drop-in on 1024x1024x40 "
                         "(one tracer, device arrays) without and with each offloaded "
                         "boundary condition, then with each other optional branch (cell "
                         "fluxes, NoDifFlux, NoAdvFlux, SmallDepths, nudging, discharges); "
                         "two JSON lines")
    ap.add_argument("--dropin-steps", type=int, default=0,
                    help="instead of the headline line, run this many per-call drop-in steps "
                         "on 1024x1024x40 (8 tracers, device arrays, FirstProperty on tracer "
                         "0) and exit: the workload of the HIP API trace that compares "
                         "allocations at two step counts")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain device copy of 512 MiB three times (the "
                         "reference workload for scaling FETCH_SIZE/WRITE_SIZE counters)")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        torch.cuda.set_device(0)
        n = 64 * 1024 * 1024
        a = torch.ones(n, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes": n * 8, "copies": 3}), flush=True)
        return

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this build measures one GPU only (the transport step's "
              "single-GPU kernel); run with --gpus 1.", file=sys.stderr)
        sys.exit(2)
    if args.boundaries:
        boundary_costs(args)
        return
    if args.dropin_steps > 0:
        dropin_steps(args.dropin_steps)
        return
    if args.suite:
        suite(args)
        return

    import numpy as np
    import torch

    from mohid_amd import _lib
    from mohid_amd import fixtures as fx
    from mohid_amd.advection_diffusion import Transport

    if not torch.cuda.is_available():
        print("bench.py: no GPU visible", file=sys.stderr)
        sys.exit(2)
    torch.cuda.set_device(0)
    _lib.lib()

    cfg = fx.config(args.config)
    t0 = time.time()
    case = fx.make_case(cfg, device="cuda")
    T = len(case.tracers)
    nwet = case.n_wet
    _log(f"case {cfg.name}: {cfg.ni}x{cfg.nj}x{cfg.nk}, {T} tracers, {nwet} water cells "
         f"({time.time() - t0:.1f}s to build)")

    prec = {"fp64": _lib.MGPU_PREC_FP64, "fast": _lib.MGPU_PREC_FAST,
            "mixed": _lib.MGPU_PREC_MIXED}[args.precision]
    kernel = {"auto": _lib.MGPU_KERNEL_LANES, "lanes": _lib.MGPU_KERNEL_LANES,
              "split": _lib.MGPU_KERNEL_SPLIT, "fused": _lib.MGPU_KERNEL_FUSED}[args.kernel]
    params = _with_implicit(case.params, args.implicit)
    tr = Transport(case.size, case.worksize, case.geom, T, precision=prec, kernel=kernel,
                   tracers_per_launch=args.tracers_per_launch, refinements=args.refinements)
    tr.set_hydro(case.hydro, params)
    for t, p in enumerate(case.tracers):
        tr.set_tracer(t, p)
    use_graph = not args.no_graph

    # ---- launch shape from the autotuner, before any timing; tuning advances
    # the state, so the tracers are reloaded afterwards
    from mohid_amd import autotune as AT
    tuned = None
    variant = {"auto": "lanes"}.get(args.kernel, args.kernel)
    if args.kernel in ("auto", "lanes", "fused") and not args.no_autotune:
        t0 = time.time()
        kernels = {"auto": (AT.LANES, AT.FUSED), "lanes": (AT.LANES,),
                   "fused": (AT.FUSED,)}[args.kernel]
        tuned = AT.autotune(tr, steps=2, kernels=kernels)
        for t, p in enumerate(case.tracers):
            tr.set_tracer(t, p)
        variant = "fused" if tuned["shape"][0] == AT.FUSED else "lanes"
        _log(f"autotuned launch shape (kernel, tracers/launch, explicit tracers/lane, column "
             f"tracers/lane) = {tuned['shape']} ({variant}) in {time.time() - t0:.1f}s")

    # ---- warm-up
    if args.warmup > 0:
        tr.run(args.warmup, use_graph=use_graph)
    torch.cuda.synchronize()

    # ---- timed region: K steps (one captured graph when use_graph)
    stream = torch.cuda.ExternalStream(tr.stream_handle())
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    if use_graph:
        tr.run(args.steps, use_graph=True)   # capture + instantiate outside the timed region
        torch.cuda.synchronize()
    torch.cuda.synchronize()
    w0 = time.perf_counter()
    ev0.record(stream)
    tr.run(args.steps, use_graph=use_graph)
    ev1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - w0
    ev_ms = ev0.elapsed_time(ev1)
    ms_per_step = 1e3 * wall / args.steps
    value = T * nwet * args.steps / wall
    _log(f"timed {args.steps} steps: {wall * 1e3:.2f} ms wall, {ev_ms:.2f} ms events, "
         f"{value:.4e} cell-updates/s")

    # ---- per-kernel durations (HIP events on the library's stream around the
    # kernels of every step, eager steps)
    ksteps = max(2, min(args.steps, 10))
    tr.timing(True, ksteps)
    tr.run(ksteps, use_graph=False)
    prep_ms, h_ms, v_ms, nsteps_t = tr.kernel_times()
    tr.timing(False)
    trans_ms = h_ms + v_ms
    # FP64 byte model: (16 + 64/T) B per tracer-cell update, T*N_wet updates
    per_unit = 16.0 + 64.0 / T
    alg_bytes = per_unit * nwet * T
    achieved = alg_bytes / (trans_ms * 1e-3) / 1e9
    kernel_key = f"transport_{variant}_{args.precision}"
    if args.implicit != "none":
        kernel_key += "_implicit"
    traffic = _traffic_from_profiles(kernel_key, cfg.name)
    _log(f"per step: prep {prep_ms:.3f} ms, explicit {h_ms:.3f} ms, vertical {v_ms:.3f} ms "
         f"({nsteps_t} steps); transport {achieved:.1f} GB/s algorithmic")

    out = {
        "metric": "tracer cell-updates/s + %HBM peak, 1024\u00b2\u00d740\u00d78 tracers, "
                  "mixed vs FP64 GPU",
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision != "mixed" else "f64 rows + f32 sweeps",
        "data": "synthetic (seeded fixtures, mohid_amd/fixtures.py)",
        "config": {
            "workload": f"{cfg.name}: {cfg.ni}x{cfg.nj}x{cfg.nk}, {T} tracers, "
                        f"{cfg.vertical}, land {cfg.land_frac:.2f}, KFloor {cfg.kfloor}, "
                        f"TVD-SuperBee H+V, dt {cfg.dt:g}s",
            "n_wet_cells": nwet,
            "precision": args.precision,
            "graph": use_graph,
            "kernel": variant,
            "launch_shape": list(tuned["shape"]) if tuned else None,
            "optimize_path": bool(case.params["Optimize"]),
            "implicit_horizontal": args.implicit,
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "kernel": ("line_rows_kernel + line_thomas_kernel + column_lane_kernel"
                       if args.implicit != "none"
                       else "walk_lane_kernel" if variant == "fused"
                       else ("explicit_lane_kernel + " +
                             ("column_mixed_kernel" if args.precision == "mixed"
                              else "column_pipe_kernel" if case.params["Optimize"]
                              else "column_lane_kernel")) if variant == "lanes"
                       else "explicit_kernel + vertical_kernel") + " (all tracers of one step)",
            "kernel_ms": trans_ms,
            "explicit_kernel_ms": h_ms,
            "vertical_kernel_ms": v_ms,
            "alg_bytes_per_launch": alg_bytes,
            "alg_bytes_model": f"(16 + 64/T) B x T x N_wet, T={T}",
            "prep_kernel_ms": prep_ms,
            "copy_ceiling": _copy_ceiling(),
        },
    }

    # ---- mixed precision beside the FP64 headline: the same K steps from the
    # same initial tracers, timed the same way, and its error against FP64
    if args.precision == "fp64" and args.implicit == "none" and not args.no_mixed:
        out["mixed"] = _mixed_vs_fp64(case, tr, args, value)
    # ---- CPU baseline: the oracle on a bounded sample of the same workload
    # (the same configuration on a smaller horizontal window), timed on the
    # host cores; the GPU path runs the same sample for the parity figure
    if not args.no_cpu:
        from oracle import oracle as ora
        # every core this process may run on, capped by the host's
        # OMP_NUM_THREADS (its CPU share) when set
        nthreads = ora.host_threads()
        scfg = fx.config(cfg.name, ni=min(cfg.ni, args.cpu_window),
                         nj=min(cfg.nj, args.cpu_window))
        scase = fx.make_case(scfg)
        sT, swet = len(scase.tracers), scase.n_wet
        inst = ora.Instance(scase, nthreads=nthreads)
        props = [np.array(t, order="F", copy=True) for t in scase.tracers]
        dt = scase.params["DTProp"]

        def ostep(s):
            for t, p in enumerate(props):
                rc = inst.step(p, first_property=(t == 0), now=(s + 1) * dt,
                               **_implicit_dirs(args.implicit, s))
                assert rc == 0, rc

        ostep(0)                      # untimed first step
        c0 = time.perf_counter()
        nsteps = 0
        while nsteps < 2 or (time.perf_counter() - c0 < args.cpu_seconds and nsteps < 50):
            ostep(nsteps + 1)
            nsteps += 1
        cwall = time.perf_counter() - c0
        inst.close()
        cval = sT * swet * nsteps / cwall
        tr2 = Transport(scase.size, scase.worksize, scase.geom, sT, precision=prec, kernel=kernel,
                        tracers_per_launch=args.tracers_per_launch)
        tr2.set_hydro(scase.hydro, _with_implicit(scase.params, args.implicit))
        for t, p in enumerate(scase.tracers):
            tr2.set_tracer(t, p)
        tr2.run(1 + nsteps, use_graph=False)
        wp = scase.hydro["WaterPoints3D"] == 1
        worst = 0.0
        for t in range(sT):
            got = tr2.get_tracer(t)
            err = np.abs(got[wp] - props[t][wp]) / np.maximum(np.abs(props[t][wp]), 1e-300)
            worst = max(worst, float(err.max()))
        tr2.close()
        out["cpu_baseline"] = {
            "value": cval, "unit": "cell-updates/s", "cores": nthreads, "kind": "port",
            "host": _host_cpu(),
            "sample": f"{nsteps} steps x {sT} tracers of {scfg.name} on a {scfg.ni}x{scfg.nj}"
                      f"x{scfg.nk} window ({swet} water cells) after 1 untimed step; "
                      f"oracle/advdiff_oracle.c with OpenMP",
            "wall_s": cwall,
        }
        out["parity_vs_oracle"] = {"sample": f"{scfg.ni}x{scfg.nj}x{scfg.nk}",
                                   "steps": 1 + nsteps, "max_rel_err": worst}
        _log(f"cpu oracle {cval:.4e} cell-updates/s on {nthreads} threads ({nsteps} steps, "
             f"{cwall:.1f}s); gpu vs oracle max rel err {worst:.3e}")
    tr.close()
    print(json.dumps(out), flush=True)


def _mixed_vs_fp64(case, tr64, args, value64):
    """MGPU_PREC_MIXED with 1 and 2 correction passes on the headline case:
    updates/s over K graph-captured steps (autotuned), and the max relative
    error of every tracer on water cells against the FP64 handle's result
    after the same K steps from the same initial tracers."""
    import torch
    from mohid_amd import _lib
    from mohid_amd import autotune as AT
    from mohid_amd.advection_diffusion import Transport
    T = len(case.tracers)
    K = args.steps
    wp = case.hydro["WaterPoints3D"] == 1
    zero = torch.zeros_like(case.hydro["VolumeZ"])
    for t, p in enumerate(case.tracers):
        tr64.set_tracer(t, p)
    tr64.run(K, use_graph=True)
    ref = []
    for t in range(T):
        o = torch.empty_like(zero)
        tr64.get_tracer(t, o)
        ref.append(o)
    res = {"steps_compared": K, "passes": {}}
    for passes in (1, 2):
        tm = Transport(case.size, case.worksize, case.geom, T, precision=_lib.MGPU_PREC_MIXED,
                       refinements=passes)
        tm.set_hydro(case.hydro, case.params)
        for t, p in enumerate(case.tracers):
            tm.set_tracer(t, p)
        shape = AT.autotune(tm, steps=2, kernels=(AT.LANES,))["shape"]
        for t, p in enumerate(case.tracers):
            tm.set_tracer(t, p)
        tm.run(K, use_graph=True)              # capture, and the K steps compared
        tm.status()
        worst = 0.0
        o = torch.empty_like(zero)
        for t in range(T):
            tm.get_tracer(t, o)
            rel = torch.where(wp, torch.abs(o - ref[t]) / torch.clamp(torch.abs(ref[t]),
                                                                      min=1e-300), zero)
            worst = max(worst, float(torch.max(rel).item()))
        torch.cuda.synchronize()
        w0 = time.perf_counter()
        tm.run(K, use_graph=True)
        torch.cuda.synchronize()
        wall = time.perf_counter() - w0
        tm.timing(True, min(K, 10))
        tm.run(min(K, 10), use_graph=False)
        pm, hm, vm, _ = tm.kernel_times()
        tm.close()
        v = T * case.n_wet * K / wall
        res["passes"][str(passes)] = {
            "value": v, "ms_per_step": 1e3 * wall / K, "vs_fp64": v / value64,
            "max_rel_err_vs_fp64": worst, "launch_shape": list(shape),
            "prep_ms": pm, "explicit_ms": hm, "column_ms": vm}
        _log(f"mixed, {passes} pass(es): {v:.4e} cell-updates/s ({v / value64:.3f} x FP64), "
             f"column kernel {vm:.3f} ms, max rel err vs FP64 after {K} steps {worst:.2e}")
    del ref
    torch.cuda.empty_cache()
    return res


def _host_cpu():
    """What the host offers the CPU baseline: its CPU model, nproc, the cores
    this process may run on and the OMP_NUM_THREADS it sets."""
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"cpu_model": model, "nproc": os.cpu_count(),
            "affinity_cores": len(os.sched_getaffinity(0)),
            "OMP_NUM_THREADS": os.environ.get("OMP_NUM_THREADS")}


def _implicit_dirs(first, step):
    """ImpExp_AdvXX/YY of a step when the implicit direction starts at
    `first` ("x" or "y") and alternates every step (WP:14681-14712)."""
    if first == "none":
        return {}
    xx = (step % 2 == 0) == (first == "x")
    return {"ImpExp_AdvXX": 1.0 if xx else 0.0, "ImpExp_AdvYY": 0.0 if xx else 1.0}


def _with_implicit(params, first):
    return dict(params, **_implicit_dirs(first, 0))


def _timed(fn, stream, k):
    """Wall seconds of k calls of fn(), with the device idle before and after."""
    import torch
    torch.cuda.synchronize()
    w0 = time.perf_counter()
    for _ in range(k):
        fn()
    torch.cuda.synchronize()
    return time.perf_counter() - w0


def _batched_rate(case, precision, steps, chunk=None, tune=True, implicit="none"):
    """Batched transport of all the case's tracers: autotuned shape, steps
    replayed as captured graphs of `chunk` steps; returns (updates/s, shape)."""
    import torch
    from mohid_amd.advection_diffusion import Transport
    from mohid_amd.autotune import autotune
    T = len(case.tracers)
    tr = Transport(case.size, case.worksize, case.geom, T, precision=precision)
    tr.set_hydro(case.hydro, _with_implicit(case.params, implicit))
    for t, p in enumerate(case.tracers):
        tr.set_tracer(t, p)
    shape = autotune(tr, steps=2)["shape"] if tune else None
    for t, p in enumerate(case.tracers):
        tr.set_tracer(t, p)
    chunk = chunk or steps
    tr.run(chunk, use_graph=True)                   # capture + warm-up
    wall = _timed(lambda: tr.run(chunk, use_graph=True), None, steps // chunk)
    tr.close()
    torch.cuda.empty_cache()
    return T * case.n_wet * (steps // chunk) * chunk / wall, shape


def dropin_steps(steps):
    """`steps` transport steps through the per-call drop-in, as the Fortran
    caller makes them: one call per tracer, FirstProperty on the first."""
    import torch
    from mohid_amd import _lib
    from mohid_amd import fixtures as fx
    from mohid_amd.advection_diffusion import AdvectionDiffusion

    torch.cuda.set_device(0)
    _lib.lib()
    case = fx.make_case(fx.config("c3"), device="cuda")
    ad = AdvectionDiffusion(case.size, case.worksize, case.geom)
    props = [t.clone() for t in case.tracers]
    for s in range(steps):
        for t, p in enumerate(props):
            ad(p, case.hydro, case.params, first_property=(t == 0))
        torch.cuda.synchronize()
        _log(f"drop-in step {s + 1} of {steps}: {len(props)} calls")
    ad.kill()
    print(json.dumps({"dropin_steps": steps, "calls": steps * len(props)}), flush=True)


def boundary_costs(args):
    """Per-call time of the drop-in with each boundary condition on the
    headline grid: the water points on the domain's edges are boundary
    points (as HorizontalMap makes them), a reference field for the open
    boundary, DecayTime 1 h."""
    import torch
    from mohid_amd import _lib
    from mohid_amd import constants as C
    from mohid_amd import fixtures as fx
    from mohid_amd.advection_diffusion import AdvectionDiffusion

    torch.cuda.set_device(0)
    _lib.lib()
    case = fx.make_case(fx.config("c3"), device="cuda")
    ws = case.worksize
    # device fields are (k, j, i) tensors: Fortran (i, j, k) memory order
    wp = case.hydro["WaterPoints3D"][ws[5]] == 1
    edge = torch.zeros_like(wp)
    edge[ws[2]:ws[3] + 1, ws[0]] = True
    edge[ws[2]:ws[3] + 1, ws[1]] = True
    edge[ws[2], ws[0]:ws[1] + 1] = True
    edge[ws[3], ws[0]:ws[1] + 1] = True
    bp = (edge & wp).to(torch.int32).contiguous()
    hy = dict(case.hydro, BoundaryPoints2D=bp)
    ref = torch.full_like(case.tracers[0], 2.0)
    ad = AdvectionDiffusion(case.size, case.worksize, case.geom)
    prop = case.tracers[0].clone()
    out = {"measure": "drop-in per call with boundary conditions, 1024x1024x40, 1 tracer, "
                      "device arrays (median of 10 calls)", "boundary_columns": int(bp.sum()),
           "unit": "ms_per_call"}
    k = 10

    def median_ms(fn):
        # per-call wall times (each call ends in a stream synchronisation);
        # the median keeps a host-side stall from one call out of the figure
        ts = []
        for _ in range(k):
            t0 = time.perf_counter()
            fn()
            ts.append(time.perf_counter() - t0)
        return 1e3 * sorted(ts)[k // 2]

    for name, extra in (("none", {}),
                        ("MassConservation", dict(ReferenceProp=ref,
                                                  BoundaryCondition=C.MASS_CONSERVATION)),
                        ("ImposedValue", dict(ReferenceProp=ref,
                                              BoundaryCondition=C.IMPOSED_VALUE)),
                        ("NullGradient", dict(ReferenceProp=ref,
                                              BoundaryCondition=C.NULL_GRADIENT)),
                        ("Cyclic", dict(ReferenceProp=ref, BoundaryCondition=C.CYCLIC))):
        prm = dict(case.params, DecayTime=3600.0, **extra)
        for _ in range(2):
            ad(prop, hy, prm)
        out[name] = median_ms(lambda: ad(prop, hy, prm))
        print(f"boundaries: {name} {out[name]:.3f} ms per call", file=sys.stderr, flush=True)
    print(json.dumps(out), flush=True)

    # the other optional branches, on the non-optimised path they belong to
    g = torch.Generator(device="cuda").manual_seed(5)
    flag = lambda: (torch.rand(case.tracers[0].shape, generator=g, device="cuda") < 0.3
                    ).to(torch.int32).contiguous()
    hf = dict(case.hydro, NoFluxU=flag(), NoFluxV=flag(), NoFluxW=flag())
    hs = dict(case.hydro, SmallDepths=(torch.rand(wp.shape, generator=g, device="cuda") < 0.4
                                       ).to(torch.int32).contiguous())
    jj, ii = torch.nonzero(wp, as_tuple=True)
    pick = torch.randperm(len(ii), generator=g, device="cuda")[:8].cpu()
    nd = len(pick)
    disch = dict(flow=[1.5] * nd, conc=[5.0] * nd, concMF=[1.0] * nd,
                 I=[int(v) for v in ii.cpu()[pick]], J=[int(v) for v in jj.cpu()[pick]],
                 K=[int(ws[5])] * nd, Kmin=[-9999999] * nd, Kmax=[-9999999] * nd,
                 vert=[1] * nd, ignore=[0] * nd, ncells=[1] * nd, bypass=[0] * nd)
    out2 = {"measure": "drop-in per call with the other optional branches, 1024x1024x40, "
                       "1 tracer, device arrays, Optimize off (median of 10 calls)",
            "unit": "ms_per_call", "discharges": nd}
    base = dict(case.params, Optimize=False)
    for name, h, extra in (("none", case.hydro, {}), ("CellFluxes", case.hydro,
                                                      dict(CellFluxes=True)),
                           ("NoDifFlux", hf, dict(NoDifFlux=True)),
                           ("NoAdvFlux", hf, dict(NoAdvFlux=True)),
                           ("SmallDepths", hs, {}),
                           ("AdvectionNudging", case.hydro,
                            dict(AdvectionNudging=True, AdvectionNudgingCells=10)),
                           ("Discharges", case.hydro, {})):
        prm = dict(base, **extra)
        if name == "Discharges":
            def call():
                ad.set_discharges(disch)
                ad(prop, h, prm)
                ad.unset_discharges()
        else:
            def call():
                ad(prop, h, prm)
        for _ in range(2):
            call()
        out2[name] = median_ms(call)
        print(f"options: {name} {out2[name]:.3f} ms per call", file=sys.stderr, flush=True)
    ad.kill()
    print(json.dumps(out2), flush=True)


def suite(args):
    """Comparison measurements beside the headline (one JSON line each)."""
    import numpy as np
    import torch
    from mohid_amd import _lib
    from mohid_amd import fixtures as fx
    from mohid_amd.advection_diffusion import AdvectionDiffusion, Transport

    torch.cuda.set_device(0)
    _lib.lib()

    def emit(rec):
        print(json.dumps(rec), flush=True)

    # 1. plain FP64 drop-in vs batched, one tracer, 512 x 512 x 50 sigma, TVD
    case = fx.make_case(fx.config("c2"), device="cuda")
    ad = AdvectionDiffusion(case.size, case.worksize, case.geom)
    prop = case.tracers[0].clone()
    k = 20
    for _ in range(2):
        ad(prop, case.hydro, case.params)
    wall = _timed(lambda: ad(prop, case.hydro, case.params), None, k)
    ad.kill()
    drop = case.n_wet * k / wall
    emit({"measure": "plain FP64 drop-in (per-call AdvectionDiffusion, device arrays)",
          "workload": "512x512x50 sigma, 1 tracer, TVD-SuperBee, dt 30 s",
          "value": drop, "unit": "cell-updates/s", "ms_per_call": 1e3 * wall / k})
    b64, shp = _batched_rate(case, _lib.MGPU_PREC_FP64, 20)
    emit({"measure": "batched FP64, same workload", "value": b64, "unit": "cell-updates/s",
          "launch_shape": shp, "vs_plain_dropin": b64 / drop})
    del case
    torch.cuda.empty_cache()

    # 2. the headline shape: plain FP64 drop-in (one AdvectionDiffusion call per
    # tracer and step, FirstProperty on the first: WP:14776-14822) vs batched
    # FP64 and FAST, 8 tracers, 1024 x 1024 x 40
    case = fx.make_case(fx.config("c3"), device="cuda")
    ad = AdvectionDiffusion(case.size, case.worksize, case.geom)
    props = [t.clone() for t in case.tracers]

    def dropin_step():
        for t, p in enumerate(props):
            ad(p, case.hydro, case.params, first_property=(t == 0))

    dropin_step()
    k = 5
    # one step at a time: an occasional host-side stall of tens of ms (seen
    # outside any kernel on the GPU timeline) would dominate a mean, so the
    # rate is the median step's
    each = [_timed(dropin_step, None, 1) for _ in range(k)]
    wall = sorted(each)[k // 2]
    ad.kill()
    del props
    drop3 = len(case.tracers) * case.n_wet / wall
    emit({"measure": "plain FP64 drop-in, 1024x1024x40 x 8 tracers (8 calls per step)",
          "value": drop3, "unit": "cell-updates/s", "ms_per_step_median": 1e3 * wall,
          "ms_each_step": [1e3 * w for w in each]})
    r64, s64 = _batched_rate(case, _lib.MGPU_PREC_FP64, 20)
    rf, sf = _batched_rate(case, _lib.MGPU_PREC_FAST, 20)
    r64_headline = r64
    emit({"measure": "batched FP64 vs FAST, 1024x1024x40 x 8 tracers",
          "fp64": r64, "fast": rf, "unit": "cell-updates/s", "fast_over_fp64": rf / r64,
          "fp64_over_plain_dropin": r64 / drop3, "fast_over_plain_dropin": rf / drop3,
          "launch_shapes": {"fp64": s64, "fast": sf}})
    del case
    torch.cuda.empty_cache()

    # 3. FAST vs FP64 error per tracer and mass drift, stiff 512 x 512 x 50
    # (8 tracers, thin surface layers, Kz up to 0.1, dt 120 s); the FP64 GPU
    # path is bit-identical to the oracle, so it stands in for it at full size
    case = fx.make_case(fx.config("c4"), device="cuda")
    steps = 50
    outs = {}
    for name, prec in (("fp64", _lib.MGPU_PREC_FP64), ("fast", _lib.MGPU_PREC_FAST)):
        tr = Transport(case.size, case.worksize, case.geom, len(case.tracers), precision=prec)
        tr.set_hydro(case.hydro, case.params)
        for t, p in enumerate(case.tracers):
            tr.set_tracer(t, p)
        tr.run(steps, use_graph=True)
        outs[name] = [tr.get_tracer(t, out=torch.empty_like(case.tracers[t]))
                      for t in range(len(case.tracers))]
        tr.close()
    wp = case.hydro["WaterPoints3D"] == 1
    V = case.hydro["VolumeZ"]
    per = []
    for t in range(len(case.tracers)):
        a, b, c0 = outs["fast"][t][wp], outs["fp64"][t][wp], case.tracers[t][wp]
        rel = (a - b).abs() / b.abs().clamp_min(1e-300)
        m0 = float((V[wp] * c0).sum())
        per.append({"tracer": case.cfg.tracers[t], "max_rel": float(rel.max()),
                    "rms_rel": float(rel.pow(2).mean().sqrt()),
                    "mass_drift_fp64": abs(float((V[wp] * b).sum()) - m0) / abs(m0),
                    "mass_drift_fast": abs(float((V[wp] * a).sum()) - m0) / abs(m0)})
    emit({"measure": "FAST vs FP64 after 50 steps, stiff 512x512x50 sigma (dz to 0.25 m, "
                     "Kz to 0.1, dt 120 s), 8 tracers", "per_tracer": per})
    del case, outs
    torch.cuda.empty_cache()

    # 4. implicit horizontal advection (THOMAS_3D line solves, direction
    # alternating every step) on the headline shape: batched FP64 and FAST,
    # per-kernel times, and the oracle on a bounded window of the same case
    case = fx.make_case(fx.config("c3"), device="cuda")
    T = len(case.tracers)
    i64, si = _batched_rate(case, _lib.MGPU_PREC_FP64, 20, implicit="x")
    ifa, sfa = _batched_rate(case, _lib.MGPU_PREC_FAST, 20, implicit="x")
    tr = Transport(case.size, case.worksize, case.geom, T)
    tr.set_hydro(case.hydro, _with_implicit(case.params, "x"))
    for t, p in enumerate(case.tracers):
        tr.set_tracer(t, p)
    tr.run(2, use_graph=False)
    tr.timing(True, 4)
    tr.run(4, use_graph=False)
    prep_ms, line_ms, col_ms, _ = tr.kernel_times()
    tr.timing(False)
    tr.close()
    del case
    torch.cuda.empty_cache()
    emit({"measure": "implicit horizontal advection (alternating x/y), 1024x1024x40 x 8 tracers",
          "fp64": i64, "fast": ifa, "unit": "cell-updates/s",
          "explicit_fp64_same_workload": r64_headline, "implicit_over_explicit": i64 / r64_headline,
          "line_kernel_ms": line_ms, "column_kernel_ms": col_ms, "prep_kernel_ms": prep_ms,
          "alg_GBps_fp64": i64 * (16 + 64 / T) / 1e9,
          "launch_shapes": {"fp64": si, "fast": sfa}})
    from oracle import oracle as ora
    scase = fx.make_case(fx.config("c3", ni=256, nj=256))
    inst = ora.Instance(scase, nthreads=ora.host_threads())
    props = [np.array(t, order="F", copy=True) for t in scase.tracers]
    dt = scase.params["DTProp"]
    c0 = time.perf_counter()
    ns = 0
    while ns < 2 or (time.perf_counter() - c0 < 10.0 and ns < 40):
        for t, p in enumerate(props):
            assert inst.step(p, first_property=(t == 0), now=(ns + 1) * dt,
                             **dict(scase.params, **_implicit_dirs("x", ns))) == 0
        ns += 1
    cw = time.perf_counter() - c0
    inst.close()
    emit({"measure": "oracle (CPU), implicit horizontal advection, 256x256x40 window of the "
                     "same case x 8 tracers", "value": len(scase.tracers) * scase.n_wet * ns / cw,
          "unit": "cell-updates/s", "threads": ora.host_threads(), "steps": ns})
    del scase, props

    # 5. large run: 2048 x 2048 x 60 sigma, 4 tracers, graphs of 50 steps
    case = fx.make_case(fx.config("c5"), device="cuda")
    r64, s64 = _batched_rate(case, _lib.MGPU_PREC_FP64, 100, chunk=50)
    rf, sf = _batched_rate(case, _lib.MGPU_PREC_FAST, 100, chunk=50)
    emit({"measure": "batched, 2048x2048x60 sigma x 4 tracers, 25 % land, 2 graphs of 50 steps",
          "n_wet_cells": case.n_wet, "fp64": r64, "fast": rf, "unit": "cell-updates/s",
          "alg_GBps_fp64": r64 * (16 + 64 / 4) / 1e9, "alg_GBps_fast": rf * (16 + 64 / 4) / 1e9,
          "launch_shapes": {"fp64": s64, "fast": sf}})


if __name__ == "__main__":
    main()
