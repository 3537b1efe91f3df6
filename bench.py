"""Benchmark: FP64 MUSCL-Hancock HLLC throughput at 8192^2 on one MI355X.

A step is one pass of the hot path over the grid: the device clock (dt from
the previous step's fused max-speed reduction), the x sweep and the y sweep
of the dimensionally split second-order scheme.  The workload is the
cylindrical blast (rho=1, p=1 inside r <= 0.4 of (1,1); rho=0.125, p=0.1
outside; [0,2]^2) on an 8192 x 8192 grid, HLLC, CFL 0.8, state resident in
HBM before the timed region.

value: cell updates per second over the timed steps (Gcell-updates/s).
roofline: the dominant kernel (the slower sweep) -- 64 algorithmic bytes per
  cell (read 32 B, write 32 B of the four FP64 planes) times nx*ny cells per
  launch, divided by its average duration from HIP events recorded on the
  launch stream over the timed region; peak 8 TB/s.  The whole-step fraction
  (128 B per cell update) is reported beside it.
cpu_baseline: the reference CPU solver compiled from its sources
  (oracle/_ref) -- or the oracle port when that build is absent -- on the
  same workload at 2048^2 for a bounded number of steps, with all host
  threads OpenMP gives it.
batched: 512 blast-wave instances (seed 0, sides log-uniform in [64,1024],
  95.3 M cells), 500 steps with per-instance dt, one persistent launch per
  step replayed from captured hipGraphs, schedule from the autotuner; its
  cell-update rate and the ratio to the single-grid value.
adjoint: the 512x256 shock-bubble adjoint case (400 HLLC steps, a
  checkpoint every 20): the backward pass (segment recompute + adjoint
  steps) per step over the device-resident forward step.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--size N]
       [--solver hllc] [--no-cpu] [--no-extras] [--leg batched|adjoint]
(--leg runs one of the two extra legs alone and prints its JSON object:
the profiling runs of those workloads use it.)
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12        # MI355X HBM3E, bytes/s
BYTES_PER_SWEEP = 64.0   # read + write of four FP64 planes per cell
METRIC = "Gcell-updates/s + % HBM roofline at 8192² FP64; batched/adjoint vs plain"


def cpu_baseline(solver: int, budget_s: float = 12.0):
    """Reference CPU solver (or the oracle port) on cyl_blast at 2048^2."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np
    import pyoracle

    from large_sys_eq_opt_amd import cases

    cfg = cases.get_n_case_config("cyl_blast", 8)   # 2048^2
    kind, params = cases.case_params("cyl_blast")
    orc = pyoracle.Oracle()
    U0 = orc.init(kind, params.array(), cfg.nx, cfg.ny, cfg.bounds)
    if pyoracle.reference_available():
        impl, kind_s = pyoracle.Reference(), "reference"
        run = lambda n: impl.run(U0, cfg.bounds, solver, cfg.cfl, max_steps=n)  # noqa: E731
    else:
        impl, kind_s = orc, "port"
        run = lambda n: impl.run(U0, cfg.dx, cfg.dy, solver, cfg.cfl, max_steps=n)  # noqa: E731
    probe = run(2)
    per_step = max(probe["seconds"] / 2.0, 1e-6)
    n = int(max(3, min(400, budget_s / per_step)))
    r = run(n)
    cells = cfg.nx * cfg.ny
    rate = cells * r["steps"] / r["seconds"]
    threads = orc.num_threads()
    return {"value": rate / 1e9, "unit": "Gcell-updates/s", "cores": threads, "kind": kind_s,
            "sample": f"cyl_blast 2048x2048 HLLC FP64, {r['steps']} steps, {r['seconds']:.2f} s, "
                      f"{threads} OpenMP threads ({'reference CPU sources' if kind_s == 'reference' else 'oracle port'})",
            "np": np.__version__}


def batched_leg(plain_rate: float, steps: int = 500):
    """512-instance ensemble, graph-replayed persistent launches."""
    from large_sys_eq_opt_amd import ensemble

    spec = ensemble.ensemble_spec(512, 0)
    b = ensemble.Batch(spec, solver=1, cfl=0.8)
    b.initialise()
    tune = ensemble.autotune(b)
    per_graph = 25
    b.capture(per_graph)
    b.clock()
    t0 = time.perf_counter()
    b.step(steps)
    _, _, st, _ = b.clock()
    wall = time.perf_counter() - t0
    info = b.info()
    b.close()
    rate = b.cells * steps / wall
    return {"value": rate / 1e9, "unit": "Gcell-updates/s", "instances": len(spec), "cells": b.cells,
            "steps": steps, "ms_per_step": wall * 1e3 / steps, "ratio_vs_plain": rate / plain_rate,
            "schedule": list(tune["schedule"]), "graph_steps": per_graph, "blocks": info["blocks"],
            "all_instances_stepped": bool((st == steps).all())}


def adjoint_leg():
    """Backward-pass cost per step over the device-resident forward step on
    the 512x256 shock-bubble adjoint case."""
    import numpy as np

    from large_sys_eq_opt_amd import adjoint as A
    from large_sys_eq_opt_amd.solver import Grid2DGPU, Stepper, synchronize

    case = A.ShockBubbleAdjoint()
    cfg, sb, N = case.cfg, case.params, case.nsteps
    rho_b = np.full((cfg.ny, cfg.nx), sb.bubble_rho)
    U, _, _ = A.smooth_bubble_state(cfg, rho_b, sb.bubble_r)
    g = Grid2DGPU.from_config(cfg)
    best = {}
    for _ in range(3):
        g.upload_from_aos(U)
        st = Stepper(g, 1, cfg.cfl, 0.0, case.bc)
        synchronize()
        t0 = time.perf_counter()
        st.advance(N)
        st.finish()
        synchronize()
        best["fwd"] = min(best.get("fwd", 1e9), time.perf_counter() - t0)
        st.close()
        case.U0.upload_from_aos(U)
        t0 = time.perf_counter()
        case.adj.forward(case.U0, N, cfg.cfl, None, case.UN)
        best["ckpt"] = min(best.get("ckpt", 1e9), time.perf_counter() - t0)
        t0 = time.perf_counter()
        case.adj.backward(case.lamN, case.lam0)
        best["bwd"] = min(best.get("bwd", 1e9), time.perf_counter() - t0)
    g.close()
    case.close()
    f, c, b = best["fwd"] / N, best["ckpt"] / N, best["bwd"] / N
    return {"workload": f"shock_bubble_adj {cfg.nx}x{cfg.ny} HLLC, {N} steps, checkpoint every 20",
            "forward_us_per_step": f * 1e6, "checkpointing_forward_us_per_step": c * 1e6,
            "backward_us_per_step": b * 1e6, "ratio_vs_forward": b / f,
            "ratio_vs_checkpointing_forward": b / c}


def copy_peak(nbytes: int = 4 << 30, reps: int = 10):
    """Measured device-to-device copy bandwidth (read + write bytes / s)."""
    import ctypes as C

    from large_sys_eq_opt_amd._lib import call

    a, b = C.c_void_p(), C.c_void_p()
    call("mh_device_alloc", nbytes, C.byref(a))
    call("mh_device_alloc", nbytes, C.byref(b))
    n = nbytes // 8
    call("mh_copy", a, b, n, None)
    call("mh_synchronize", None)
    t0 = time.perf_counter()
    for _ in range(reps):
        call("mh_copy", a, b, n, None)
    call("mh_synchronize", None)
    dt = (time.perf_counter() - t0) / reps
    call("mh_device_free", a)
    call("mh_device_free", b)
    return 2.0 * nbytes / dt


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--size", type=int, default=8192)
    ap.add_argument("--solver", default="hllc")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extras", action="store_true", help="skip the batched and adjoint legs")
    ap.add_argument("--leg", choices=["batched", "adjoint"], default=None,
                    help="run only this extra leg (batched: against --plain-rate Gcell/s)")
    ap.add_argument("--plain-rate", type=float, default=32.0)
    ap.add_argument("--copy-only", type=int, default=0,
                    help="only run the copy kernel on buffers of this many MiB (counter calibration)")
    args = ap.parse_args()
    if args.copy_only:
        bw = copy_peak(args.copy_only << 20, reps=3)
        print(json.dumps({"copy_bytes": args.copy_only << 20, "copy_GB_per_s": bw / 1e9}))
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; the multi-GPU path (y-slab decomposition "
                                   "with ghost-row exchange over xGMI) is a design, not a measured result"}))
        sys.exit(2)
    if args.leg == "batched":
        print(json.dumps(batched_leg(args.plain_rate * 1e9)))
        return
    if args.leg == "adjoint":
        print(json.dumps(adjoint_leg()))
        return

    from large_sys_eq_opt_amd import Grid2DGPU, Stepper, cases, parse_riemann_solver

    solver = parse_riemann_solver(args.solver)
    n = args.size
    if n % 256:
        raise SystemExit("--size must be a multiple of 256 (cyl_blast scale factor)")
    cfg = cases.get_n_case_config("cyl_blast", n // 256)
    kind, params = cases.case_params("cyl_blast")
    U = Grid2DGPU.from_config(cfg)
    U.init_state(kind, params.array())
    st = Stepper(U, solver, cfg.cfl, 0.0, cases.case_boundary("cyl_blast"),
                 log_capacity=args.steps + args.warmup + 8)
    if args.warmup:
        st.advance(args.warmup)
    st.clock()   # synchronises: inputs resident, warm-up done
    t0 = time.perf_counter()
    ms_total, ms_k = st.advance_profiled(args.steps)   # synchronises at the end
    wall = time.perf_counter() - t0
    t_sim, nsteps, dts = st.clock()
    if nsteps != args.steps + args.warmup:
        raise SystemExit(f"stepper ran {nsteps} steps, expected {args.steps + args.warmup}")
    cells = cfg.nx * cfg.ny
    sec = wall
    rate = cells * args.steps / sec
    names = ["clock", "sweep_x", "sweep_y"]
    kern = {names[i]: {"avg_ms": ms_k[i]} for i in range(3)}
    for nm in ("sweep_x", "sweep_y"):
        kern[nm]["GB_per_s"] = BYTES_PER_SWEEP * cells / (kern[nm]["avg_ms"] * 1e-3) / 1e9
    dom = max(("sweep_x", "sweep_y"), key=lambda k: kern[k]["avg_ms"])
    achieved = BYTES_PER_SWEEP * cells / (kern[dom]["avg_ms"] * 1e-3)
    traffic = None
    tfile = os.path.join(ROOT, "profiles", f"traffic_{dom}_{args.solver.lower()}_cyl_blast_{n}.json")
    if os.path.exists(tfile):
        with open(tfile) as f:
            traffic = json.load(f).get("bytes_per_launch")
    out = {
        "metric": METRIC,
        "value": rate / 1e9,
        "unit": "Gcell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": sec * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"cyl_blast {n}x{n} {args.solver.upper()} FP64 MUSCL-Hancock, x-then-y split, "
                               f"device-resident dt (clock + x sweep + y sweep per step)",
                   "nx": cfg.nx, "ny": cfg.ny, "solver": args.solver, "cfl": cfg.cfl, "precision": "fp64"},
        "roofline": {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK, "traffic": traffic, "kernel": dom,
                     "algorithmic_bytes_per_launch": BYTES_PER_SWEEP * cells},
        "step_hbm_frac": rate * 2 * BYTES_PER_SWEEP / HBM_PEAK,
        "kernels": kern,
        "event_ms_total": ms_total,
        "sim_time": t_sim,
    }
    st.close()
    U.close()
    out["copy_peak_GB_per_s"] = copy_peak() / 1e9
    out["roofline"]["frac_of_copy_peak"] = achieved / (out["copy_peak_GB_per_s"] * 1e9)
    if not args.no_extras:
        out["batched"] = batched_leg(rate)
        out["adjoint"] = adjoint_leg()
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(int(solver))
        out["cpu_baseline"].pop("np", None)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out))


if __name__ == "__main__":
    main()
