"""Benchmark: FP64 MUSCL-Hancock HLLC throughput at 8192^2 on one MI355X.

A step is one pass of the hot path over the grid: the device clock (dt from
the previous step's fused max-speed reduction) and the split step -- on
this grid the fused x-then-y step kernel (one read and one write of the
state per step; the two-sweep form with --fused 0).  The workload is the
cylindrical blast (rho=1, p=1 inside r <= 0.4 of (1,1); rho=0.125, p=0.1
outside; [0,2]^2) on an 8192 x 8192 grid, HLLC, CFL 0.8, state resident in
HBM before the timed region.

value: cell updates per second over the timed steps (Gcell-updates/s).
roofline: the dominant kernel -- its algorithmic bytes per launch at 128 B
  per cell update (the split step's x sweep and y sweep each reading and
  writing the four FP64 planes: the per-unit figure the target is set in)
  times the cell updates one launch performs (a whole step for the fused
  kernel, half a step's bytes for one sweep), divided by its average
  duration from HIP events recorded on the launch stream over the timed
  region; peak 8 TB/s.  "traffic" is the launch's measured memory traffic
  (two counter passes, profiles/).  The fused kernel reads and writes the
  state once, 64 B per cell update; that rate is reported beside it
  ("kernel_bytes"), and the kernel is bound by FP64 VALU issue rather than
  by HBM.  The whole-step fraction (128 B per cell update over the step
  time, clock included) is "step_hbm_frac".
cpu_baseline: the reference CPU solver compiled from its sources
  (oracle/_ref) -- or the oracle port when that build is absent -- on the
  same workload (8192^2 HLLC) for a bounded sample (median of 3 runs of a
  few steps, ~20 s in all), OpenMP threads bound to cores; the CPU model and
  core counts are recorded.
batched: 512 blast-wave instances (seed 0, sides log-uniform in [64,1024],
  95.3 M cells), 500 steps with per-instance dt replayed from captured
  hipGraphs under the autotuner's schedule (the line names it: "mode" is
  the fused step, dispatched launches or the persistent kernel), its cell-update rate and
  ratio to the single-grid value; the persistent kernel's best schedule (one
  launch per step with a grid barrier between the phases, or two launches)
  is timed over the same 500 steps beside it ("persistent").
adjoint: the 512x256 shock-bubble adjoint case (400 HLLC steps, a
  checkpoint every 20): the backward pass (segment recompute + adjoint
  steps) per step over the device-resident forward step, and the whole
  gradient (checkpointing forward + backward) over it.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--size N]
       [--solver hllc] [--fused 1|0] [--no-cpu] [--no-extras]
       [--leg batched|adjoint|snapshot|cpu|slab]
(--leg runs one extra leg alone and prints its JSON object: the profiling
runs of those workloads use it; "snapshot" measures the cost of the device
loop's snapshots at --size, "cpu" times the CPU baseline alone at --size
and --solver, "slab" one GPU's y-slab of --size^2 split over 2, 4 and 8
GPUs.)
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


def host_cpu() -> dict:
    """The host's CPU model and core counts (lscpu), and how many of its
    logical CPUs this process may run on (its affinity mask)."""
    info = {"model": None, "physical_cores": None, "logical_cpus": os.cpu_count(),
            "affinity_cpus": len(os.sched_getaffinity(0))}
    try:
        txt = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout
        for line in txt.splitlines():
            if line.startswith("Model name:"):
                info["model"] = line.split(":", 1)[1].strip()
        par = subprocess.run(["lscpu", "-p=CPU,CORE,SOCKET"], capture_output=True, text=True, timeout=20).stdout
        cores, mine = set(), set()
        allowed = os.sched_getaffinity(0)
        for line in par.splitlines():
            if line.startswith("#"):
                continue
            cpu, core, sock = (int(x) for x in line.split(",")[:3])
            cores.add((sock, core))
            if cpu in allowed:
                mine.add((sock, core))
        info["physical_cores"] = len(cores)
        info["physical_cores_in_affinity"] = len(mine)
    except (OSError, ValueError, subprocess.SubprocessError):
        pass
    return info


def cpu_baseline(solver: int, n: int = 8192, samples: int = 3, budget_s: float = 20.0):
    """The reference CPU solver compiled from its sources (oracle/_ref; the
    oracle port where that build is absent) on the benchmark workload itself
    -- cyl_blast at n^2 -- for a bounded sample: one calibration step, then
    `samples` runs of k steps sized to ~budget_s in total; the median rate.
    OpenMP threads: OMP_NUM_THREADS as the environment sets it (else the
    physical cores this process may use), bound close to cores
    (OMP_PROC_BIND=close, OMP_PLACES=cores, set before the library loads, as
    scripts/slurm/cpu_all.sh:110-111 does)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np
    import pyoracle

    from large_sys_eq_opt_amd import cases

    cfg = cases.get_n_case_config("cyl_blast", n // 256)
    kind, params = cases.case_params("cyl_blast")
    orc = pyoracle.Oracle()
    U0 = orc.init(kind, params.array(), cfg.nx, cfg.ny, cfg.bounds)
    if pyoracle.reference_available():
        impl, kind_s = pyoracle.Reference(), "reference"
        run = lambda k: impl.run(U0, cfg.bounds, solver, cfg.cfl, max_steps=k)  # noqa: E731
    else:
        impl, kind_s = orc, "port"
        run = lambda k: impl.run(U0, cfg.dx, cfg.dy, solver, cfg.cfl, max_steps=k)  # noqa: E731
    probe = run(1)
    k = int(max(1, min(200, budget_s / samples / max(probe["seconds"], 1e-6))))
    rates = []
    for _ in range(samples):
        r = run(k)
        rates.append(cfg.nx * cfg.ny * r["steps"] / r["seconds"])
    rate = float(np.median(rates))
    threads = orc.num_threads()
    cpu = _HOST_CPU or host_cpu()
    omp = {v: os.environ.get(v) for v in ("OMP_NUM_THREADS", "OMP_PROC_BIND", "OMP_PLACES")}
    return {"value": rate / 1e9, "unit": "Gcell-updates/s", "cores": threads, "kind": kind_s,
            "sample": f"cyl_blast {cfg.nx}x{cfg.ny} {['HLL', 'HLLC', 'exact', 'FORCE'][solver]} FP64 (the benchmark "
                      f"workload), median of {samples} runs of {k} steps from the initial state "
                      f"({', '.join(f'{x / 1e9:.4f}' for x in rates)} Gcell-updates/s), {threads} OpenMP threads, "
                      f"{'reference CPU sources' if kind_s == 'reference' else 'oracle port'}",
            "cpu_model": cpu["model"], "physical_cores": cpu.get("physical_cores"),
            "physical_cores_in_affinity": cpu.get("physical_cores_in_affinity"),
            "affinity_cpus": cpu["affinity_cpus"], "omp_env": omp}


_HOST_CPU = None


def _set_omp_env():
    """OpenMP placement of the CPU baseline, set before any OpenMP runtime
    loads (and the host's CPU read before any thread is bound).  An
    OMP_NUM_THREADS already in the environment is kept.  Only the CPU
    baseline's own process sets it (cpu_baseline_subprocess): a bound
    OpenMP runtime pins the thread that starts it, and the threads that
    thread starts later (the snapshot writer's) would inherit that."""
    global _HOST_CPU
    _HOST_CPU = host_cpu()
    os.environ.setdefault("OMP_PROC_BIND", "close")
    os.environ.setdefault("OMP_PLACES", "cores")
    if "OMP_NUM_THREADS" not in os.environ:
        n = _HOST_CPU.get("physical_cores_in_affinity") or _HOST_CPU["affinity_cpus"]
        os.environ["OMP_NUM_THREADS"] = str(max(1, n))


def cpu_baseline_subprocess(solver_name: str, n: int):
    """cpu_baseline in a process of its own (bench.py --leg cpu), so that its
    OpenMP placement does not bind this process's threads."""
    r = subprocess.run([sys.executable, os.path.abspath(__file__), "--leg", "cpu", "--size", str(n), "--solver",
                        solver_name], capture_output=True, text=True, timeout=900)
    if r.returncode != 0:
        raise RuntimeError("CPU baseline failed: " + r.stderr[-2000:])
    return json.loads(r.stdout.strip().splitlines()[-1])


MODES = {0: "persistent: one launch per step, grid barrier between the x and y phases",
         1: "persistent: two launches per step (x phase, y phase)",
         2: "dispatched: two launches per step (x tiles; y tiles + IEEE redo), one tile per wavefront",
         3: "fused: one launch per step (an x/y wavefront pair per y tile, x rows handed to y through LDS) "
            "+ its IEEE redo and clock commit"}


def _timed_batch(spec, schedule, steps, per_graph=25, size_class=(0, 0, 0)):
    from large_sys_eq_opt_amd import ensemble

    b = ensemble.Batch(spec, solver=1, cfl=0.8)
    b.set_schedule(*schedule)
    b.set_size_class(*size_class)
    b.initialise()
    b.capture(per_graph)
    # one untimed replay (the arena's first use, clocks up), then back to the
    # initial states and clocks; the captured graph stays valid
    b.step(per_graph)
    b.initialise()
    b.clock()
    t0 = time.perf_counter()
    b.step(steps)
    _, _, st, _ = b.clock()
    wall = time.perf_counter() - t0
    info = b.info()
    cells = b.cells
    b.close()
    return cells * steps / wall, wall, info, bool((st == steps).all())


def batched_leg(plain_rate: float, steps: int = 500):
    """512-instance ensemble, graph-replayed, under the autotuned schedule and
    under the persistent kernel."""
    from large_sys_eq_opt_amd import ensemble

    spec = ensemble.ensemble_spec(512, 0)
    b = ensemble.Batch(spec, solver=1, cfl=0.8)
    b.initialise()
    tune = ensemble.autotune(b)
    b.close()
    sched, cls = tuple(tune["schedule"]), tuple(tune["size_class"])
    per_graph = 25
    rate, wall, info, ok = _timed_batch(spec, sched, steps, per_graph, cls)
    pers = min((t, c) for c, t in tune["timings"] if c[2] in (0, 1))[1]   # the persistent kernel's best
    prate, pwall, pinfo, pok = _timed_batch(spec, pers, steps, per_graph)
    return {"value": rate / 1e9, "unit": "Gcell-updates/s", "instances": len(spec),
            "cells": sum(x["side"] ** 2 for x in spec), "steps": steps, "ms_per_step": wall * 1e3 / steps,
            "ratio_vs_plain": rate / plain_rate, "schedule": list(sched), "mode": MODES[sched[2]],
            "size_class": {"side_below": cls[0], "ychunk": cls[1], "xrows": cls[2]},
            "graph_steps": per_graph, "blocks": info["blocks"], "all_instances_stepped": ok,
            "persistent": {"schedule": list(pers), "mode": MODES[pers[2]], "value": prate / 1e9,
                           "ms_per_step": pwall * 1e3 / steps, "ratio_vs_plain": prate / plain_rate,
                           "blocks": pinfo["blocks"], "all_instances_stepped": pok}}


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
            "ratio_vs_checkpointing_forward": b / c, "whole_gradient_ratio_vs_forward": (c + b) / f}


def slab_leg(n: int = 8192, steps: int = 200, warmup: int = 30):
    """One GPU's share of the y-slab decomposition of the n^2 workload over
    G = 2, 4, 8 GPUs: an n x n/G slab through the blast's centre, HLLC, its
    top and bottom ghost rows held at a fixed state (a stand-in for the
    neighbours' halo rows; the exchange itself is not modelled), the device
    loop timed as the plain leg."""
    from large_sys_eq_opt_amd import Grid2DGPU, Stepper, cases
    from large_sys_eq_opt_amd.solver import Boundary, synchronize

    kind, params = cases.case_params("cyl_blast")
    inner = [1.0, 0.0, 0.0, 1.0 / 0.4]   # the blast's inner state, conserved
    bc = Boundary((0, 0, 1, 1), [[0.0] * 4, [0.0] * 4, inner, inner])
    out = []
    for g in (2, 4, 8):
        ny = n // g
        h = 2.0 / n
        cfg = cases.CaseConfig(n, ny, 2, 0.0, 2.0, 1.0 - 0.5 * ny * h, 1.0 + 0.5 * ny * h, 0.8, 0.0)
        U = Grid2DGPU.from_config(cfg)
        U.init_state(kind, params.array())
        st = Stepper(U, 1, cfg.cfl, 0.0, bc, log_capacity=steps + warmup + 8)
        st.advance(warmup)
        st.clock()
        t0 = time.perf_counter()
        _, ms_k = st.advance_profiled(steps)
        wall = time.perf_counter() - t0
        st.clock()
        st.close()
        U.close()
        synchronize()
        rate = n * ny * steps / wall
        out.append({"gpus": g, "slab": f"{n}x{ny}", "per_gpu_Gcell_updates_per_s": rate / 1e9,
                    "ms_per_step": wall * 1e3 / steps, "kernel_ms": [round(x, 4) for x in ms_k],
                    "sum_of_slabs_Gcell_updates_per_s": g * rate / 1e9})
    return {"workload": f"cyl_blast {n}x{n} HLLC split into y-slabs, one slab per GPU, ghost rows held fixed",
            "slabs": out}


def snapshot_leg(n: int = 8192, solver: int = 1, csv_dir: str | None = None):
    """Snapshot cost on the benchmark workload (cyl_blast n^2, run to
    t_end = 0.25 in the device-resident loop): the loop alone; the same loop
    holding at the 5 snapshot times, each snapshot copied on the device and
    then to pinned host memory on the copy stream while the next steps run;
    and one snapshot written as the reference's CSV by the library's
    writer (to csv_dir or a temporary directory, removed afterwards; skipped
    without 4x its size free)."""
    import shutil
    import tempfile

    from large_sys_eq_opt_amd import Grid2DGPU, Snapshot, Stepper, cases

    cfg = cases.get_n_case_config("cyl_blast", n // 256)
    kind, params = cases.case_params("cyl_blast")
    bc = cases.case_boundary("cyl_blast")
    times = [cfg.t_end * k / 5.0 for k in range(1, 6)]
    out = {"workload": f"cyl_blast {n}x{n} HLLC to t_end {cfg.t_end}, device-resident loop, 5 snapshots"}
    U = Grid2DGPU.from_config(cfg)
    snap = None
    for with_snaps in (False, True):
        U.init_state(kind, params.array())
        st = Stepper(U, solver, cfg.cfl, cfg.t_end, bc)
        if with_snaps and snap is None:
            snap = Snapshot(U, 2)
        t0 = time.perf_counter()
        nk, t, steps, waits = 0, 0.0, 0, 0.0
        while True:
            if with_snaps:
                st.set_stop(times[nk] if nk < 5 else 0.0)
            st.advance(256)
            t, steps, _ = st.clock()
            while with_snaps and nk < 5 and t >= times[nk]:
                if nk >= 2:   # a slot is reused: its previous copy must have landed
                    w0 = time.perf_counter()
                    snap.planes(nk % 2)
                    waits += time.perf_counter() - w0
                st.finish()
                snap.capture(U, nk % 2)
                nk += 1
            if t >= cfg.t_end:
                break
        if with_snaps:
            snap.planes(0)
            snap.planes(1)
        wall = time.perf_counter() - t0
        st.close()
        key = "loop_with_5_snapshots" if with_snaps else "loop"
        out[key] = {"seconds": wall, "steps": steps}
        if with_snaps:
            out[key]["host_wait_seconds"] = waits
    out["snapshot_overhead_seconds"] = out["loop_with_5_snapshots"]["seconds"] - out["loop"]["seconds"]
    out["state_bytes"] = 4 * 8 * (cfg.nx + 4) * (cfg.ny + 4)
    d = csv_dir or tempfile.mkdtemp(prefix="mh_snap_")
    try:
        need = 4 * 70 * cfg.nx * cfg.ny
        free = shutil.disk_usage(d).free
        if free < need:
            out["csv"] = {"skipped": f"{free / 1e9:.1f} GB free in {d}, about {need / 4e9:.1f} GB needed"}
        else:
            fn = os.path.join(d, "gpu_cyl_blast_hllc_snapshot_5.csv")
            t0 = time.perf_counter()
            snap.write_csv(1, fn, 6)
            w = time.perf_counter() - t0
            size = os.path.getsize(fn)
            out["csv"] = {"seconds": w, "bytes": size, "MB_per_s": size / w / 1e6, "rows": cfg.nx * cfg.ny,
                          "threads": min(16, os.cpu_count() or 1)}
            os.remove(fn)
    finally:
        if not csv_dir:
            shutil.rmtree(d, ignore_errors=True)
    snap.close()
    U.close()
    return out


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


def smooth_state(n):
    """An (n+4)^2 ghosted AoS state with no two equal neighbours: smooth
    density, velocity and pressure waves across the unit square (rho in
    [0.7, 1.3], p in [0.6, 1.4], |u|, |v| <= 0.3)."""
    import numpy as np

    y, x = (np.mgrid[0:n + 4, 0:n + 4] + 0.5) / (n + 4)
    rho = 1.0 + 0.3 * np.sin(2 * np.pi * x + 0.3) * np.cos(2 * np.pi * y + 1.1)
    u = 0.3 * np.sin(2 * np.pi * (y + 0.37 * x) + 0.7)
    v = 0.2 * np.cos(2 * np.pi * (x - 0.41 * y) + 2.0)
    p = 1.0 + 0.4 * np.cos(2 * np.pi * (x + 0.83 * y) + 0.4)
    U = np.empty((n + 4, n + 4, 4))
    U[..., 0] = rho
    U[..., 1] = rho * u
    U[..., 2] = rho * v
    U[..., 3] = p / 0.4 + 0.5 * rho * (u * u + v * v)
    return U


def main():
    if "--leg" in sys.argv and "cpu" in sys.argv:
        _set_omp_env()
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=30)
    ap.add_argument("--size", type=int, default=8192)
    ap.add_argument("--solver", default="hllc")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extras", action="store_true", help="skip the batched and adjoint legs")
    ap.add_argument("--leg", choices=["batched", "adjoint", "snapshot", "cpu", "slab"], default=None,
                    help="run only this extra leg (batched: against --plain-rate Gcell/s)")
    ap.add_argument("--plain-rate", type=float, default=32.0)
    ap.add_argument("--state", choices=["case", "smooth"], default="case",
                    help="case: the cylindrical blast's initial state; smooth: a smooth state with no two equal "
                         "neighbours (no uniform region), on the same grid")
    ap.add_argument("--fused", type=int, choices=[0, 1, 2], default=1,
                    help="1: the fused step where it applies on grids of 2^20 cells or more (default); 2: on every grid; 0: the x sweep and the y sweep")
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
    if args.leg == "snapshot":
        print(json.dumps(snapshot_leg(args.size)))
        return
    if args.leg == "slab":
        print(json.dumps(slab_leg(args.size)))
        return
    if args.leg == "cpu":
        from large_sys_eq_opt_amd import parse_riemann_solver

        print(json.dumps(cpu_baseline(int(parse_riemann_solver(args.solver)), args.size, 3,
                                      20.0 if args.size >= 8192 else 12.0)))
        return

    from large_sys_eq_opt_amd import Grid2DGPU, Stepper, cases, parse_riemann_solver
    from large_sys_eq_opt_amd.solver import set_launch_shape

    # the box's copy bandwidth first: it also brings the GPU's clocks up from
    # idle before the stepper's warm-up (from idle, the first ~25 ms of steps
    # run up to 25 % slower)
    copy_bw = copy_peak()
    set_launch_shape(fused=args.fused)
    solver = parse_riemann_solver(args.solver)
    n = args.size
    if n % 256:
        raise SystemExit("--size must be a multiple of 256 (cyl_blast scale factor)")
    cfg = cases.get_n_case_config("cyl_blast", n // 256)
    kind, params = cases.case_params("cyl_blast")
    U = Grid2DGPU.from_config(cfg)
    U.init_state(kind, params.array())
    if args.state == "smooth":
        U.upload_from_aos(smooth_state(n))
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
    fused = ms_k[2] == 0.0
    names = ["clock", "step_fused"] if fused else ["clock", "sweep_x", "sweep_y"]
    kern = {names[i]: {"avg_ms": ms_k[i]} for i in range(len(names))}
    for nm in names[1:]:
        # bytes the launch itself reads and writes: the state once
        kern[nm]["GB_per_s"] = BYTES_PER_SWEEP * cells / (kern[nm]["avg_ms"] * 1e-3) / 1e9
    dom = max(names[1:], key=lambda k: kern[k]["avg_ms"])
    # algorithmic bytes per launch at 128 B per cell update: a whole step
    # for the fused kernel, one sweep's half for a sweep
    alg_bytes = (2 if fused else 1) * BYTES_PER_SWEEP * cells
    achieved = alg_bytes / (kern[dom]["avg_ms"] * 1e-3)
    traffic = None
    tfile = os.path.join(ROOT, "profiles", f"traffic_{dom}_{args.solver.lower()}_cyl_blast_{n}.json")
    if args.state == "case" and os.path.exists(tfile):
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
        "config": {"workload": (f"cyl_blast {n}x{n}" if args.state == "case" else f"smooth non-uniform state {n}x{n}")
                               + f" {args.solver.upper()} FP64 MUSCL-Hancock, x-then-y split, "
                               + ("device-resident dt (clock + fused x-then-y step per step)" if fused else
                                  "device-resident dt (clock + x sweep + y sweep per step)"),
                   "nx": cfg.nx, "ny": cfg.ny, "solver": args.solver, "cfl": cfg.cfl, "precision": "fp64"},
        "roofline": {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK, "traffic": traffic, "kernel": dom,
                     "algorithmic_bytes_per_launch": alg_bytes,
                     "kernel_bytes": {"read_write_per_launch": BYTES_PER_SWEEP * cells,
                                      "GB_per_s": kern[dom]["GB_per_s"],
                                      "frac": kern[dom]["GB_per_s"] * 1e9 / HBM_PEAK}},
        "step_hbm_frac": rate * 2 * BYTES_PER_SWEEP / HBM_PEAK,
        "kernels": kern,
        "event_ms_total": ms_total,
        "sim_time": t_sim,
    }
    st.close()
    U.close()
    out["copy_peak_GB_per_s"] = copy_bw / 1e9
    out["roofline"]["frac_of_copy_peak"] = achieved / (out["copy_peak_GB_per_s"] * 1e9)
    if not args.no_extras:
        out["batched"] = batched_leg(rate)
        out["adjoint"] = adjoint_leg()
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline_subprocess(args.solver, n)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out))


if __name__ == "__main__":
    main()
