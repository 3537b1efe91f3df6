"""Time loop, snapshot schedule and snapshot output.

``run_case`` is the loop of scripts/gpu/main_gpu.cu:243-299 and
scripts/cpu/main_cpu.cpp:206-262: while t < t_end: dt = compute_dt (the
boundary fill is fused into the previous step), dt = min(dt, t_end - t),
advance, swap, t += dt; a snapshot is written at the first step with
t >= k*t_end/num_snapshots.  ``mode="host"`` computes every dt with a host
round trip exactly as the reference does; ``mode="device"`` keeps dt and t
on the device (same arithmetic): the clock holds at each snapshot time, the
state is copied on the device and then to pinned host memory on a copy
stream while the next steps run, and a worker thread writes the CSV.

Snapshot CSV: header ``i,j,x,y,rho,rhou,rhov,E``, one row per interior cell,
i and j including the ghost offset, 6 significant digits
(scripts/cpu/main_cpu.cpp:119-164), formatted by the library on several
host threads; ``precision=17`` writes round-trip exact values instead.
"""
from __future__ import annotations

import os
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np

from . import cases as _cases
from .solver import (Boundary, Grid2DGPU, Snapshot, Stepper, advance_first_order_gpu, advance_second_order_gpu,
                     compute_dt_gpu, parse_riemann_solver, riemann_solver_to_string, write_csv_aos)


def snapshot_name(prefix: str, case_name: str, solver_name: str, scale: int, k: int, outdir="outputs") -> str:
    return os.path.join(outdir, f"{prefix}_{case_name}_{solver_name}_n{scale}_snapshot_{k}.csv")


def write_grid_csv(U_aos: np.ndarray, cfg, filename: str, precision: int = 6) -> None:
    """scripts/cpu/main_cpu.cpp:136-164 (ostream default = 6 significant digits)."""
    os.makedirs(os.path.dirname(filename) or ".", exist_ok=True)
    write_csv_aos(U_aos, cfg, filename, precision)


def read_grid_csv(filename: str, nx: int, ny: int) -> np.ndarray:
    """Reads a snapshot CSV back into an interior array of shape (ny, nx, 4)."""
    d = np.loadtxt(filename, delimiter=",", skiprows=1)
    out = np.zeros((ny, nx, 4))
    out[d[:, 1].astype(int) - 2, d[:, 0].astype(int) - 2] = d[:, 4:8]
    return out


def run_case(case, solver, scale: int = 1, *, mode: str = "host", num_snapshots: int = 0, write_output=False,
             outdir="outputs", precision=6, order: int = 2, max_steps: int | None = None, t_end: float | None = None,
             bc: Boundary | None = None, U0: np.ndarray | None = None, keep_snapshots=False, log=None,
             cfl: float | None = None):
    """Runs a case through the HIP solver; returns dict(U, t, steps, dt, snap_t, seconds).
    t_end and cfl override the case's values."""
    case_id = _cases.parse_case_id(case)
    s = parse_riemann_solver(solver)
    cfg = _cases.get_n_case_config(case_id, scale)
    if t_end is not None:
        cfg.t_end = t_end
    if cfl is not None:
        cfg.cfl = float(cfl)
    bc = _cases.case_boundary(case_id) if bc is None else bc
    U = Grid2DGPU.from_config(cfg)
    if U0 is None:
        kind, params = _cases.case_params(case_id)
        U.init_state(kind, params.array())
    else:
        U.upload_from_aos(U0)
    if max_steps is None and not cfg.t_end > 0:
        raise ValueError("run_case: give max_steps when the case has no t_end")
    limit = max_steps if max_steps is not None else (1 << 62)
    step = 0
    snap_times = [cfg.t_end * float(k) / float(num_snapshots) for k in range(1, num_snapshots + 1)]
    snaps, snap_t, dts = [], [], []
    t, step = 0.0, 0
    t0 = time.perf_counter()
    if mode == "device":
        # the case's t_end always bounds the run (0: none); max_steps may
        # end it earlier, as in host mode
        cap = int(min(limit, 1 << 22)) if max_steps is not None else 1 << 20
        st = Stepper(U, s, cfg.cfl, cfg.t_end if cfg.t_end > 0 else 0.0, bc, log_capacity=cap)
        want = bool(num_snapshots) and (write_output or keep_snapshots)
        snap = Snapshot(U, 2) if want else None
        pool = ThreadPoolExecutor(1) if want else None
        pending = [None, None]
        if keep_snapshots:
            snaps = [None] * num_snapshots
        names = {"case": _cases.case_id_to_string(case_id), "solver": riemann_solver_to_string(s)}

        def emit(slot, k, t_k):
            if keep_snapshots:
                snaps[k] = snap.aos(slot)
            if write_output:
                fn = snapshot_name("gpu", names["case"], names["solver"], scale, k + 1, outdir)
                os.makedirs(os.path.dirname(fn) or ".", exist_ok=True)
                snap.write_csv(slot, fn, precision)
                if log:
                    log(f"[GPU] Wrote snapshot {k + 1} at t = {t_k:.6g} -> {fn}")

        nk = 0
        while True:
            if num_snapshots:
                st.set_stop(snap_times[nk] if nk < num_snapshots else 0.0)
            st.advance(int(min(256, limit - step)))
            t, step, _ = st.clock()
            while nk < num_snapshots and t >= snap_times[nk]:
                if want:
                    slot = nk % 2
                    if pending[slot] is not None:
                        pending[slot].result()
                    st.finish()   # every ghost cell, as the host loop's state has them
                    snap.capture(U, slot)
                    pending[slot] = pool.submit(emit, slot, nk, t)
                snap_t.append(t)
                nk += 1
            if step >= limit or (cfg.t_end > 0 and t >= cfg.t_end):
                break
        t, step, dts = st.clock()
        st.finish()
        st.close()
        if want:
            for f in pending:
                if f is not None:
                    f.result()
            pool.shutdown()
            snap.close()
    else:
        tmp, nxt = Grid2DGPU.from_config(cfg), Grid2DGPU.from_config(cfg)
        from .solver import apply_boundary_gpu

        apply_boundary_gpu(U, bc)
        nk = 0
        while (cfg.t_end > 0 and t < cfg.t_end or cfg.t_end <= 0) and step < limit:
            dt = compute_dt_gpu(U, None, cfg.cfl)
            if cfg.t_end > 0 and t + dt > cfg.t_end:
                dt = cfg.t_end - t
            if order == 1:
                advance_first_order_gpu(U, nxt, dt, s, bc)
            else:
                advance_second_order_gpu(U, tmp, nxt, None, dt, s, bc)
            U.swap(nxt)
            t += dt
            step += 1
            dts.append(dt)
            while nk < num_snapshots and t >= snap_times[nk]:
                if write_output or keep_snapshots:
                    A = U.download_to_aos()
                    if keep_snapshots:
                        snaps.append(A)
                    if write_output:
                        fn = snapshot_name("gpu", _cases.case_id_to_string(case_id), riemann_solver_to_string(s),
                                           scale, nk + 1, outdir)
                        write_grid_csv(A, cfg, fn, precision)
                        if log:
                            log(f"[GPU] Wrote snapshot {nk + 1} at t = {t:.6g} -> {fn}")
                snap_t.append(t)
                nk += 1
        tmp.close()
        nxt.close()
    seconds = time.perf_counter() - t0
    out = dict(U=U.download_to_aos(), t=t, steps=step, dt=np.asarray(dts), snap_t=np.asarray(snap_t),
               snapshots=snaps, seconds=seconds, cfg=cfg)
    U.close()
    return out
