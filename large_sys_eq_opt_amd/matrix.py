"""The reference's benchmark matrix on the HIP solver: every resolution scale
x case x Riemann solver run through the command line, one summary row per
run (scripts/slurm/gpu_all.sh:183-234, header :156).

    python -m large_sys_eq_opt_amd.matrix [--scales 1 2 4] [--cases shock_bubble blast_wave]
        [--solvers hll hllc exact force] [--csv validation/gpu_matrix.csv] [--logdir logs]
        [--host-loop] [--tag TAG]

Each run is ``cli.main([n, --case, c, --solver, s, --no-output, --device-loop])``
with its messages written to a log file, as the reference tees each
``main_gpu`` run into logs/; nx, ny, total_cells and the step count are read
back from that log the way the reference's summary does (the lines
"[GPU] nx : ...", "[GPU] Total steps = ...").  Columns:

    arch,case,solver,n,nx,ny,total_cells,total_steps,real_seconds,user_seconds,sys_seconds,max_rss_kb,exit_status,log_file

real_seconds is the run's wall time inside this process (the reference's
/usr/bin/time also counts process start-up); user/sys are this process's
CPU-time deltas over the run; max_rss_kb is the process's peak resident set
so far (getrusage), not the run's alone.  --host-loop uses the
reference-compatible host loop (a dt round trip per step) instead of the
device-resident one; both give the same step counts.
"""
from __future__ import annotations

import argparse
import os
import re
import resource
import sys
import time

HEADER = ("arch,case,solver,n,nx,ny,total_cells,total_steps,real_seconds,user_seconds,sys_seconds,max_rss_kb,"
          "exit_status,log_file")


def _value(pattern: str, text: str):
    m = re.search(pattern, text, re.M)
    return m.group(1) if m else ""


def run_one(n: int, case: str, solver: str, logdir: str, tag: str, device_loop: bool = True) -> dict:
    from .cli import main as cli_main

    os.makedirs(logdir, exist_ok=True)
    log_file = os.path.join(logdir, f"gpu_clean_{case}_{solver}_n{n}_{tag}.log")
    argv = [str(n), "--case", case, "--solver", solver, "--no-output"] + (["--device-loop"] if device_loop else [])
    r0 = resource.getrusage(resource.RUSAGE_SELF)
    t0 = time.perf_counter()
    with open(log_file, "w") as f:
        rc = cli_main(argv, out=f, err=f)
    wall = time.perf_counter() - t0
    r1 = resource.getrusage(resource.RUSAGE_SELF)
    with open(log_file) as f:
        text = f.read()
    return {
        "arch": "gpu", "case": case, "solver": solver, "n": n,
        "nx": _value(r"^\[GPU\] nx\s*:\s*(\d+)", text),
        "ny": _value(r"^\[GPU\] ny\s*:\s*(\d+)", text),
        "total_cells": _value(r"^\[GPU\] total_cells\s*:\s*(\d+)", text),
        "total_steps": _value(r"^\[GPU\] Total steps = (\d+)", text),
        "real_seconds": f"{wall:.2f}",
        "user_seconds": f"{r1.ru_utime - r0.ru_utime:.2f}",
        "sys_seconds": f"{r1.ru_stime - r0.ru_stime:.2f}",
        "max_rss_kb": str(r1.ru_maxrss),
        "exit_status": str(rc),
        "log_file": log_file,
    }


def run_matrix(scales, case_names, solvers, csv_path: str, logdir: str = "logs", tag: str = "local",
               device_loop: bool = True, log=print) -> list:
    """Runs scales x cases x solvers in the reference's order (scale outermost)
    and writes the summary CSV; stops at the first failing run, as the
    reference's script does.  Returns the rows."""
    rows = []
    os.makedirs(os.path.dirname(csv_path) or ".", exist_ok=True)
    with open(csv_path, "w") as out:
        out.write(HEADER + "\n")
        for n in scales:
            for case in case_names:
                for solver in solvers:
                    row = run_one(int(n), case, solver, logdir, tag, device_loop)
                    rows.append(row)
                    out.write(",".join(str(row[k]) for k in HEADER.split(",")) + "\n")
                    out.flush()
                    if log:
                        log(f"[matrix] {case} {solver} n={n}: {row['total_steps']} steps in {row['real_seconds']} s "
                            f"(exit {row['exit_status']})")
                    if row["exit_status"] != "0":
                        return rows
    return rows


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--scales", type=int, nargs="+", default=[1, 2, 4])
    ap.add_argument("--cases", nargs="+", default=["shock_bubble", "blast_wave"])
    ap.add_argument("--solvers", nargs="+", default=["hll", "hllc", "exact", "force"])
    ap.add_argument("--csv", default=os.path.join("validation", "gpu_matrix.csv"))
    ap.add_argument("--logdir", default="logs")
    ap.add_argument("--tag", default="local")
    ap.add_argument("--host-loop", action="store_true")
    a = ap.parse_args(sys.argv[1:] if argv is None else argv)
    rows = run_matrix(a.scales, a.cases, a.solvers, a.csv, a.logdir, a.tag, not a.host_loop)
    return 0 if rows and all(r["exit_status"] == "0" for r in rows) else 1


if __name__ == "__main__":
    sys.exit(main())
