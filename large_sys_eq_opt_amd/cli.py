"""Command line of the HIP solver with the arguments, defaults and messages of
the reference's GPU driver (scripts/gpu/main_gpu.cu:25-34 options, :85-133
parsing, :175-315 run):

    python -m large_sys_eq_opt_amd.cli [resolution_scale] [--output | --no-output | --timing-only]
                                       [--case NAME] [--solver hll|hllc|exact|force]

resolution_scale (default 1) multiplies the case's grid; --output writes five
CSV snapshots (outputs/gpu_<case>_<solver>_n<scale>_snapshot_<k>.csv) at the
first steps with t >= k t_end / 5; the defaults are the shock-bubble case and
the HLL solver.  Errors print "[GPU] Error: ..." and exit with status 1.

Options with no reference counterpart: --device-loop keeps dt and t on the
device (no per-step host round trip; the clock holds at each snapshot time
and the snapshot travels to the host while the next steps run),
--precision N writes N significant digits (17: round-trip exact), --outdir
DIR writes the snapshots there.
"""
from __future__ import annotations

import sys
from dataclasses import dataclass

from . import cases as _cases
from .solver import parse_riemann_solver, riemann_solver_to_string


@dataclass
class RunOptions:   # scripts/gpu/main_gpu.cu:25-34
    resolution_scale: int = 1
    write_output: bool = False
    num_snapshots: int = 5
    case_name: str = "shock_bubble"
    solver_name: str = "hll"
    device_loop: bool = False
    precision: int = 6
    outdir: str = "outputs"


def parse_run_options(argv) -> RunOptions:
    """scripts/gpu/main_gpu.cu:85-133 (argv without the program name)."""
    opts = RunOptions()
    argv = list(argv)
    if len(argv) >= 1:
        try:
            opts.resolution_scale = int(argv[0])
        except ValueError:
            raise ValueError(f"resolution_scale must be an integer, got {argv[0]!r}") from None
    i = 1
    while i < len(argv):
        arg = argv[i]
        if arg == "--output":
            opts.write_output = True
            opts.num_snapshots = 5
        elif arg in ("--no-output", "--timing-only"):
            opts.write_output = False
            opts.num_snapshots = 0
        elif arg == "--case":
            if i + 1 >= len(argv):
                raise ValueError("--case requires a value, e.g. shock_bubble or blast_wave.")
            i += 1
            opts.case_name = argv[i]
            _cases.parse_case_id(opts.case_name)   # validated early, as the reference does
        elif arg == "--solver":
            if i + 1 >= len(argv):
                raise ValueError("--solver requires a value: hll, hllc, exact, or force.")
            i += 1
            opts.solver_name = riemann_solver_to_string(parse_riemann_solver(argv[i]))
        elif arg == "--device-loop":
            opts.device_loop = True
        elif arg in ("--precision", "--outdir"):
            if i + 1 >= len(argv):
                raise ValueError(f"{arg} requires a value.")
            i += 1
            if arg == "--precision":
                opts.precision = int(argv[i])
                if not 1 <= opts.precision <= 17:
                    raise ValueError("--precision must be between 1 and 17.")
            else:
                opts.outdir = argv[i]
        else:
            raise ValueError(f"Unknown argument: {arg}")
        i += 1
    if opts.resolution_scale <= 0:
        raise ValueError("resolution_scale must be positive.")
    return opts


def _g(v: float) -> str:   # std::ostream's default formatting of a double
    return format(float(v), ".6g")


def main(argv=None, out=None, err=None) -> int:
    """Runs the reference driver's loop on the GPU; returns the exit status."""
    from .driver import run_case

    out = out or sys.stdout
    err = err or sys.stderr
    say = lambda s: print(s, file=out, flush=True)   # noqa: E731
    try:
        opts = parse_run_options(sys.argv[1:] if argv is None else argv)
        cfg = _cases.get_n_case_config(_cases.parse_case_id(opts.case_name), opts.resolution_scale)
        say(f"[GPU] case_name        : {opts.case_name}")
        say(f"[GPU] riemann_solver   : {opts.solver_name}")
        say(f"[GPU] resolution_scale : {opts.resolution_scale}")
        say(f"[GPU] nx               : {cfg.nx}")
        say(f"[GPU] ny               : {cfg.ny}")
        say(f"[GPU] total_cells      : {cfg.nx * cfg.ny}")
        say(f"[GPU] write_output     : {'true' if opts.write_output else 'false'}")
        say(f"[GPU] num_snapshots    : {opts.num_snapshots if opts.write_output else 0}")
        nsnap = opts.num_snapshots if opts.write_output else 0
        r = run_case(opts.case_name, opts.solver_name, opts.resolution_scale,
                     mode="device" if opts.device_loop else "host", num_snapshots=nsnap,
                     write_output=opts.write_output, outdir=opts.outdir, precision=opts.precision, log=say)
        say("[GPU] Finished.")
        say(f"[GPU] Final time = {_g(r['t'])}")
        say(f"[GPU] Total steps = {r['steps']}")
        if not opts.write_output:
            say("[GPU] Output disabled. No files were written.")
        return 0
    except Exception as e:   # the reference catches std::exception and returns 1
        print(f"[GPU] Error: {e}", file=err, flush=True)
        return 1


if __name__ == "__main__":
    sys.exit(main())
