"""Host logic of the command line (scripts/gpu/main_gpu.cu:85-133) and of the
snapshot comparer (visualization/compare_mse.py); no GPU needed."""
from __future__ import annotations

import io
import os

import numpy as np
import pytest


def test_run_options_defaults_and_flags():
    from large_sys_eq_opt_amd.cli import parse_run_options

    o = parse_run_options([])
    assert (o.resolution_scale, o.write_output, o.num_snapshots, o.case_name, o.solver_name) == \
        (1, False, 5, "shock_bubble", "hll")   # main_gpu.cu:25-34
    o = parse_run_options(["4", "--case", "blast_wave", "--solver", "HLLC", "--output"])
    assert (o.resolution_scale, o.write_output, o.num_snapshots, o.case_name, o.solver_name) == \
        (4, True, 5, "blast_wave", "hllc")
    o = parse_run_options(["2", "--output", "--timing-only"])
    assert not o.write_output and o.num_snapshots == 0
    o = parse_run_options(["1", "--no-output", "--device-loop", "--precision", "17", "--outdir", "x"])
    assert o.device_loop and o.precision == 17 and o.outdir == "x"
    o = parse_run_options(["1", "--device-loop", "--output"])
    assert o.device_loop and o.write_output and o.num_snapshots == 5


@pytest.mark.parametrize("argv", [["0"], ["-3"], ["two"], ["1", "--case"], ["1", "--case", "nope"],
                                  ["1", "--solver", "roe"], ["1", "--solver"], ["1", "--fast"],
                                  ["1", "--precision", "30"]])
def test_run_options_errors(argv):
    from large_sys_eq_opt_amd.cli import main, parse_run_options

    with pytest.raises(ValueError):
        parse_run_options(argv)
    out, err = io.StringIO(), io.StringIO()
    assert main(argv, out=out, err=err) == 1
    assert err.getvalue().startswith("[GPU] Error: ") and out.getvalue() == ""


def _write(path, U, cfg):
    from large_sys_eq_opt_amd.driver import write_grid_csv

    write_grid_csv(U, cfg, str(path), precision=17)


def _cfg(nx, ny):
    from large_sys_eq_opt_amd.cases import CaseConfig

    return CaseConfig(nx, ny, 2, 0.0, 1.0, 0.0, 2.0, 0.8, 0.1)


def test_compare_pair_mse_and_linf(tmp_path):
    from large_sys_eq_opt_amd.compare import compare_pair

    rng = np.random.default_rng(0)
    nx, ny = 9, 7
    U = rng.uniform(0.5, 2.0, (ny + 4, nx + 4, 4))
    V = U.copy()
    V[3, 5, 2] += 0.25        # one cell of rho*v
    V[6, 2, 0] -= 0.5         # one cell of rho
    cfg = _cfg(nx, ny)
    _write(tmp_path / "a.csv", U, cfg)
    _write(tmp_path / "b.csv", V, cfg)
    r = compare_pair(str(tmp_path / "a.csv"), str(tmp_path / "b.csv"))
    n = nx * ny
    assert r["mse_rhov"] == pytest.approx(0.0625 / n, rel=1e-12)
    assert r["mse_rho"] == pytest.approx(0.25 / n, rel=1e-12)
    assert r["mse_rhou"] == 0.0 and r["mse_E"] == 0.0
    assert r["mse_all_fields"] == pytest.approx((0.0625 + 0.25) / (4 * n), rel=1e-12)
    assert r["linf_rho"] == pytest.approx(0.5) and r["linf_rhov"] == pytest.approx(0.25)
    assert r["linf_all_fields"] == pytest.approx(0.5)
    same = compare_pair(str(tmp_path / "a.csv"), str(tmp_path / "a.csv"))
    assert all(v == 0.0 for v in same.values())


def test_compare_rejects_mismatched_grids(tmp_path):
    from large_sys_eq_opt_amd.compare import compare_pair

    _write(tmp_path / "a.csv", np.ones((11, 13, 4)), _cfg(9, 7))
    _write(tmp_path / "b.csv", np.ones((11, 12, 4)), _cfg(8, 7))
    with pytest.raises(ValueError):
        compare_pair(str(tmp_path / "a.csv"), str(tmp_path / "b.csv"))


def test_compare_main_writes_reference_columns(tmp_path):
    """Snapshot pairing by id, a missing snapshot reported, and the summary
    file with the columns of the reference's mse_results/mse_summary_all_cases.csv
    (case,solver,num_snapshots,mean_mse_all_fields,mean_mse_rho,...,mean_mse_E)."""
    from large_sys_eq_opt_amd.compare import main
    from large_sys_eq_opt_amd.driver import snapshot_name

    nx, ny = 6, 5
    cfg = _cfg(nx, ny)
    out = tmp_path / "outputs"
    for k in (1, 2, 3):
        U = np.full((ny + 4, nx + 4, 4), float(k))
        _write(snapshot_name("cpu", "blast_wave", "hllc", 1, k, str(out)), U, cfg)
        if k != 3:
            U[4, 4, 3] += 1e-3 * k
            _write(snapshot_name("gpu", "blast_wave", "hllc", 1, k, str(out)), U, cfg)
    lines = []
    res = tmp_path / "mse"
    assert main(["--outdir", str(out), "--results", str(res), "--cases", "blast_wave", "--solvers", "hllc"],
                log=lines.append) == 0
    assert any("missing gpu snapshots [3]" in s for s in lines)
    with open(res / "mse_summary_all_cases.csv") as f:
        head = f.readline().strip().split(",")
        row = f.readline().strip().split(",")
    assert head[:8] == ["case", "solver", "num_snapshots", "mean_mse_all_fields", "mean_mse_rho",
                        "mean_mse_rhou", "mean_mse_rhov", "mean_mse_E"]
    d = dict(zip(head, row))
    n = nx * ny
    assert int(d["num_snapshots"]) == 2
    assert float(d["mean_mse_E"]) == pytest.approx(((1e-3) ** 2 + (2e-3) ** 2) / 2 / n, rel=1e-9)
    assert float(d["linf_E"]) == pytest.approx(2e-3, rel=1e-9)
    assert os.path.exists(res / "mse_detail_blast_wave_hllc.csv")
    assert main(["--outdir", str(out), "--results", str(res), "--cases", "shock_bubble"], log=lines.append) == 1
