"""The command line (scripts/gpu/main_gpu.cu) end to end on the GPU: its
messages, its five snapshot files, and those snapshots against the oracle's
through the snapshot comparer (zero MSE and zero largest difference: HLLC is
bit-identical, the files carry 17 significant digits)."""
from __future__ import annotations

import io

import pytest

pytestmark = pytest.mark.gpu


def test_cli_run_matches_oracle_snapshots(hip, oracle, tmp_path):
    from large_sys_eq_opt_amd import cases
    from large_sys_eq_opt_amd.cli import main
    from large_sys_eq_opt_amd.compare import main as compare_main
    from large_sys_eq_opt_amd.driver import snapshot_name, write_grid_csv

    out = tmp_path / "outputs"
    buf = io.StringIO()
    rc = main(["1", "--case", "cyl_blast", "--solver", "hllc", "--output", "--outdir", str(out), "--precision", "17"],
              out=buf)
    text = buf.getvalue().splitlines()
    assert rc == 0, text
    cfg = cases.get_n_case_config("cyl_blast", 1)
    assert text[:8] == ["[GPU] case_name        : cyl_blast", "[GPU] riemann_solver   : hllc",
                        "[GPU] resolution_scale : 1", f"[GPU] nx               : {cfg.nx}",
                        f"[GPU] ny               : {cfg.ny}", f"[GPU] total_cells      : {cfg.nx * cfg.ny}",
                        "[GPU] write_output     : true", "[GPU] num_snapshots    : 5"]
    assert sum(s.startswith("[GPU] Wrote snapshot ") for s in text) == 5
    kind, params = cases.case_params("cyl_blast")
    U0 = oracle.init(kind, params.array(), cfg.nx, cfg.ny, cfg.bounds)
    o = oracle.run(U0, cfg.dx, cfg.dy, 1, cfg.cfl, t_end=cfg.t_end, nsnap=5, keep_snapshots=True)
    assert text[-3:] == ["[GPU] Finished.", f"[GPU] Final time = {format(o['t'], '.6g')}",
                         f"[GPU] Total steps = {o['steps']}"]
    for k in range(5):
        write_grid_csv(o["snapshots"][k], cfg, snapshot_name("cpu", "cyl_blast", "hllc", 1, k + 1, str(out)), 17)
    res = tmp_path / "mse"
    assert compare_main(["--outdir", str(out), "--results", str(res), "--cases", "cyl_blast", "--solvers", "hllc"],
                        log=lambda s: None) == 0
    with open(res / "mse_summary_all_cases.csv") as f:
        head = f.readline().strip().split(",")
        row = dict(zip(head, f.readline().strip().split(",")))
    assert int(row["num_snapshots"]) == 5
    assert float(row["mean_mse_all_fields"]) == 0.0 and float(row["linf_all_fields"]) == 0.0
