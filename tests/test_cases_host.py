"""Host-side case setup, solver selection and snapshot I/O (no GPU)."""
from __future__ import annotations

import numpy as np
import pytest

from conftest import same_bits

from large_sys_eq_opt_amd import cases
from large_sys_eq_opt_amd.driver import read_grid_csv, snapshot_name, write_grid_csv
from large_sys_eq_opt_amd.solver import parse_riemann_solver, riemann_solver_to_string


def test_case_configs_match_reference():
    c = cases.get_case_config("shock_bubble")   # src/test_cases.cpp:134-145
    assert (c.nx, c.ny, c.ng, c.x_max, c.y_max, c.cfl, c.t_end) == (500, 197, 2, 0.225, 0.089, 0.4, 0.0011741)
    c = cases.get_case_config("blast_wave")     # src/test_cases.cpp:147-158
    assert (c.nx, c.ny, c.x_max, c.cfl, c.t_end) == (500, 500, 1.0, 0.4, 0.2)
    c = cases.get_n_case_config("cyl_blast", 32)
    assert (c.nx, c.ny, c.x_max, c.cfl) == (8192, 8192, 2.0, 0.8)
    with pytest.raises(ValueError):
        cases.get_n_case_config("blast_wave", 0)
    with pytest.raises(ValueError):
        cases.parse_case_id("no_such_case")


def test_post_shock_state():
    # Mach 1.22 into (rho, p) = (1, 1): values quoted for the nondimensional case
    rho2, u2, p2 = cases.post_shock_state(1.0, 1.0, 1.22)
    assert abs(rho2 - 1.376364) < 1e-6 and abs(u2 - 0.394729) < 1e-6 and abs(p2 - 1.5698) < 1e-4
    # Rankine-Hugoniot: mass and momentum flux continuity in the shock frame
    s = 1.22 * np.sqrt(1.4)
    assert abs(1.0 * s - rho2 * (s - u2)) < 1e-12
    assert abs(1.0 * s * s + 1.0 - (rho2 * (s - u2) ** 2 + p2)) < 1e-12


@pytest.mark.parametrize("name", ["shock_bubble", "blast_wave", "cyl_blast", "shock_bubble_nd"])
def test_host_initial_state_matches_oracle(oracle, name):
    cfg = cases.get_case_config(name)
    kind, params = cases.case_params(name)
    assert same_bits(cases.initialise_grid(name, cfg), oracle.init(kind, params.array(), cfg.nx, cfg.ny,
                                                                           cfg.bounds))


def test_reference_initial_grid(reference):
    for case_id, name in ((0, "shock_bubble"), (1, "blast_wave")):
        U, _ = reference.init_case(case_id, 1)
        mine, _ = cases.make_n_grid(name, 1)
        assert same_bits(U, mine)


def test_solver_names():
    assert parse_riemann_solver("hllc") == 1 and parse_riemann_solver("EXACT") == 2
    assert riemann_solver_to_string(3) == "force"
    with pytest.raises(ValueError):
        parse_riemann_solver("roe")


def test_snapshot_csv_format_and_roundtrip(tmp_path):
    cfg = cases.get_case_config("cyl_blast")
    cfg.nx, cfg.ny = 6, 4
    U = cases.initialise_grid("cyl_blast", cfg)
    fn = str(tmp_path / "s.csv")
    write_grid_csv(U, cfg, fn)
    lines = open(fn).read().splitlines()
    assert lines[0] == "i,j,x,y,rho,rhou,rhov,E"           # scripts/cpu/main_cpu.cpp:140
    assert lines[1].startswith("2,2,0.166667,0.25,")        # indices include ng; 6 significant digits
    assert len(lines) == 1 + 24
    exact = str(tmp_path / "e.csv")
    write_grid_csv(U, cfg, exact, precision=17)
    assert same_bits(read_grid_csv(exact, 6, 4), U[2:-2, 2:-2])
    assert snapshot_name("gpu", "blast_wave", "hllc", 2, 3) == "outputs/gpu_blast_wave_hllc_n2_snapshot_3.csv"


def _csv_rows_python(U, cfg, precision):
    """The snapshot rows as the reference's ostream prints them
    (scripts/cpu/main_cpu.cpp:136-164), formatted one cell at a time."""
    dx = (cfg.x_max - cfg.x_min) / float(cfg.nx)
    dy = (cfg.y_max - cfg.y_min) / float(cfg.ny)
    g = lambda v: format(float(v), f".{precision}g")  # noqa: E731
    rows = ["i,j,x,y,rho,rhou,rhov,E"]
    for j in range(2, 2 + cfg.ny):
        y = cfg.y_min + (float(j - 2) + 0.5) * dy
        for i in range(2, 2 + cfg.nx):
            x = cfg.x_min + (float(i - 2) + 0.5) * dx
            rows.append(",".join([str(i), str(j), g(x), g(y)] + [g(v) for v in U[j, i]]))
    return rows


@pytest.mark.parametrize("precision, nthreads", [(6, 3), (17, 0), (3, 1)])
def test_native_csv_writer_matches_per_cell_formatting(tmp_path, precision, nthreads):
    """The library's multi-threaded writer gives the same text as formatting
    every value with %.{p}g in cell order, on a grid whose rows span several
    formatting blocks and rounds, with values of every sign and magnitude
    (including -0, denormals and 1e300)."""
    import numpy as np

    from large_sys_eq_opt_amd.solver import write_csv_aos

    cfg = cases.CaseConfig(300, 700, 2, -1.5, 2.0, 0.25, 3.0, 0.8, 0.0)
    rng = np.random.default_rng(precision)
    U = rng.standard_normal((cfg.ny + 4, cfg.nx + 4, 4)) * 10.0 ** rng.integers(-300, 300, (cfg.ny + 4, cfg.nx + 4, 4))
    U[5, 5] = [-0.0, 0.0, 5e-324, -1e300]
    U[7, 9] = [1.0, -1.0, 123456789.0, 0.1]
    fn = str(tmp_path / "n.csv")
    write_csv_aos(U, cfg, fn, precision, nthreads)
    assert open(fn).read().splitlines() == _csv_rows_python(U, cfg, precision)


def test_ensemble_spec_is_the_committed_table():
    import json
    import os

    from large_sys_eq_opt_amd.ensemble import ensemble_spec

    path = os.path.join(os.path.dirname(__file__), "golden", "ensemble_512_seed0.json")
    assert ensemble_spec(512, 0) == json.load(open(path))
