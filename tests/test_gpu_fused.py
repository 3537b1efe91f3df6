"""The fused split step (one pass per step: an x wavefront feeding a y
wavefront through LDS) gives the x sweep followed by the y sweep bit for bit.

Device-loop runs against the oracle (the launch shape's fused = 2 takes the
fused step below the default's 2^20-cell threshold): ragged heights and chunk
sizes (partial last chunks, chunks shorter than the 32-row window), widths
that end in a partial 64-column strip (1, 2, 62 and 63 columns, and grids
narrower than one strip), states on both sides of the
fast-division range (tiles redone by the IEEE pass), tiny momenta (the HLL /
HLLC numerators that take the IEEE division), a state so cold that every
tile is redone, and inflow boundaries; the stepper's state alternates
between its two grids and has to come back into the caller's grid.  At the
benchmark size the fused steps are compared with the two-sweep steps.
"""
from __future__ import annotations

import numpy as np
import pytest

from conftest import interior, rel_linf, same_bits
from test_gpu_fast_range import scaled_field, tiny_momentum_field

pytestmark = pytest.mark.gpu

BITWISE_SOLVERS = [0, 1, 3]


def _run(hip, cfg, solver, U0, steps, shape, bc=None):
    from large_sys_eq_opt_amd import Grid2DGPU, Stepper
    from large_sys_eq_opt_amd.solver import get_launch_shape, set_launch_shape

    saved = get_launch_shape()
    try:
        set_launch_shape(**shape)
        g = Grid2DGPU.from_config(cfg)
        g.upload_from_aos(U0)
        st = Stepper(g, solver, cfg.cfl, 0.0, bc, log_capacity=steps)
        st.advance(steps)
        st.finish()
        _, n, dts = st.clock()
        st.close()
        out = g.download_to_aos()
        g.close()
    finally:
        set_launch_shape(**saved)
    return n, dts, out


def _check(o, n, dts, out, steps):
    assert n == steps
    assert same_bits(dts, o["dt"])
    assert same_bits(interior(out), interior(o["U"]))


@pytest.mark.parametrize("solver", BITWISE_SOLVERS)
@pytest.mark.parametrize("nx,ny,seg", [(64, 40, 16), (192, 150, 37), (320, 257, 128), (128, 300, 0)])
def test_fused_mixed_range_bitwise(hip, oracle, solver, nx, ny, seg):
    """Eleven steps (an odd count: the state ends in the stepper's second
    grid and is swapped back) of the two-blob field, part of it outside the
    fast range."""
    from large_sys_eq_opt_amd import cases

    steps = 11
    cfg = cases.CaseConfig(nx, ny, 2, 0.0, 1.0, 0.0, 1.0, 0.8, 0.0)
    U0 = scaled_field(oracle, nx, ny, 7)
    o = oracle.run(U0, cfg.dx, cfg.dy, solver, cfg.cfl, max_steps=steps)
    n, dts, out = _run(hip, cfg, solver, U0, steps, dict(fused=2, seg_y=seg, min_tiles=0))
    _check(o, n, dts, out, steps)


@pytest.mark.parametrize("solver", BITWISE_SOLVERS)
@pytest.mark.parametrize("nx,ny,seg", [(65, 40, 16), (66, 33, 0), (126, 70, 0), (127, 90, 37), (300, 131, 64),
                                       (30, 50, 0), (1, 20, 0), (63, 64, 32)])
def test_fused_partial_strips_bitwise(hip, oracle, solver, nx, ny, seg):
    """Widths that are not a multiple of 64: the last strip's lanes past the
    grid's edge hold the ghost cells and compute on copies of them; seven
    steps of the two-blob field."""
    from large_sys_eq_opt_amd import cases

    steps = 7
    cfg = cases.CaseConfig(nx, ny, 2, 0.0, 1.0, 0.0, 1.0, 0.8, 0.0)
    U0 = scaled_field(oracle, nx, ny, 11)
    o = oracle.run(U0, cfg.dx, cfg.dy, solver, cfg.cfl, max_steps=steps)
    n, dts, out = _run(hip, cfg, solver, U0, steps, dict(fused=2, seg_y=seg, min_tiles=0))
    _check(o, n, dts, out, steps)


@pytest.mark.parametrize("solver", BITWISE_SOLVERS)
@pytest.mark.parametrize("nx,ny,seg,min_tiles,tail", [(192, 150, 37, 8, 4), (320, 257, 64, 40, 16), (130, 300, 0, 8, 7),
                                                      (64, 96, 48, 2, 1), (200, 64, 16, 16, 16)])
def test_fused_tail_chunks_bitwise(hip, oracle, solver, nx, ny, seg, min_tiles, tail):
    """The last rows of every strip in short chunks dispatched last (the
    launch shape's fused_tail; min_tiles sets how many): ragged big and tail
    chunks, a tail of one-row chunks, and a tail as long as the big chunks
    (none taken); nine steps of the two-blob field."""
    from large_sys_eq_opt_amd import cases

    steps = 9
    cfg = cases.CaseConfig(nx, ny, 2, 0.0, 1.0, 0.0, 1.0, 0.8, 0.0)
    U0 = scaled_field(oracle, nx, ny, 13)
    o = oracle.run(U0, cfg.dx, cfg.dy, solver, cfg.cfl, max_steps=steps)
    n, dts, out = _run(hip, cfg, solver, U0, steps, dict(fused=2, seg_y=seg, min_tiles=min_tiles, fused_tail=tail))
    _check(o, n, dts, out, steps)


@pytest.mark.parametrize("solver", BITWISE_SOLVERS)
def test_fused_blast_bitwise(hip, oracle, solver):
    """The cylindrical blast on 256x256 ([0,2]^2), 30 steps, default shape."""
    from large_sys_eq_opt_amd import cases

    steps = 30
    cfg = cases.CaseConfig(256, 256, 2, 0.0, 2.0, 0.0, 2.0, 0.8, 0.0)
    kind, params = cases.case_params("cyl_blast")
    U0 = oracle.init(kind, params.array(), 256, 256, cfg.bounds)
    o = oracle.run(U0, cfg.dx, cfg.dy, solver, cfg.cfl, max_steps=steps)
    n, dts, out = _run(hip, cfg, solver, U0, steps, dict(fused=2))
    _check(o, n, dts, out, steps)


@pytest.mark.parametrize("solver", BITWISE_SOLVERS)
def test_fused_tiny_momenta_bitwise(hip, oracle, solver):
    from large_sys_eq_opt_amd import cases

    nx, ny, steps = 320, 260, 12
    cfg = cases.CaseConfig(nx, ny, 2, 0.0, 1.0, 0.0, 1.0, 0.8, 0.0)
    U0 = tiny_momentum_field(oracle, nx, ny, 5)
    o = oracle.run(U0, cfg.dx, cfg.dy, solver, cfg.cfl, max_steps=steps)
    n, dts, out = _run(hip, cfg, solver, U0, steps, dict(fused=2))
    _check(o, n, dts, out, steps)


def test_fused_every_tile_redone(hip, oracle):
    """A cold, nearly still gas (p = 2e-13, below the fast range and the
    reconstruction's floor; |u|, |v| ~ 1e-8): dt/dx exceeds 2^20, so every
    tile of every step goes to the redo list without a fast pass."""
    from large_sys_eq_opt_amd import cases

    nx, ny, steps = 192, 130, 5
    cfg = cases.CaseConfig(nx, ny, 2, 0.0, 1.0, 0.0, 1.0, 0.8, 0.0)
    rng = np.random.default_rng(3)
    V = np.empty((ny + 4, nx + 4, 4))
    V[..., 0] = 1.0 + 0.2 * rng.random((ny + 4, nx + 4))
    V[..., 1] = 1e-8 * rng.standard_normal((ny + 4, nx + 4))
    V[..., 2] = 1e-8 * rng.standard_normal((ny + 4, nx + 4))
    V[..., 3] = 2e-13 * (1.0 + 0.1 * rng.random((ny + 4, nx + 4)))
    U0 = oracle.prim_to_cons(V.reshape(-1, 4)).reshape(ny + 4, nx + 4, 4)
    o = oracle.run(U0, cfg.dx, cfg.dy, 1, cfg.cfl, max_steps=steps)
    assert o["dt"][0] / cfg.dx > 2.0 ** 20
    n, dts, out = _run(hip, cfg, 1, U0, steps, dict(fused=2))
    _check(o, n, dts, out, steps)


@pytest.mark.parametrize("solver", BITWISE_SOLVERS)
def test_fused_inflow_bitwise(hip, oracle, solver):
    """The shock-bubble case (inflow at the left) on 256x128, 25 steps:
    the intermediate state's ghost rows and the inflow columns."""
    from large_sys_eq_opt_amd import cases

    steps, case = 25, "shock_bubble_nd"
    cfg = cases.get_case_config(case)
    kind, params = cases.case_params(case)
    nx, ny = 256, 128
    c = cases.CaseConfig(nx, ny, 2, cfg.x_min, cfg.x_max, cfg.y_min, cfg.y_max, cfg.cfl, 0.0)
    U0 = oracle.init(kind, params.array(), nx, ny, c.bounds)
    bc = cases.case_boundary(case)
    o = oracle.run(U0, c.dx, c.dy, solver, c.cfl, max_steps=steps, inflow_left=bc.left_inflow_state)
    n, dts, out = _run(hip, c, solver, U0, steps, dict(fused=2, seg_y=48, min_tiles=0), bc=bc)
    _check(o, n, dts, out, steps)


def test_fused_inflow_partial_strip_bitwise(hip, oracle):
    """The shock-bubble case on 200x100 (a last strip of 8 columns), HLLC,
    20 steps."""
    from large_sys_eq_opt_amd import cases

    steps, case = 20, "shock_bubble_nd"
    cfg = cases.get_case_config(case)
    kind, params = cases.case_params(case)
    nx, ny = 200, 100
    c = cases.CaseConfig(nx, ny, 2, cfg.x_min, cfg.x_max, cfg.y_min, cfg.y_max, cfg.cfl, 0.0)
    U0 = oracle.init(kind, params.array(), nx, ny, c.bounds)
    bc = cases.case_boundary(case)
    o = oracle.run(U0, c.dx, c.dy, 1, c.cfl, max_steps=steps, inflow_left=bc.left_inflow_state)
    n, dts, out = _run(hip, c, 1, U0, steps, dict(fused=2, seg_y=0, min_tiles=0), bc=bc)
    _check(o, n, dts, out, steps)


def test_fused_equals_sweeps_8192(hip):
    """Three HLLC steps of the 8192^2 blast: fused and two-sweep steppers
    give the same dt sequence and state bit for bit."""
    from large_sys_eq_opt_amd import Grid2DGPU, Stepper, cases
    from large_sys_eq_opt_amd.solver import get_launch_shape, set_launch_shape

    n, steps = 8192, 3
    cfg = cases.get_n_case_config("cyl_blast", n // 256)
    kind, params = cases.case_params("cyl_blast")
    res = []
    saved = get_launch_shape()
    try:
        for fused in (1, 0):
            set_launch_shape(fused=fused)
            g = Grid2DGPU.from_config(cfg)
            g.init_state(kind, params.array())
            st = Stepper(g, 1, cfg.cfl, 0.0, cases.case_boundary("cyl_blast"), log_capacity=steps)
            st.advance(steps)
            st.finish()
            _, k, dts = st.clock()
            st.close()
            assert k == steps
            res.append((dts, g.download_to_aos()[::7, ::5]))
            g.close()
    finally:
        set_launch_shape(**saved)
    assert same_bits(res[0][0], res[1][0])
    assert same_bits(res[0][1], res[1][1])


EXACT = 2


def _exact_case(name, nx, ny):
    from large_sys_eq_opt_amd import cases

    cfg = cases.get_case_config(name)
    kind, params = cases.case_params(name)
    c = cases.CaseConfig(nx, ny, 2, cfg.x_min, cfg.x_max, cfg.y_min, cfg.y_max, cfg.cfl, 0.0)
    return c, kind, params, cases.case_boundary(name)


@pytest.mark.parametrize("case, nx, ny, steps, shape", [
    ("cyl_blast", 256, 256, 30, dict(seg_y=0, min_tiles=8192)),
    ("cyl_blast", 200, 131, 20, dict(seg_y=37, min_tiles=16, fused_tail=5)),
    ("shock_bubble_nd", 256, 128, 25, dict(seg_y=48, min_tiles=0)),
    ("shock_bubble_nd", 130, 90, 25, dict(seg_y=0, min_tiles=8, fused_tail=4)),
])
def test_fused_exact_equals_sweeps(hip, oracle, case, nx, ny, steps, shape):
    """The exact solver's fused step: its fast pass solves the wavefronts
    whose faces all join equal states (the closed form) and leaves every
    tile with another face to the redo pass (the full solver).  The state
    and dt sequence equal the two-sweep steps (whose y sweep takes the same
    fast pass and redo list) bit for bit, and the oracle to the exact
    solver's tolerance (1e-10 relative state, 1e-12 relative dt: the device
    pow); ragged chunks, tail chunks, partial strips and inflow."""
    cfg, kind, params, bc = _exact_case(case, nx, ny)
    U0 = oracle.init(kind, params.array(), nx, ny, cfg.bounds)
    inflow = bc.left_inflow_state if case == "shock_bubble_nd" else None
    o = oracle.run(U0, cfg.dx, cfg.dy, EXACT, cfg.cfl, max_steps=steps, inflow_left=inflow)
    n1, d1, u1 = _run(hip, cfg, EXACT, U0, steps, dict(fused=2, **shape), bc=bc)
    n0, d0, u0 = _run(hip, cfg, EXACT, U0, steps, dict(fused=0, **shape), bc=bc)
    assert n1 == n0 == steps
    assert same_bits(d1, d0)
    assert same_bits(interior(u1), interior(u0))
    assert np.max(np.abs(d1 - o["dt"]) / o["dt"]) <= 1e-12
    assert rel_linf(u1, o["U"]) <= 1e-10


def test_fused_exact_mixed_range(hip, oracle):
    """The two-blob field (part of it outside the fast range, where the
    whole tile goes to the redo pass) under the exact solver: fused = two
    sweeps bit for bit, oracle to 1e-10.  The redo pass runs each piece
    with the fast division sequences first; the pieces that reach the
    out-of-range blob fail that verdict and are solved again with the IEEE
    sequences, the others keep the first result."""
    from large_sys_eq_opt_amd import cases

    nx, ny, steps = 192, 150, 9
    cfg = cases.CaseConfig(nx, ny, 2, 0.0, 1.0, 0.0, 1.0, 0.8, 0.0)
    U0 = scaled_field(oracle, nx, ny, 7)
    o = oracle.run(U0, cfg.dx, cfg.dy, EXACT, cfg.cfl, max_steps=steps)
    n1, d1, u1 = _run(hip, cfg, EXACT, U0, steps, dict(fused=2, seg_y=37, min_tiles=0))
    n0, d0, u0 = _run(hip, cfg, EXACT, U0, steps, dict(fused=0, seg_y=37, min_tiles=0))
    assert n1 == n0 == steps
    assert same_bits(d1, d0) and same_bits(interior(u1), interior(u0))
    assert rel_linf(u1, o["U"]) <= 1e-10


def test_fused_exact_8192_equals_sweeps(hip):
    """Two exact steps of the 8192^2 blast (the benchmark grid; fused by
    default there): fused and two-sweep steppers agree bit for bit."""
    from large_sys_eq_opt_amd import Grid2DGPU, Stepper, cases
    from large_sys_eq_opt_amd.solver import get_launch_shape, set_launch_shape

    n, steps = 8192, 2
    cfg = cases.get_n_case_config("cyl_blast", n // 256)
    kind, params = cases.case_params("cyl_blast")
    res = []
    saved = get_launch_shape()
    try:
        for fused in (1, 0):
            set_launch_shape(fused=fused)
            g = Grid2DGPU.from_config(cfg)
            g.init_state(kind, params.array())
            st = Stepper(g, EXACT, cfg.cfl, 0.0, cases.case_boundary("cyl_blast"), log_capacity=steps)
            st.advance(steps)
            st.finish()
            _, k, dts = st.clock()
            st.close()
            assert k == steps
            res.append((dts, g.download_to_aos()[::7, ::5]))
            g.close()
    finally:
        set_launch_shape(**saved)
    assert same_bits(res[0][0], res[1][0])
    assert same_bits(res[0][1], res[1][1])


def _smooth_state(nx, ny):
    """A ghosted AoS state with no two equal neighbours (smooth density,
    velocity and pressure waves): no wavefront of the exact solver's fast
    pass is uniform."""
    y, x = np.mgrid[0:ny + 4, 0:nx + 4] + 0.5
    x, y = x / (nx + 4), y / (ny + 4)
    rho = 1.0 + 0.3 * np.sin(2 * np.pi * x + 0.3) * np.cos(2 * np.pi * y + 1.1)
    u = 0.3 * np.sin(2 * np.pi * (y + 0.37 * x) + 0.7)
    v = 0.2 * np.cos(2 * np.pi * (x - 0.41 * y) + 2.0)
    p = 1.0 + 0.4 * np.cos(2 * np.pi * (x + 0.83 * y) + 0.4)
    U = np.empty((ny + 4, nx + 4, 4))
    U[..., 0], U[..., 1], U[..., 2] = rho, rho * u, rho * v
    U[..., 3] = p / 0.4 + 0.5 * rho * (u * u + v * v)
    return U


def _cold_state(oracle, nx, ny, seed):
    """A cold, nearly still gas below the fast range (every tile redone)."""
    rng = np.random.default_rng(seed)
    V = np.empty((ny + 4, nx + 4, 4))
    V[..., 0] = 1.0 + 0.2 * rng.random((ny + 4, nx + 4))
    V[..., 1] = 1e-8 * rng.standard_normal((ny + 4, nx + 4))
    V[..., 2] = 1e-8 * rng.standard_normal((ny + 4, nx + 4))
    V[..., 3] = 2e-13 * (1.0 + 0.1 * rng.random((ny + 4, nx + 4)))
    return oracle.prim_to_cons(V.reshape(-1, 4)).reshape(ny + 4, nx + 4, 4)


@pytest.mark.parametrize("solver", [1, EXACT])
def test_many_tiles_redone(hip, oracle, solver):
    """A redo list long enough that each redone tile is split into fewer
    pieces than its rows allow (2560x2048: 640 fused tiles of 128 rows, 640
    y tiles of 128 rows with the redo list): every tile goes to the redo
    pass (HLLC: a cold gas below the fast range; exact: a state with no
    uniform wavefront), and every row of every tile must be redone.  Three
    steps, fused and two sweeps, against the oracle."""
    from large_sys_eq_opt_amd import cases

    nx, ny, steps = 2560, 2048, 3
    cfg = cases.CaseConfig(nx, ny, 2, 0.0, 1.0, 0.0, 1.0, 0.8, 0.0)
    U0 = _cold_state(oracle, nx, ny, 9) if solver == 1 else _smooth_state(nx, ny)
    o = oracle.run(U0, cfg.dx, cfg.dy, solver, cfg.cfl, max_steps=steps)
    for shape in (dict(fused=2, seg_y=128), dict(fused=0, seg_y=128, min_tiles=0)):
        n, dts, out = _run(hip, cfg, solver, U0, steps, shape)
        assert n == steps, shape
        if solver == EXACT:
            assert np.max(np.abs(dts - o["dt"]) / o["dt"]) <= 1e-12, shape
            assert rel_linf(out, o["U"]) <= 1e-10, shape
        else:
            assert same_bits(dts, o["dt"]), shape
            assert same_bits(interior(out), interior(o["U"])), shape


@pytest.mark.parametrize("state", ["smooth", "blast"])
def test_exact_stepper_changes_form(hip, oracle, state):
    """Under the default launch shape an exact-solver stepper on 4096^2
    starts on the fused step and, at a clock read after 10 steps, changes to
    the two sweeps where its fast passes left most tiles to the redo pass (a
    state with no uniform wavefront), and stays on the fused step on the
    cylindrical blast; 12 steps equal a stepper held to the two sweeps bit
    for bit either way."""
    from large_sys_eq_opt_amd import Grid2DGPU, Stepper, cases
    from large_sys_eq_opt_amd.solver import get_launch_shape, set_launch_shape

    n = 4096
    cfg = cases.get_n_case_config("cyl_blast", n // 256)
    kind, params = cases.case_params("cyl_blast")
    bc = cases.case_boundary("cyl_blast")

    def start(g):
        if state == "smooth":
            g.upload_from_aos(_smooth_state(n, n))
        else:
            g.init_state(kind, params.array())

    saved = get_launch_shape()
    try:
        set_launch_shape(fused=1)
        g = Grid2DGPU.from_config(cfg)
        start(g)
        st = Stepper(g, EXACT, cfg.cfl, 0.0, bc, log_capacity=16)
        st.advance(10)
        st.clock()
        _, k = st.advance_profiled(2)
        _, steps, dts = st.clock()
        st.close()
        assert steps == 12
        assert (k[2] > 0.0) == (state == "smooth"), k   # the y sweep ran: the two sweeps
        out = g.download_to_aos()[::5, ::3]
        g.close()
        set_launch_shape(fused=0)
        g = Grid2DGPU.from_config(cfg)
        start(g)
        st = Stepper(g, EXACT, cfg.cfl, 0.0, bc, log_capacity=16)
        st.advance(12)
        _, steps0, dts0 = st.clock()
        st.close()
        ref = g.download_to_aos()[::5, ::3]
        g.close()
    finally:
        set_launch_shape(**saved)
    assert steps0 == 12
    assert same_bits(dts, dts0)
    assert same_bits(out, ref)
