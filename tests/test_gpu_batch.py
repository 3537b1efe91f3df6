"""Batched ensemble: each instance equals a single-grid run of the same
initial state bit for bit (and the oracle), under every schedule (the fused
mode 3 included: most instance sides end in a partial 64-column strip) and
when the steps replay from a captured hipGraph."""
from __future__ import annotations

import numpy as np
import pytest

from conftest import interior, same_bits

pytestmark = pytest.mark.gpu

SPEC = [{"side": s, "r0": r, "p_in": p} for s, r, p in
        [(64, 0.3, 10.0), (67, 0.45, 1.5), (100, 0.2, 80.0), (129, 0.35, 3.0), (31, 0.5, 50.0),
         (16, 0.4, 2.0), (90, 0.12, 99.0), (200, 0.25, 7.0)]]
STEPS = 12


def oracle_instance(oracle, s, solver, steps):
    from large_sys_eq_opt_amd.ensemble import blast_params

    n = s["side"]
    U0 = oracle.init(2, blast_params(s), n, n, (0.0, 2.0, 0.0, 2.0))
    return oracle.run(U0, 2.0 / n, 2.0 / n, solver, 0.8, max_steps=steps)


# (xseg, ychunk, mode, blocks_per_cu, xrows): persistent fused / two-launch,
# dispatched, ragged segments and chunks, several rows per x tile
@pytest.mark.parametrize("schedule", [(0, 256, 0, 0, 1), (0, 32, 1, 0, 1), (48, 0, 0, 1, 2), (0, 64, 2, 0, 1),
                                      (60, 16, 2, 0, 3), (0, 64, 2, 0, 8), (0, 128, 2, 0, 1000), (0, 96, 1, 0, 7),
                                      (0, 128, 3, 0, 1), (0, 37, 3, 0, 1), (0, 0, 3, 0, 1)])
def test_batch_instances_match_oracle(hip, oracle, schedule):
    from large_sys_eq_opt_amd.ensemble import Batch

    b = Batch(SPEC, solver=1)
    b.set_schedule(*schedule)
    b.initialise()
    b.step(STEPS)
    t, dt, steps, status = b.clock()
    assert (steps == STEPS).all() and (status == 0).all()
    for k, s in enumerate(SPEC):
        o = oracle_instance(oracle, s, 1, STEPS)
        U = b.download(k)
        assert same_bits(interior(U), interior(o["U"])), (k, s)
        assert dt[k] == o["dt"][-1] and t[k] == np.float64(sum_seq(o["dt"]))
    b.close()


@pytest.mark.parametrize("schedule, size_class", [((0, 64, 2, 0, 4), (100, 16, 3)), ((0, 256, 0, 0, 1), (1000, 0, 5)),
                                                  ((60, 128, 2, 0, 8), (50, 40, 0)), ((0, 64, 3, 0, 1), (100, 16, 0))])
def test_batch_size_class_matches_oracle(hip, oracle, schedule, size_class):
    """Instances below a side take their own y chunks and x tile rows (the
    autotuner's size classes): every instance still equals its own oracle
    run bit for bit."""
    from large_sys_eq_opt_amd.ensemble import Batch

    b = Batch(SPEC, solver=1)
    b.set_schedule(*schedule)
    b.set_size_class(*size_class)
    b.initialise()
    b.step(STEPS)
    t, dt, steps, status = b.clock()
    assert (steps == STEPS).all() and (status == 0).all()
    for k, s in enumerate(SPEC):
        o = oracle_instance(oracle, s, 1, STEPS)
        assert same_bits(interior(b.download(k)), interior(o["U"])), (k, s)
        assert dt[k] == o["dt"][-1] and t[k] == np.float64(sum_seq(o["dt"]))
    b.close()


def sum_seq(dts):
    t = 0.0
    for d in dts:
        t += d
    return t


@pytest.mark.parametrize("mode", [0, 3])
def test_batch_graph_replay_matches_direct_launches(hip, oracle, mode):
    from large_sys_eq_opt_amd.ensemble import Batch

    a = Batch(SPEC, solver=1)
    a.set_schedule(0, 64, mode, 0, 1)
    a.initialise()
    a.capture(4)
    a.step(STEPS)          # 3 graph replays of 4 steps
    b = Batch(SPEC, solver=1)
    b.set_schedule(0, 64, mode, 0, 1)
    b.initialise()
    b.step(STEPS)          # direct launches
    ta, _, sa, _ = a.clock()
    tb, _, sb, _ = b.clock()
    assert (sa == STEPS).all() and np.array_equal(ta, tb)
    for k in range(len(SPEC)):
        assert same_bits(interior(a.download(k)), interior(b.download(k)))
    a.close()
    b.close()


@pytest.mark.parametrize("mode", [0, 3])
@pytest.mark.parametrize("solver", [0, 2, 3])
def test_batch_other_solvers(hip, oracle, solver, mode):
    """HLL, exact (within 1e-10: its Newton iteration is not bitwise) and
    FORCE, under the persistent and the fused schedule."""
    from large_sys_eq_opt_amd.ensemble import Batch

    spec = SPEC[:4]
    b = Batch(spec, solver=solver)
    b.set_schedule(0, 64, mode, 0, 1)
    b.initialise()
    b.step(6)
    for k, s in enumerate(spec):
        o = oracle_instance(oracle, s, solver, 6)
        U = b.download(k)
        if solver == 2:
            assert np.max(np.abs(interior(U) - interior(o["U"]))) <= 1e-10 * np.max(np.abs(interior(o["U"])))
        else:
            assert same_bits(interior(U), interior(o["U"]))
    b.close()


def test_batch_ensemble_512_seed0_500_steps(hip, oracle):
    """The 512-instance ensemble from seed 0 (sides 64..1024, 95.3 M cells)
    advanced 500 HLLC steps, 25 steps per captured graph, under the
    dispatched schedule the benchmark's tuner picks and the persistent one:
    every instance takes 500 steps with status 0, and the smallest, the
    largest and six seeded random instances equal their own oracle runs bit
    for bit (state, last dt and t).  In 500 steps the blast waves of the
    smaller instances cross the transmissive edges."""
    from large_sys_eq_opt_amd.ensemble import Batch, ensemble_spec

    spec = ensemble_spec(512, 0)
    sides = np.array([s["side"] for s in spec])
    rng = np.random.default_rng(3)
    picks = sorted({int(np.argmin(sides)), int(np.argmax(sides)), *rng.choice(len(spec), 6, replace=False).tolist()})
    ref = {k: oracle_instance(oracle, spec[k], 1, 500) for k in picks}
    for schedule in [(0, 128, 2, 0, 4), (0, 256, 0, 0, 1), (0, 64, 3, 0, 1)]:
        b = Batch(spec, solver=1)
        b.set_schedule(*schedule)
        b.initialise()
        b.capture(25)
        b.step(500)
        t, dt, steps, status = b.clock()
        assert (steps == 500).all() and (status == 0).all(), schedule
        for k in picks:
            o = ref[k]
            assert same_bits(interior(b.download(k)), interior(o["U"])), (schedule, k, spec[k])
            assert dt[k] == o["dt"][-1] and t[k] == np.float64(sum_seq(o["dt"])), (schedule, k)
        b.close()


def test_batch_fused_mode_switches(hip, oracle):
    """Five fused steps (the instances' states end in their scratch grids),
    a switch to the persistent schedule (which puts them back), seven more
    steps, a switch back to the fused mode and three more: every instance
    equals its 15-step oracle run."""
    from large_sys_eq_opt_amd.ensemble import Batch

    b = Batch(SPEC, solver=1)
    b.set_schedule(0, 64, 3, 0, 1)
    b.initialise()
    b.step(5)
    b.set_schedule(0, 256, 0, 0, 1)
    b.step(7)
    b.set_schedule(0, 48, 3, 0, 1)
    b.step(3)
    t, dt, steps, status = b.clock()
    assert (steps == 15).all() and (status == 0).all()
    for k, s in enumerate(SPEC):
        o = oracle_instance(oracle, s, 1, 15)
        assert same_bits(interior(b.download(k)), interior(o["U"])), (k, s)
        assert dt[k] == o["dt"][-1] and t[k] == np.float64(sum_seq(o["dt"]))
    b.close()


def test_batch_fused_download_without_clock(hip, oracle):
    """download() after an odd number of fused steps, with no clock() call
    in between, reads the current state."""
    from large_sys_eq_opt_amd.ensemble import Batch

    b = Batch(SPEC[:3], solver=1)
    b.set_schedule(0, 64, 3, 0, 1)
    b.initialise()
    b.step(3)
    for k, s in enumerate(SPEC[:3]):
        o = oracle_instance(oracle, s, 1, 3)
        assert same_bits(interior(b.download(k)), interior(o["U"])), (k, s)
    b.close()


def test_batch_exact_fused_equals_dispatched(hip, oracle):
    """The exact solver under the fused mode (its fast pass solves the
    uniform wavefronts, the redo pass every tile with another face, in
    pieces) equals the dispatched two-sweep schedule bit for bit on all
    eight instances over 20 steps, and its own oracle runs to 1e-10."""
    from large_sys_eq_opt_amd.ensemble import Batch

    out = []
    for schedule in [(0, 64, 3, 0, 1), (0, 64, 2, 0, 4)]:
        b = Batch(SPEC, solver=2)
        b.set_schedule(*schedule)
        b.initialise()
        b.step(20)
        t, dt, steps, status = b.clock()
        assert (steps == 20).all() and (status == 0).all(), schedule
        out.append(([b.download(k) for k in range(len(SPEC))], t, dt))
        b.close()
    for k in range(len(SPEC)):
        assert same_bits(interior(out[0][0][k]), interior(out[1][0][k])), k
    assert same_bits(out[0][1], out[1][1]) and same_bits(out[0][2], out[1][2])
    for k in (0, 3, 7):
        o = oracle_instance(oracle, SPEC[k], 2, 20)
        U = interior(out[0][0][k])
        assert np.max(np.abs(U - interior(o["U"]))) <= 1e-10 * np.max(np.abs(interior(o["U"]))), k
