"""Batched ensembles of independent blast waves and the schedule autotuner.

Ensemble specification (seed 0, 512 instances by default): per instance, in
this order, three draws u1, u2, u3 from SplitMix64 mapped to [0, 1) by the top
53 bits, then
    side  = round(exp(ln 64 + u1 * (ln 1024 - ln 64)))   (log-uniform in [64, 1024])
    r0    = 0.1 + 0.4 * u2                                (uniform in [0.1, 0.5])
    p_in  = exp(u3 * ln 100)                              (log-uniform in [1, 100])
Each instance is a side x side grid on [0,2]^2 with the blast centred at
(1,1): rho = 1, p = p_in for r <= r0, rho = 0.125, p = 0.1 outside,
transmissive boundaries, CFL 0.8.

``Batch`` drives the batch kernels (csrc/batch.hip) through the C ABI;
``autotune`` times candidate schedules (tile shapes, the persistent kernel
with one or two launches per step or dispatched launches, workgroups per CU)
and then size classes (smaller instances with their own tile shapes) on the
batch itself and keeps the fastest.
"""
from __future__ import annotations

import ctypes as C
import json
import math
import os
import time

import numpy as np

from . import _lib
from ._lib import GridDesc, call
from .solver import Grid2DGPU, _stream

MASK64 = (1 << 64) - 1


def splitmix64(seed: int):
    x = seed & MASK64
    while True:
        x = (x + 0x9E3779B97F4A7C15) & MASK64
        z = x
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
        yield z ^ (z >> 31)


def ensemble_spec(n: int = 512, seed: int = 0):
    g = splitmix64(seed)
    u = lambda: (next(g) >> 11) * (1.0 / 9007199254740992.0)  # noqa: E731
    out = []
    for _ in range(n):
        u1, u2, u3 = u(), u(), u()
        side = int(round(math.exp(math.log(64.0) + u1 * (math.log(1024.0) - math.log(64.0)))))
        out.append({"side": side, "r0": 0.1 + 0.4 * u2, "p_in": math.exp(u3 * math.log(100.0))})
    return out


def blast_params(inst):
    """mh_init_state kind-2 parameters (x0, y0, r0, rho_in, p_in, rho_out, p_out)."""
    return np.array([1.0, 1.0, inst["r0"], 1.0, inst["p_in"], 0.125, 0.1])


class Batch:
    def __init__(self, spec, solver=1, cfl=0.8, domain=(0.0, 2.0, 0.0, 2.0)):
        self.spec = list(spec)
        self.n = len(self.spec)
        self.solver = int(solver)
        descs = (GridDesc * self.n)()
        for k, s in enumerate(self.spec):
            descs[k] = GridDesc(s["side"], s["side"], *map(float, domain))
        self._h = C.c_void_p()
        call("mh_batch_create", self.n, descs, self.solver, float(cfl), None, C.byref(self._h))
        self.cells = sum(s["side"] * s["side"] for s in self.spec)
        self.schedule = (0, 256, 0, 0, 1)
        self.size_class = (0, 0, 0)

    def grid(self, k: int) -> Grid2DGPU:
        """A Grid2DGPU-like view of instance k (non-owning)."""
        h = C.c_void_p()
        call("mh_batch_grid", self._h, int(k), C.byref(h))
        g = Grid2DGPU.__new__(Grid2DGPU)
        g._h = h
        s = self.spec[k]
        g.nx = g.ny = s["side"]
        g.ng = 2
        g.x_min, g.x_max, g.y_min, g.y_max = 0.0, 2.0, 0.0, 2.0
        g.close = lambda: None   # owned by the batch
        return g

    def initialise(self):
        for k, s in enumerate(self.spec):
            self.grid(k).init_state(2, blast_params(s))
        call("mh_batch_prepare", self._h, None)

    def set_schedule(self, xseg=0, ychunk=256, mode=0, blocks_per_cu=0, xrows=1):
        """x tiles: row segments of xseg cells (0 = whole rows), xrows rows
        each; y tiles: chunks of ychunk rows; mode 0: one persistent launch
        per step, 1: two persistent launches, 2: two dispatched launches with
        one tile per wavefront, 3: the fused step (one x/y wavefront pair per
        y tile; instance states alternate between their grids and scratch
        grids and are back in the grids after sync(), clock() or download())."""
        call("mh_batch_set_schedule", self._h, int(xseg), int(xrows), int(ychunk), int(mode), int(blocks_per_cu))
        self.schedule = (int(xseg), int(ychunk), int(mode), int(blocks_per_cu), int(xrows))

    def set_size_class(self, side_below=0, ychunk=0, xrows=0):
        """Instances with fewer than side_below cells a side take y chunks of
        ychunk rows and x tiles of xrows rows (0: the schedule's)."""
        call("mh_batch_set_size_class", self._h, int(side_below), int(ychunk), int(xrows))
        self.size_class = (int(side_below), int(ychunk), int(xrows))

    def capture(self, steps_per_graph: int, stream=None):
        call("mh_batch_capture", self._h, int(steps_per_graph), _stream(stream))

    def step(self, nsteps: int, stream=None):
        call("mh_batch_step", self._h, int(nsteps), _stream(stream))

    def clock(self, stream=None):
        t, dt = np.zeros(self.n), np.zeros(self.n)
        st, ss = np.zeros(self.n, np.int32), np.zeros(self.n, np.int32)
        call("mh_batch_clock", self._h, t.ctypes.data_as(C.c_void_p), dt.ctypes.data_as(C.c_void_p),
             st.ctypes.data_as(C.c_void_p), ss.ctypes.data_as(C.c_void_p), _stream(stream))
        return t, dt, st, ss

    def info(self):
        a, b, c = C.c_int(), C.c_int(), C.c_int()
        call("mh_batch_info", self._h, C.byref(a), C.byref(b), C.byref(c))
        return {"blocks": a.value, "x_tiles": b.value, "y_tiles": c.value}

    def sync(self, stream=None):
        """Waits for the enqueued steps and puts every instance's state in its
        grid (fused steps leave some in their scratch grids)."""
        call("mh_batch_sync", self._h, _stream(stream))

    def download(self, k: int):
        self.sync()
        return self.grid(k).download_to_aos()

    def close(self):
        if self._h:
            _lib.load().mh_batch_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# Size classes tried on the best dispatched or fused schedule: (side below which an
# instance takes them, y chunk rows, x tile rows); (0, 0, 0) is none.
SIZE_CLASSES = [(0, 0, 0)] + [(side, yc, xr) for side in (256, 512)
                              for yc, xr in ((32, 8), (64, 4), (64, 8), (128, 8), (32, 16), (256, 8), (0, 16))]

DEFAULT_CANDIDATES = [
    # (xseg, ychunk, mode, blocks_per_cu, xrows)
    (0, 256, 0, 0, 1), (0, 128, 1, 0, 1), (0, 256, 2, 0, 4), (0, 192, 2, 0, 4), (0, 128, 2, 0, 4),
    (0, 96, 2, 0, 4), (0, 64, 2, 0, 4), (0, 128, 2, 0, 1), (0, 128, 2, 0, 8),
    (0, 256, 3, 0, 1), (0, 192, 3, 0, 1), (0, 128, 3, 0, 1), (0, 96, 3, 0, 1), (0, 64, 3, 0, 1),
]


def _signature(batch: Batch) -> str:
    sides = sorted(s["side"] for s in batch.spec)
    return f"n{batch.n}-cells{batch.cells}-s{batch.solver}-min{sides[0]}-max{sides[-1]}"


def _time(batch, steps, set_up):
    set_up()
    batch.initialise()
    batch.step(1)
    batch.clock()
    t0 = time.perf_counter()
    batch.step(steps)
    batch.clock()
    return (time.perf_counter() - t0) / steps


def autotune(batch: Batch, candidates=None, steps: int = 8, cache: str | None = None, log=None, rounds: int = 2,
             size_classes=None):
    """Times each candidate schedule for `steps` steps (after one untimed
    step), in `rounds` interleaved rounds, and keeps the one with the fastest
    best round (one short sample per candidate let timing noise pick a
    schedule 9 % slower than the best); where that is a dispatched or fused
    schedule, then times the size classes (SIZE_CLASSES: smaller instances with their
    own tile shapes) on it the same way.  Applies the winner and
    re-initialises the batch.  Results are cached per ensemble signature in
    `cache` (JSON)."""
    sig = _signature(batch)
    if cache and os.path.exists(cache):
        with open(cache) as f:
            known = json.load(f)
        if sig in known:
            best = tuple(known[sig]["schedule"])
            cls = tuple(known[sig].get("size_class", (0, 0, 0)))
            batch.set_schedule(*best)
            batch.set_size_class(*cls)
            batch.initialise()
            return {"schedule": best, "size_class": cls, "cached": True, "timings": known[sig].get("timings")}
    cands = list(candidates or DEFAULT_CANDIDATES)
    best_dt = {}
    batch.set_size_class(0, 0, 0)
    for _ in range(max(1, rounds)):
        for cand in cands:
            dt = _time(batch, steps, lambda: batch.set_schedule(*cand))
            best_dt[cand] = min(dt, best_dt.get(cand, dt))
            if log:
                log(f"autotune {cand}: {dt * 1e3:.3f} ms/step")
    timings = sorted((t, c) for c, t in best_dt.items())
    best = timings[0][1]
    cls_dt = {(0, 0, 0): timings[0][0]}
    if best[2] >= 2:
        batch.set_schedule(*best)
        classes = list(size_classes or SIZE_CLASSES)
        if best[2] == 3:
            # the fused step has no x tiles: classes that differ only in x
            # tile rows are one class
            classes = list(dict.fromkeys((side, yc, 0) for side, yc, _ in classes))
        for _ in range(max(1, rounds)):
            for cls in classes:
                dt = _time(batch, steps, lambda: batch.set_size_class(*cls))
                cls_dt[cls] = min(dt, cls_dt.get(cls, dt))
                if log:
                    log(f"autotune size class {cls}: {dt * 1e3:.3f} ms/step")
    cls_best = min((t, c) for c, t in cls_dt.items())[1]
    batch.set_schedule(*best)
    batch.set_size_class(*cls_best)
    batch.initialise()
    res = {"schedule": best, "size_class": cls_best, "cached": False, "timings": [(c, t) for t, c in timings],
           "size_class_timings": sorted(((c, t) for c, t in cls_dt.items()), key=lambda x: x[1])}
    if cache:
        known = {}
        if os.path.exists(cache):
            with open(cache) as f:
                known = json.load(f)
        known[sig] = {"schedule": list(best), "size_class": list(cls_best),
                      "timings": [[list(c), t] for t, c in timings]}
        os.makedirs(os.path.dirname(cache) or ".", exist_ok=True)
        with open(cache, "w") as f:
            json.dump(known, f, indent=1)
    return res
