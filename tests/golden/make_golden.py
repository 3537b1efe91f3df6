"""Writes the golden fixtures of tests/golden/ from the reference CPU solver.

The reference's own C++ sources are compiled by oracle/Makefile (target
``ref``) into oracle/_ref/libmuscl_ref.so; this script calls that library
and stores inputs and outputs as float64 arrays in ``golden.npz`` (no
pickled objects).  Run it in a tree where the reference is present:

    make -C oracle ref && python tests/golden/make_golden.py

Contents
  flux_{UL,UR}                 face states: random physical pairs plus pairs
                               that drive each fallback branch
  flux_s{S}_d{D}               riemann_flux (head/riemann.hpp:650-672) for
                               solver S (0 HLL, 1 HLLC, 2 exact, 3 FORCE),
                               direction D
  rec_{Um,Uc,Up}, rec_dtd      reconstruction inputs
  rec_{L,R}_d{D}_k{K}          reconstruct_cell_muscl_hancock
                               (src/cpu/solver_cpu.cpp:68-99) at dt/d = rec_dtd[K]
  star_{VL,VR}, star_{p,u,ok}  exact_star_pressure_velocity (:379-443)
  run_{case}_s{S}_{U0,U,dt}    the reference loop (scripts/cpu/main_cpu.cpp:
                               221-262) on small grids: initial state, final
                               state and dt sequence
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, ROOT)

import pyoracle  # noqa: E402

# Small runs: (name, IC kind, IC params, nx, ny, bounds, cfl, steps, left inflow)
RUNS = [
    ("cylblast24", 2, [1.0, 1.0, 0.4, 1.0, 1.0, 0.125, 0.1], 24, 24, (0.0, 2.0, 0.0, 2.0), 0.8, 30, False),
    ("sbnd48", 3, [1.0, 1.0, 1.22, 0.1, 0.138, 0.5, 0.5, 0.2], 48, 24, (0.0, 2.0, 0.0, 1.0), 0.8, 30, True),
    ("refblast20", 2, [0.5, 0.5, 0.1, 1.0, 100.0, 1.0, 1.0], 20, 20, (0.0, 1.0, 0.0, 1.0), 0.4, 25, False),
]


def face_pairs(orc, rng):
    def prims(n):
        V = np.empty((n, 4))
        V[:, 0] = rng.uniform(0.1, 2.0, n)
        V[:, 1:3] = rng.uniform(-1.5, 1.5, (n, 2))
        V[:, 3] = rng.uniform(0.05, 5.0, n)
        return V

    VL = [prims(160)]
    VR = [prims(160)]
    VL.append(np.array([[1.0, 5.0, 0.1, 1.0], [1.0, -5.0, 0.0, 1.0], [1.0, 0.0, 0.0, 1000.0],
                        [1.0, 0.3, -0.2, 1.0], [1.0, -3.0, 0.0, 0.4], [1.0, 2.0, 0.0, 0.4],
                        [1.0, 0.0, 0.0, 1.0], [1.0, 0.75, 0.0, 1.0], [0.445, 0.698, 0.0, 3.528],
                        [5.99924, 19.5975, 0.0, 460.894]]))
    VR.append(np.array([[1.0, 5.0, 0.0, 1.0], [1.0, -5.0, 0.0, 1.0], [1.0, 0.0, 0.0, 0.01],
                        [1.0, 0.3, -0.2, 1.0], [1.0, 3.0, 0.0, 0.4], [1.0, -2.0, 0.0, 0.4],
                        [0.125, 0.0, 0.0, 0.1], [0.125, 0.0, 0.0, 0.1], [0.5, 0.0, 0.0, 0.571],
                        [5.99242, -6.19633, 0.0, 46.0950]]))
    UL = orc.prim_to_cons(np.concatenate(VL))
    UR = orc.prim_to_cons(np.concatenate(VR))
    bad = np.array([[1.0, 0.0, 0.0, -1.0], [1.0, 0.0, 0.0, 0.0], [1.0, 0.5, 0.0, 0.1]])
    UL = np.concatenate([UL, bad, UR[:3]])
    UR = np.concatenate([UR, UL[:3], bad])
    return UL, UR


def main():
    ref = pyoracle.Reference()
    orc = pyoracle.Oracle()
    rng = np.random.default_rng(20240601)
    out = {}
    UL, UR = face_pairs(orc, rng)
    out["flux_UL"], out["flux_UR"] = UL, UR
    for s in range(4):
        for d in range(2):
            out[f"flux_s{s}_d{d}"] = ref.riemann_flux(UL, UR, d, s)

    n = 200
    Vs = [np.column_stack([rng.uniform(0.1, 2.0, n), rng.uniform(-1, 1, n), rng.uniform(-1, 1, n),
                           rng.uniform(0.05, 5.0, n)]) for _ in range(3)]
    Vs[0][:10, 3] = 1e-13     # neighbour below the pressure floor
    Vs[2][10:20, 0] = 5e-13   # neighbour below the density floor
    Um, Uc, Up = (orc.prim_to_cons(v) for v in Vs)
    out["rec_Um"], out["rec_Uc"], out["rec_Up"] = Um, Uc, Up
    out["rec_dtd"] = np.array([0.05, 0.4, 3.0])
    for d in range(2):
        for k, dtd in enumerate(out["rec_dtd"]):
            L, R = ref.reconstruct(Um, Uc, Up, float(dtd), d)
            out[f"rec_L_d{d}_k{k}"], out[f"rec_R_d{d}_k{k}"] = L, R

    VL = np.column_stack([rng.uniform(0.1, 2, 100), rng.uniform(-2, 2, 100), np.zeros(100), rng.uniform(0.01, 10, 100)])
    VR = np.column_stack([rng.uniform(0.1, 2, 100), rng.uniform(-2, 2, 100), np.zeros(100), rng.uniform(0.01, 10, 100)])
    out["star_VL"], out["star_VR"] = VL, VR
    ps, us, ok = ref.exact_star(VL, VR, 0)
    out["star_p"], out["star_u"], out["star_ok"] = ps, us, ok.astype(np.float64)

    for name, kind, params, nx, ny, bounds, cfl, steps, inflow in RUNS:
        U0 = orc.init(kind, np.array(params), nx, ny, bounds)
        infl = None
        if inflow:
            rho0, p0, mach = params[0], params[1], params[2]
            a0 = np.sqrt(1.4 * p0 / rho0)
            M2 = mach * mach
            rho2 = rho0 * ((1.4 + 1.0) * M2) / ((1.4 - 1.0) * M2 + 2.0)
            p2 = p0 * (1.0 + 2.0 * 1.4 / (1.4 + 1.0) * (M2 - 1.0))
            u2 = 2.0 * a0 / (1.4 + 1.0) * (mach - 1.0 / mach)
            infl = orc.prim_to_cons(np.array([[rho2, u2, 0.0, p2]]))[0]
            out[f"run_{name}_inflow"] = infl
        out[f"run_{name}_U0"] = U0
        for s in range(4):
            r = ref.run(U0, bounds, s, cfl, max_steps=steps, inflow_left=infl)
            out[f"run_{name}_s{s}_U"] = r["U"][2:-2, 2:-2].copy()
            out[f"run_{name}_s{s}_dt"] = r["dt"]
    meta = np.array([[kind, nx, ny, cfl, steps, 1.0 if inflow else 0.0, *bounds]
                     for name, kind, params, nx, ny, bounds, cfl, steps, inflow in RUNS])
    out["run_meta"] = meta
    path = os.path.join(HERE, "golden.npz")
    np.savez_compressed(path, **out)
    print(f"wrote {path}: {len(out)} arrays, {os.path.getsize(path)} bytes")


if __name__ == "__main__":
    main()
