"""ctypes wrapper over the oracle and the compiled reference CPU solver.

TEST INFRASTRUCTURE ONLY.  Only tests/, ``__graft_entry__.smoke()`` and the
CPU-baseline leg of ``bench.py`` import this module, and only as the checker
or the timed CPU baseline -- never as the product path.

* ``Oracle``   -- ``oracle/_build/liboracle.so``: the C++ restatement of the
  reference CPU solver (src/cpu/solver_cpu.cpp, head/riemann.hpp,
  head/physics.hpp) plus a forward-mode AD instantiation.
* ``Reference`` -- ``oracle/_ref/libmuscl_ref.so``: the reference's own CPU
  sources compiled by ``oracle/Makefile`` (present only where it was built).

States are numpy float64 arrays of shape ``(ny+4, nx+4, 4)`` -- the
reference's AoS ``Conserved`` layout with two ghost layers
(head/cpu/grid_cpu.hpp:86-88).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "_build", "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libmuscl_ref.so")
REFERENCE_TREE = "/root/reference"

SOLVERS = {"hll": 0, "hllc": 1, "exact": 2, "force": 3}
_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_D = C.c_double
_I = C.c_int
_P = C.c_void_p


def build(reference: bool | None = None) -> None:
    """Compile the oracle (and the reference harness when the tree exists)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    if reference is None:
        reference = os.path.isdir(os.path.join(REFERENCE_TREE, "src"))
    if reference:
        subprocess.run(["make", "-s", "-C", HERE, "ref"], check=True)


def _ptr(a):
    return None if a is None else a.ctypes.data_as(_P)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def bc_arrays(inflow_left=None):
    """Boundary spec: transmissive on every edge, optionally a Dirichlet inflow
    state (conserved, 4 values) on the left edge."""
    kind = np.zeros(4, dtype=np.int32)
    state = np.zeros(16, dtype=np.float64)
    if inflow_left is not None:
        kind[0] = 1
        state[0:4] = inflow_left
    return kind, state


class Oracle:
    def __init__(self, path: str = ORACLE_SO):
        if not os.path.exists(path):
            raise FileNotFoundError(f"oracle library not built: {path} (run make -C oracle)")
        L = self.lib = C.CDLL(path)
        L.orc_num_threads.restype = _I
        L.orc_riemann_flux.argtypes = [_I, _P, _P, _I, _I, _P]
        L.orc_exact_star.argtypes = [_I, _P, _P, _I, _P, _P, _P, _P]
        L.orc_reconstruct.argtypes = [_I, _P, _P, _P, _D, _I, _P, _P]
        L.orc_prim_to_cons.argtypes = [_I, _P, _P]
        L.orc_cons_to_prim.argtypes = [_I, _P, _P]
        L.orc_init.argtypes = [_I, _P, _I, _I, _D, _D, _D, _D, _P]
        L.orc_init.restype = _I
        L.orc_init_smooth.argtypes = [_P, _I, _I, _D, _D, _D, _D, _P, _P, _D, _D, _P, _P]
        L.orc_apply_boundary.argtypes = [_I, _I, _P, _P, _P]
        L.orc_compute_dt.argtypes = [_I, _I, _D, _D, _P, _D, _P]
        L.orc_compute_dt.restype = _D
        L.orc_advance.argtypes = [_I, _I, _D, _D, _P, _P, _P, _D, _I, _I, _P, _P]
        L.orc_run.argtypes = [_I, _I, _D, _D, _P, _I, _I, _D, _D, _I, _P, _P, _P, _P, _P, _I, _P, _P, _P]
        L.orc_run.restype = _I
        L.orc_run_tangent.argtypes = [_I, _I, _D, _D, _P, _P, _I, _I, _P, _P, _P]
        L.orc_sweep.argtypes = [_I, _I, _D, _D, _P, _P, _D, _I, _I]
        L.orc_sweep_tangent.argtypes = [_I, _I, _D, _D, _P, _P, _D, _I, _I, _P, _P]

    def num_threads(self) -> int:
        return self.lib.orc_num_threads()

    # -- pointwise kernels ---------------------------------------------------
    def riemann_flux(self, UL, UR, direction: int, solver: int):
        UL, UR = _f64(UL), _f64(UR)
        F = np.empty_like(UL)
        self.lib.orc_riemann_flux(UL.shape[0], _ptr(UL), _ptr(UR), direction, solver, _ptr(F))
        return F

    def exact_star(self, VL, VR, direction: int = 0):
        VL, VR = _f64(VL), _f64(VR)
        n = VL.shape[0]
        ps, us = np.empty(n), np.empty(n)
        it, ok = np.empty(n, np.int32), np.empty(n, np.int32)
        self.lib.orc_exact_star(n, _ptr(VL), _ptr(VR), direction, _ptr(ps), _ptr(us), _ptr(it), _ptr(ok))
        return ps, us, it, ok

    def reconstruct(self, Um, Uc, Up, dt_over_d: float, direction: int):
        Um, Uc, Up = _f64(Um), _f64(Uc), _f64(Up)
        L, R = np.empty_like(Uc), np.empty_like(Uc)
        self.lib.orc_reconstruct(Uc.shape[0], _ptr(Um), _ptr(Uc), _ptr(Up), dt_over_d, direction, _ptr(L), _ptr(R))
        return L, R

    def prim_to_cons(self, V):
        V = _f64(V)
        U = np.empty_like(V)
        self.lib.orc_prim_to_cons(V.shape[0], _ptr(V), _ptr(U))
        return U

    def cons_to_prim(self, U):
        U = _f64(U)
        V = np.empty_like(U)
        self.lib.orc_cons_to_prim(U.shape[0], _ptr(U), _ptr(V))
        return V

    # -- grids ---------------------------------------------------------------
    def init(self, kind: int, params, nx: int, ny: int, bounds):
        U = np.empty((ny + 4, nx + 4, 4))
        p = _f64(params if params is not None else np.zeros(8))
        rc = self.lib.orc_init(kind, _ptr(p), nx, ny, *map(float, bounds), _ptr(U))
        if rc != 0:
            raise ValueError(f"unknown initial-condition kind {kind}")
        return U

    def init_smooth(self, params, nx, ny, bounds, rho_b, R, d_rho_b=None, d_R=0.0):
        U = np.empty((ny + 4, nx + 4, 4))
        dU = np.empty_like(U)
        p = _f64(params)
        rb = _f64(rho_b)
        drb = None if d_rho_b is None else _f64(d_rho_b)
        self.lib.orc_init_smooth(_ptr(p), nx, ny, *map(float, bounds), _ptr(rb), _ptr(drb), R, d_R, _ptr(U), _ptr(dU))
        return U, dU

    def apply_boundary(self, U, inflow_left=None):
        ny, nx = U.shape[0] - 4, U.shape[1] - 4
        k, s = bc_arrays(inflow_left)
        self.lib.orc_apply_boundary(nx, ny, _ptr(U), _ptr(k), _ptr(s))
        return U

    def compute_dt(self, U, dx, dy, cfl):
        ny, nx = U.shape[0] - 4, U.shape[1] - 4
        smax = C.c_double(0.0)
        dt = self.lib.orc_compute_dt(nx, ny, dx, dy, _ptr(_f64(U)), cfl, C.byref(smax))
        return dt, smax.value

    def advance(self, U, dx, dy, dt, solver, order=2, inflow_left=None):
        ny, nx = U.shape[0] - 4, U.shape[1] - 4
        U = _f64(U)
        tmp, new = np.zeros_like(U), np.zeros_like(U)
        k, s = bc_arrays(inflow_left)
        self.lib.orc_advance(nx, ny, dx, dy, _ptr(U), _ptr(tmp), _ptr(new), dt, solver, order, _ptr(k), _ptr(s))
        return new, tmp

    def sweep(self, U, dx, dy, dt, direction, solver):
        ny, nx = U.shape[0] - 4, U.shape[1] - 4
        U = _f64(U)
        out = U.copy()
        self.lib.orc_sweep(nx, ny, dx, dy, _ptr(U), _ptr(out), dt, direction, solver)
        return out

    def run(self, U, dx, dy, solver, cfl, t_end=0.0, max_steps=1 << 30, order=2, inflow_left=None,
            dt_in=None, nsnap=0, keep_snapshots=False):
        """The reference time loop (scripts/cpu/main_cpu.cpp:206-262)."""
        ny, nx = U.shape[0] - 4, U.shape[1] - 4
        U = _f64(U).copy()
        cap = max_steps if max_steps < (1 << 24) else 1 << 20
        dt_log = np.zeros(cap)
        k, s = bc_arrays(inflow_left)
        t_out, secs = C.c_double(0.0), C.c_double(0.0)
        snap_t = np.zeros(max(nsnap, 1))
        snaps = np.zeros((max(nsnap, 1),) + U.shape) if keep_snapshots and nsnap else None
        dti = None if dt_in is None else _f64(dt_in)
        steps = self.lib.orc_run(nx, ny, dx, dy, _ptr(U), solver, order, cfl, t_end, max_steps, _ptr(k), _ptr(s),
                                 _ptr(dti), _ptr(dt_log), C.byref(t_out), nsnap, _ptr(snap_t), _ptr(snaps),
                                 C.byref(secs))
        if steps < 0:
            raise RuntimeError("compute_dt: non-positive maximum wave speed")
        return dict(U=U, steps=steps, t=t_out.value, dt=dt_log[:steps].copy(), snap_t=snap_t[:nsnap],
                    snapshots=snaps, seconds=secs.value)

    def sweep_tangent(self, U, dU, dx, dy, dt, direction, solver):
        """Tangent of one sweep: (sweep(U), d sweep(U)[dU]), dU over every cell."""
        ny, nx = U.shape[0] - 4, U.shape[1] - 4
        U, dU = _f64(U), _f64(dU)
        out, dout = U.copy(), np.zeros_like(U)
        self.lib.orc_sweep_tangent(nx, ny, dx, dy, _ptr(U), _ptr(dU), dt, direction, solver, _ptr(out), _ptr(dout))
        return out, dout

    def run_tangent(self, U, dU, dx, dy, solver, dt_seq, inflow_left=None):
        ny, nx = U.shape[0] - 4, U.shape[1] - 4
        U, dU = _f64(U).copy(), _f64(dU).copy()
        k, s = bc_arrays(inflow_left)
        dts = _f64(dt_seq)
        self.lib.orc_run_tangent(nx, ny, dx, dy, _ptr(U), _ptr(dU), solver, len(dts), _ptr(dts), _ptr(k), _ptr(s))
        return U, dU


class Reference:
    """The reference's own CPU solver, compiled from its sources."""

    def __init__(self, path: str = REF_SO):
        if not os.path.exists(path):
            raise FileNotFoundError(f"reference harness not built: {path}")
        L = self.lib = C.CDLL(path)
        L.ref_riemann_flux.argtypes = [_I, _P, _P, _I, _I, _P]
        L.ref_reconstruct.argtypes = [_I, _P, _P, _P, _D, _I, _P, _P]
        L.ref_exact_star.argtypes = [_I, _P, _P, _I, _P, _P, _P]
        L.ref_case_config.argtypes = [_I, _I, _P, _P, _P, _P, _P]
        L.ref_case_config.restype = _I
        L.ref_init_case.argtypes = [_I, _I, _P]
        L.ref_compute_dt.argtypes = [_I, _I, _P, _P, _D]
        L.ref_compute_dt.restype = _D
        L.ref_run.argtypes = [_I, _I, _P, _P, _I, _I, _D, _D, _I, _P, _P, _P, _P, _I, _P, _P]
        L.ref_run.restype = _I

    def riemann_flux(self, UL, UR, direction, solver):
        UL, UR = _f64(UL), _f64(UR)
        F = np.empty_like(UL)
        self.lib.ref_riemann_flux(UL.shape[0], _ptr(UL), _ptr(UR), direction, solver, _ptr(F))
        return F

    def reconstruct(self, Um, Uc, Up, dt_over_d, direction):
        Um, Uc, Up = _f64(Um), _f64(Uc), _f64(Up)
        L, R = np.empty_like(Uc), np.empty_like(Uc)
        self.lib.ref_reconstruct(Uc.shape[0], _ptr(Um), _ptr(Uc), _ptr(Up), dt_over_d, direction, _ptr(L), _ptr(R))
        return L, R

    def exact_star(self, VL, VR, direction=0):
        VL, VR = _f64(VL), _f64(VR)
        n = VL.shape[0]
        ps, us, ok = np.empty(n), np.empty(n), np.empty(n, np.int32)
        self.lib.ref_exact_star(n, _ptr(VL), _ptr(VR), direction, _ptr(ps), _ptr(us), _ptr(ok))
        return ps, us, ok

    def case_config(self, case_id: int, n: int = 1):
        nx, ny = C.c_int(0), C.c_int(0)
        b = np.zeros(4)
        cfl, t_end = C.c_double(0), C.c_double(0)
        ng = self.lib.ref_case_config(case_id, n, C.byref(nx), C.byref(ny), _ptr(b), C.byref(cfl), C.byref(t_end))
        return dict(nx=nx.value, ny=ny.value, ng=ng, bounds=tuple(b), cfl=cfl.value, t_end=t_end.value)

    def init_case(self, case_id: int, n: int = 1):
        c = self.case_config(case_id, n)
        U = np.empty((c["ny"] + 4, c["nx"] + 4, 4))
        self.lib.ref_init_case(case_id, n, _ptr(U))
        return U, c

    def compute_dt(self, U, bounds, cfl):
        ny, nx = U.shape[0] - 4, U.shape[1] - 4
        return self.lib.ref_compute_dt(nx, ny, _ptr(_f64(bounds)), _ptr(_f64(U)), cfl)

    def run(self, U, bounds, solver, cfl, t_end=0.0, max_steps=1 << 30, order=2, inflow_left=None, dt_in=None,
            nsnap=0):
        ny, nx = U.shape[0] - 4, U.shape[1] - 4
        U = _f64(U).copy()
        cap = max_steps if max_steps < (1 << 24) else 1 << 20
        dt_log = np.zeros(cap)
        t_out, secs = C.c_double(0.0), C.c_double(0.0)
        snap_t = np.zeros(max(nsnap, 1))
        inflow = None if inflow_left is None else _f64(inflow_left)
        dti = None if dt_in is None else _f64(dt_in)
        steps = self.lib.ref_run(nx, ny, _ptr(_f64(bounds)), _ptr(U), solver, order, cfl, t_end, max_steps,
                                 _ptr(inflow), _ptr(dti), _ptr(dt_log), C.byref(t_out), nsnap, _ptr(snap_t),
                                 C.byref(secs))
        return dict(U=U, steps=steps, t=t_out.value, dt=dt_log[:steps].copy(), snap_t=snap_t[:nsnap],
                    seconds=secs.value)


def reference_available() -> bool:
    return os.path.exists(REF_SO)
