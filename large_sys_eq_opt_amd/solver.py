"""Python mirror of the reference GPU solver API over the HIP C ABI.

Reference API (paths relative to the reference repository) -> here:
  Grid2DGPU (head/gpu/grid_gpu.cuh:11-154)                 -> Grid2DGPU
  GpuWorkspace, init/free_gpu_workspace (solver_gpu.cuh:8-23) -> GpuWorkspace
  compute_dt_gpu (head/gpu/solver_gpu.cuh:25-29)            -> compute_dt_gpu
  advance_first_order_gpu (:31-36)                          -> advance_first_order_gpu
  advance_second_order_gpu (:38-44)                         -> advance_second_order_gpu
  apply_transmissive_boundary_gpu (head/gpu/boundary_gpu.cuh:5) -> apply_transmissive_boundary_gpu
  enum class RiemannSolver (head/riemann.hpp:13-18)          -> RiemannSolver
Every call goes to libmh_euler.so; nothing here computes on the host.
"""
from __future__ import annotations

import ctypes as C
import enum

import numpy as np

from . import _lib
from ._lib import BcDesc, GridDesc, MHError, call


class RiemannSolver(enum.IntEnum):  # head/riemann.hpp:13-18
    HLL = 0
    HLLC = 1
    Exact = 2
    FORCE = 3


def parse_riemann_solver(name) -> RiemannSolver:  # scripts/cpu/main_cpu.cpp:32-50
    if isinstance(name, (RiemannSolver, int)):
        return RiemannSolver(int(name))
    table = {"hll": 0, "HLL": 0, "hllc": 1, "HLLC": 1, "exact": 2, "Exact": 2, "EXACT": 2,
             "force": 3, "Force": 3, "FORCE": 3}
    if name not in table:
        raise ValueError(f"Unknown Riemann solver: {name}. Supported solvers are: hll, hllc, exact, force.")
    return RiemannSolver(table[name])


def riemann_solver_to_string(s) -> str:  # scripts/cpu/main_cpu.cpp:52-65
    return ["hll", "hllc", "exact", "force"][int(s)]


class Boundary:
    """Boundary spec: per edge (left, right, bottom, top) transmissive or a
    fixed inflow state (conserved)."""

    TRANSMISSIVE, INFLOW = 0, 1

    def __init__(self, kinds=(0, 0, 0, 0), states=None):
        self.kinds = tuple(int(k) for k in kinds)
        self.states = np.zeros((4, 4)) if states is None else np.asarray(states, dtype=np.float64).reshape(4, 4)

    @classmethod
    def transmissive(cls):
        return cls()

    @classmethod
    def inflow_left(cls, state):
        st = np.zeros((4, 4))
        st[0] = np.asarray(state, dtype=np.float64).reshape(4)
        return cls((1, 0, 0, 0), st)

    @property
    def left_inflow_state(self):
        return self.states[0].copy() if self.kinds[0] == 1 else None

    def desc(self) -> BcDesc:
        d = BcDesc()
        for e in range(4):
            d.kind[e] = self.kinds[e]
            for c in range(4):
                d.state[e][c] = float(self.states[e, c])
        return d


def _bc_ptr(bc):
    if bc is None:
        return None
    return C.byref(bc.desc())


def _stream(stream):
    if stream is None:
        return None
    if isinstance(stream, int):
        return C.c_void_p(stream)
    return C.c_void_p(int(getattr(stream, "cuda_stream", stream)))


def synchronize(stream=None):
    """Waits for the work queued on a stream (the null stream by default)."""
    call("mh_synchronize", _stream(stream))


class Grid2DGPU:
    """Device grid: four FP64 planes (rho, rhou, rhov, E), two ghost layers."""

    def __init__(self, nx, ny, ng=2, x_min=0.0, x_max=1.0, y_min=0.0, y_max=1.0):
        if ng != 2:
            raise ValueError("only ng = 2 is supported (the reference's value)")
        self._h = C.c_void_p()
        desc = GridDesc(int(nx), int(ny), float(x_min), float(x_max), float(y_min), float(y_max))
        call("mh_grid_create", C.byref(desc), C.byref(self._h))
        self.nx, self.ny, self.ng = int(nx), int(ny), 2
        self.x_min, self.x_max, self.y_min, self.y_max = float(x_min), float(x_max), float(y_min), float(y_max)
        dx, dy, rp, ps = C.c_double(), C.c_double(), C.c_int64(), C.c_int64()
        call("mh_grid_info", self._h, None, C.byref(dx), C.byref(dy), C.byref(rp), C.byref(ps))
        self.dx, self.dy, self.pitch, self.plane_stride = dx.value, dy.value, rp.value, ps.value

    @classmethod
    def from_config(cls, cfg):
        return cls(cfg.nx, cfg.ny, cfg.ng, cfg.x_min, cfg.x_max, cfg.y_min, cfg.y_max)

    @property
    def handle(self):
        return self._h

    @property
    def total_nx(self):
        return self.nx + 4

    @property
    def total_ny(self):
        return self.ny + 4

    def close(self):
        if self._h:
            _lib.load().mh_grid_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def upload_from_aos(self, host):
        a = np.ascontiguousarray(host, dtype=np.float64)
        if a.shape != (self.ny + 4, self.nx + 4, 4):
            raise ValueError(f"expected shape {(self.ny + 4, self.nx + 4, 4)}, got {a.shape}")
        call("mh_grid_upload_aos", self._h, a.ctypes.data_as(C.c_void_p))

    def download_to_aos(self):
        a = np.empty((self.ny + 4, self.nx + 4, 4))
        call("mh_grid_download_aos", self._h, a.ctypes.data_as(C.c_void_p))
        return a

    def plane_ptr(self, c: int) -> int:
        p = C.c_void_p()
        call("mh_grid_plane_ptr", self._h, int(c), C.byref(p))
        return p.value

    def copy_interior_from_device(self, ptr: int, stream=None):
        """Raw form: `ptr` must address 4*ny*nx float64 device values; prefer
        copy_interior_from_tensor, which checks the buffer."""
        call("mh_grid_copy_interior_from_device", self._h, C.c_void_p(ptr), _stream(stream))

    def copy_interior_to_device(self, ptr: int, stream=None):
        """Raw form of copy_interior_to_tensor."""
        call("mh_grid_copy_interior_to_device", self._h, C.c_void_p(ptr), _stream(stream))

    def _check_tensor(self, t):
        """The C ABI takes a bare pointer: check here that the tensor is what
        it will read or write (a float32 or misshapen tensor would be
        overrun on the device)."""
        import torch

        if not isinstance(t, torch.Tensor) or not t.is_cuda:
            raise TypeError("expected a torch tensor on a GPU")
        if t.dtype != torch.float64:
            raise TypeError(f"expected float64, got {t.dtype}")
        if tuple(t.shape) != (4, self.ny, self.nx) or not t.is_contiguous():
            raise ValueError(f"expected a contiguous (4, {self.ny}, {self.nx}) tensor, got {tuple(t.shape)}")

    def copy_interior_from_tensor(self, t, stream=None):
        """Interior <- a contiguous (4, ny, nx) float64 CUDA tensor."""
        self._check_tensor(t)
        self.copy_interior_from_device(t.data_ptr(), stream)

    def copy_interior_to_tensor(self, t, stream=None):
        """A contiguous (4, ny, nx) float64 CUDA tensor <- interior."""
        self._check_tensor(t)
        self.copy_interior_to_device(t.data_ptr(), stream)

    def swap(self, other: "Grid2DGPU"):
        call("mh_grid_swap", self._h, other._h)

    def init_state(self, kind: int, params=None, stream=None):
        p = None if params is None else np.ascontiguousarray(params, dtype=np.float64)
        call("mh_init_state", self._h, int(kind), None if p is None else p.ctypes.data_as(C.c_void_p),
             _stream(stream))
        self._keep = p


class GpuWorkspace:
    """Kept for the reference signature; the HIP path needs no workspace."""

    def __init__(self, grid=None):
        self.grid = grid


def init_gpu_workspace(ws: GpuWorkspace, grid: Grid2DGPU):
    ws.grid = grid


def free_gpu_workspace(ws: GpuWorkspace):
    ws.grid = None


def compute_dt_gpu(grid: Grid2DGPU, ws: GpuWorkspace | None, cfl: float, stream=None) -> float:
    """dt = cfl*min(dx,dy)/max(|u|+a, |v|+a) over interior cells; synchronises."""
    dt, m = C.c_double(), C.c_double()
    call("mh_compute_dt", grid.handle, float(cfl), C.byref(dt), C.byref(m), _stream(stream))
    return dt.value


def max_signal_speed_gpu(grid: Grid2DGPU, stream=None) -> float:
    dt, m = C.c_double(), C.c_double()
    lib = _lib.load()
    rc = lib.mh_compute_dt(grid.handle, 1.0, C.byref(dt), C.byref(m), _stream(stream))
    if rc not in (_lib.MH_OK, _lib.MH_ERR_WAVESPEED):
        raise MHError(rc, lib.mh_last_error().decode())
    return m.value


def advance_second_order_gpu(Uold, Utmp, Unew, ws, dt, solver, bc: Boundary | None = None, stream=None):
    call("mh_advance_second_order", Uold.handle, Utmp.handle, Unew.handle, float(dt),
         int(parse_riemann_solver(solver)), _bc_ptr(bc), _stream(stream))


def advance_first_order_gpu(Uold, Unew, dt, solver, bc: Boundary | None = None, stream=None):
    call("mh_advance_first_order", Uold.handle, Unew.handle, float(dt), int(parse_riemann_solver(solver)),
         _bc_ptr(bc), _stream(stream))


def sweep_gpu(Uin, Uout, dt, direction, solver, stream=None):
    call("mh_sweep", Uin.handle, Uout.handle, float(dt), int(direction), int(parse_riemann_solver(solver)),
         _stream(stream))


SHAPE_FIELDS = ("seg_x", "seg_y", "y_waves", "x_layout", "x_wpb", "y_wpb", "min_tiles", "fused", "fused_tail")


def set_launch_shape(**kw):
    """Sets fields of the sweeps' launch shape (see mh_launch_shape):
    seg_x, seg_y, y_waves, x_layout, x_wpb, y_wpb, min_tiles, fused,
    fused_tail."""
    cur = get_launch_shape()
    cur.update({k: int(v) for k, v in kw.items()})
    shape = _lib.LaunchShape(*(cur[f] for f in SHAPE_FIELDS))
    call("mh_set_launch_shape", C.byref(shape))


def get_launch_shape() -> dict:
    shape = _lib.LaunchShape()
    call("mh_get_launch_shape", C.byref(shape))
    return {f: getattr(shape, f) for f in SHAPE_FIELDS}


def apply_boundary_gpu(grid: Grid2DGPU, bc: Boundary | None = None, stream=None):
    call("mh_fill_ghosts", grid.handle, _bc_ptr(bc), _stream(stream))


def apply_transmissive_boundary_gpu(grid: Grid2DGPU, stream=None):  # src/gpu/boundary_gpu.cu:71-84
    apply_boundary_gpu(grid, None, stream)


class Stepper:
    """Device-resident time loop over state U: dt from the previous step's
    fused max-speed reduction, clamped to t_end, no host round trip."""

    def __init__(self, U: Grid2DGPU, solver, cfl, t_end=0.0, bc: Boundary | None = None, log_capacity=0,
                 stream=None):
        self._h = C.c_void_p()
        self.U = U
        self.log_capacity = int(log_capacity)
        call("mh_stepper_create", U.handle, int(solver), float(cfl), float(t_end), _bc_ptr(bc), self.log_capacity,
             C.byref(self._h), _stream(stream))

    def advance(self, nsteps: int, stream=None):
        call("mh_stepper_advance", self._h, int(nsteps), _stream(stream))

    def advance_profiled(self, nsteps: int, stream=None):
        """Advances with HIP events around every launch; returns (total ms,
        [clock, x sweep, y sweep] average ms per step)."""
        tot = C.c_double()
        ks = (C.c_double * 3)()
        call("mh_stepper_advance_profiled", self._h, int(nsteps), _stream(stream), C.byref(tot), ks)
        return tot.value, list(ks)

    def clock(self, stream=None):
        t, n = C.c_double(), C.c_int()
        log = np.zeros(max(self.log_capacity, 1))
        call("mh_stepper_clock", self._h, C.byref(t), C.byref(n), log.ctypes.data_as(C.c_void_p), _stream(stream))
        return t.value, n.value, log[: min(n.value, self.log_capacity)].copy()

    def finish(self, stream=None):
        call("mh_stepper_finish", self._h, _stream(stream))

    def set_stop(self, t_stop: float, stream=None):
        """Holds the loop after the first step with t >= t_stop (a snapshot
        point); enqueued steps past it do nothing until the next set_stop."""
        call("mh_stepper_set_stop", self._h, float(t_stop), _stream(stream))

    def close(self):
        if self._h:
            _lib.load().mh_stepper_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Snapshot:
    """Asynchronous snapshot staging for grids shaped like ``like``: each of
    ``nbuf`` slots holds a device copy and pinned host planes.  capture()
    enqueues the device copy on the caller's stream and the host copy on the
    snapshot's own stream; wait()/aos()/write_csv() wait for that slot."""

    def __init__(self, like: Grid2DGPU, nbuf: int = 2):
        self._h = C.c_void_p()
        self.nx, self.ny, self.nbuf = like.nx, like.ny, int(nbuf)
        call("mh_snapshot_create", like.handle, self.nbuf, C.byref(self._h))

    def capture(self, U: Grid2DGPU, slot: int, stream=None):
        call("mh_snapshot_capture", self._h, U.handle, int(slot), _stream(stream))

    def planes(self, slot: int):
        """The slot's four host planes as (ny+4, nx+4) numpy views (valid until
        the slot is captured again)."""
        p = (C.c_void_p * 4)()
        call("mh_snapshot_wait", self._h, int(slot), p)
        shape = (self.ny + 4, self.nx + 4)
        n = shape[0] * shape[1]
        return [np.ctypeslib.as_array(C.cast(p[c], C.POINTER(C.c_double)), shape=(n,)).reshape(shape)
                for c in range(4)]

    def aos(self, slot: int) -> np.ndarray:
        """A copy of the slot in the reference AoS layout, (ny+4, nx+4, 4)."""
        return np.stack(self.planes(slot), axis=-1)

    def write_csv(self, slot: int, path: str, precision: int = 6, nthreads: int = 0):
        call("mh_snapshot_write_csv", self._h, int(slot), path.encode(), int(precision), int(nthreads))

    def close(self):
        if self._h:
            _lib.load().mh_snapshot_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def write_csv_aos(U_aos: np.ndarray, cfg, filename: str, precision: int = 6, nthreads: int = 0):
    """The reference's snapshot CSV (scripts/cpu/main_cpu.cpp:136-164) of a host
    AoS state, formatted by the library on several threads."""
    a = np.ascontiguousarray(U_aos, dtype=np.float64)
    if a.shape != (cfg.ny + 4, cfg.nx + 4, 4):
        raise ValueError(f"expected shape {(cfg.ny + 4, cfg.nx + 4, 4)}, got {a.shape}")
    desc = GridDesc(int(cfg.nx), int(cfg.ny), float(cfg.x_min), float(cfg.x_max), float(cfg.y_min), float(cfg.y_max))
    call("mh_write_csv_aos", a.ctypes.data_as(C.c_void_p), C.byref(desc), int(precision), int(nthreads),
         filename.encode())
