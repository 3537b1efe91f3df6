"""Discrete adjoint of the split MUSCL-Hancock step (all four Riemann
solvers; for the exact solver, the implicit derivative of p* at the converged
Newton root) and its PyTorch binding.

``Adjoint`` wraps the C ABI workspace (csrc/adjoint.hip): the vector-Jacobian
product of one sweep or one split step at a given dt, and N-step
forward/backward with a checkpoint every K steps (each K-step segment is
recomputed once in reverse, so memory is N/K + K states).  The time-step
sequence is that of the primal run and is held fixed.

``EulerSolve`` is a ``torch.autograd.Function``: forward runs N steps from a
(4, ny, nx) float64 CUDA tensor of conserved interior states and returns the
final state; backward runs the adjoint.

The shock-bubble adjoint case (``shock_bubble_adj``: 512x256 over [0,2]x[0,1],
400 HLLC steps) uses a smoothed bubble so that the objective is
differentiable in the bubble radius: the bubble fraction of a cell is
f = clamp((R - r)/h + 1/2, 0, 1) with r the distance of the cell centre from
(0.5, 0.5) and h the mean cell size, and rho = f*rho_b + (1-f)*rho_ambient
with a per-cell bubble density rho_b.  The objective is
J = sum over cell centres in [1.2,1.6]x[0.3,0.7] of rho*dx*dy at the final
step.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from . import cases as _cases
from ._lib import GridDesc, call
from .solver import (Boundary, Grid2DGPU, _bc_ptr, _stream, advance_second_order_gpu, apply_boundary_gpu)


class Adjoint:
    def __init__(self, cfg, solver=1, bc: Boundary | None = None, checkpoint_every: int = 20, max_steps: int = 0):
        self.cfg = cfg
        desc = GridDesc(cfg.nx, cfg.ny, cfg.x_min, cfg.x_max, cfg.y_min, cfg.y_max)
        self._h = C.c_void_p()
        self.bc = bc
        call("mh_adjoint_create", C.byref(desc), int(solver), _bc_ptr(bc), int(checkpoint_every), int(max_steps),
             C.byref(self._h))
        self.solver = int(solver)
        self.nsteps = 0
        # Each forward() replaces the stored trajectory (checkpoints and dt
        # sequence) that backward() reverses; callers that keep a trajectory
        # for later note this counter and check it before backward().
        self.generation = 0

    def step_vjp(self, U: Grid2DGPU, dt: float, lam_in: Grid2DGPU, lam_out: Grid2DGPU, stream=None):
        call("mh_adjoint_step_vjp", self._h, U.handle, float(dt), lam_in.handle, lam_out.handle, _stream(stream))

    def sweep_vjp(self, U: Grid2DGPU, dt: float, direction: int, lam_in: Grid2DGPU, lam_out: Grid2DGPU, stream=None):
        call("mh_adjoint_sweep_vjp", self._h, U.handle, float(dt), int(direction), lam_in.handle, lam_out.handle,
             _stream(stream))

    def forward(self, U0: Grid2DGPU, nsteps: int, cfl: float, dt_in=None, UN: Grid2DGPU | None = None, stream=None):
        dts = np.zeros(max(int(nsteps), 1))
        din = None if dt_in is None else np.ascontiguousarray(dt_in, dtype=np.float64)
        call("mh_adjoint_forward", self._h, U0.handle, int(nsteps), float(cfl),
             None if din is None else din.ctypes.data_as(C.c_void_p), dts.ctypes.data_as(C.c_void_p),
             None if UN is None else UN.handle, _stream(stream))
        self.nsteps = int(nsteps)
        self.generation += 1
        return dts[:nsteps].copy()

    def backward(self, lamN: Grid2DGPU, lam0: Grid2DGPU, stream=None):
        call("mh_adjoint_backward", self._h, lamN.handle, lam0.handle, _stream(stream))

    def close(self):
        if self._h:
            _lib.load().mh_adjoint_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def zero_grid(g: Grid2DGPU, stream=None):
    call("mh_grid_zero", g.handle, _stream(stream))


# ---------------------------------------------------------------------------
# shock-bubble adjoint case
# ---------------------------------------------------------------------------
ADJ_WINDOW = (1.2, 1.6, 0.3, 0.7)


def window_mask(cfg, window=ADJ_WINDOW):
    """Boolean (ny, nx) mask of cell centres inside the objective window."""
    x = cfg.x_min + (np.arange(cfg.nx) + 0.5) * cfg.dx
    y = cfg.y_min + (np.arange(cfg.ny) + 0.5) * cfg.dy
    mx = (x >= window[0]) & (x <= window[1])
    my = (y >= window[2]) & (y <= window[3])
    return my[:, None] & mx[None, :]


def _pad(a, cfg):
    """A scalar or an interior (ny, nx) array as a (ny+4, nx+4) array (ghost
    cells take the nearest interior value; the boundary fill overwrites the
    state there before it is read)."""
    a = np.asarray(a, dtype=np.float64)
    if a.ndim == 0:
        return np.full((cfg.ny + 4, cfg.nx + 4), float(a))
    if a.shape == (cfg.ny + 4, cfg.nx + 4):
        return a
    return np.pad(a.reshape(cfg.ny, cfg.nx), 2, mode="edge")


def smooth_bubble_state(cfg, rho_b, R, params=None):
    """Host (numpy) smoothed shock-bubble initial state, ghosts included:
    returns U (ny+4, nx+4, 4), dU/drho_b and dU/dR per cell (the state's
    derivative along rho_b and R).  rho_b is a scalar or per cell."""
    p = params or _cases.SHOCK_BUBBLE_ND
    rho_b = _pad(rho_b, cfg)
    rho2, u2, p2 = _cases.post_shock_state(p.rho0, p.p0, p.mach)
    X, Y = _cases.cell_centers(cfg)
    h = 0.5 * (cfg.dx + cfg.dy)
    behind = X < p.shock_x
    rho_a = np.where(behind, rho2, p.rho0)
    u = np.where(behind, u2, 0.0)
    pr = np.where(behind, p2, p.p0)
    r = np.sqrt((X - p.bubble_x) ** 2 + (Y - p.bubble_y) ** 2)
    f_raw = (R - r) / h + 0.5
    f = np.clip(f_raw, 0.0, 1.0)
    dfdR = np.where((f_raw > 0.0) & (f_raw < 1.0), 1.0 / h, 0.0)
    rho = f * rho_b + (1.0 - f) * rho_a
    U = _cases.make_conserved(rho, u, np.zeros_like(u), pr)
    dUdrho = np.stack([np.ones_like(u), u, np.zeros_like(u), 0.5 * (u * u)], axis=-1)
    return U, dUdrho * f[..., None], dUdrho * ((rho_b - rho_a) * dfdR)[..., None]


def objective_weights(cfg, window=ADJ_WINDOW):
    """dJ/dU^N as a host (ny+4, nx+4, 4) array: dx*dy on the density of the
    window's cells, zero elsewhere (J is linear in U^N)."""
    w = np.zeros((cfg.ny + 4, cfg.nx + 4, 4))
    w[2:-2, 2:-2, 0] = np.where(window_mask(cfg, window), cfg.dx * cfg.dy, 0.0)
    return w


def bubble_support(cfg, R, params=None):
    """Interior cells whose initial density depends on the bubble density
    (bubble fraction > 0 at radius R)."""
    p = params or _cases.SHOCK_BUBBLE_ND
    X, Y = _cases.cell_centers(cfg)
    r = np.sqrt((X - p.bubble_x) ** 2 + (Y - p.bubble_y) ** 2)[2:-2, 2:-2]
    return (R - r) / (0.5 * (cfg.dx + cfg.dy)) + 0.5 > 0.0


class ShockBubbleAdjoint:
    """The smoothed shock-bubble adjoint case: 512x256 cells over [0,2]x[0,1]
    (dx = dy = 1/256), left inflow of the post-shock state, HLLC, 400 steps
    at CFL 0.8 with a checkpoint every 20 steps.  J is the density integral
    over a window (default [1.2,1.6]x[0.3,0.7]) at the last step; parameters
    are the per-cell bubble density rho_b (ny, nx) and the radius R.

    At step 400 (t ~ 0.39) the incident shock is near x = 0.68 and no wave
    has reached x = 1.2: over the default window J is the undisturbed
    density integral and its gradient vanishes to rounding (~1e-21).  A
    window the flow reaches, such as [0.4,0.8]x[0.3,0.7], exercises it."""

    def __init__(self, nsteps=400, checkpoint_every=20, solver=1, nx=None, ny=None, window=ADJ_WINDOW):
        case = "shock_bubble_adj"
        cfg = _cases.get_case_config(case)
        if nx is not None:
            cfg = _cases.CaseConfig(nx, ny, cfg.ng, cfg.x_min, cfg.x_max, cfg.y_min, cfg.y_max, cfg.cfl, cfg.t_end)
        self.cfg = cfg
        self.params = _cases.SHOCK_BUBBLE_ND
        self.nsteps = int(nsteps)
        self.bc = _cases.case_boundary(case)
        self.adj = Adjoint(cfg, solver, self.bc, checkpoint_every, self.nsteps)
        self.U0 = Grid2DGPU.from_config(cfg)
        self.UN = Grid2DGPU.from_config(cfg)
        self.lamN = Grid2DGPU.from_config(cfg)
        self.lam0 = Grid2DGPU.from_config(cfg)
        self.window = tuple(window)
        self.weights = objective_weights(cfg, self.window)
        self.lamN.upload_from_aos(self.weights)
        self.dts = None
        self._dU = None
        self._gen = -1

    def objective(self, rho_b, R, dt_in=None):
        """Forward run from the smoothed state; returns J.  Without dt_in the
        run takes CFL steps (recorded in self.dts) through the checkpointing
        forward, and gradient() reverses it.  With dt_in (frozen time steps,
        e.g. for finite differences) it is a plain forward run that leaves
        the stored trajectory alone."""
        U, dUr, dUR = smooth_bubble_state(self.cfg, rho_b, R, self.params)
        if dt_in is None:
            self.U0.upload_from_aos(U)
            self.dts = self.adj.forward(self.U0, self.nsteps, self.cfg.cfl, None, self.UN)
            self._dU = (dUr, dUR)
            self._gen = self.adj.generation
            return float(np.sum(self.weights[..., 0] * self.UN.download_to_aos()[..., 0]))
        dt_in = np.asarray(dt_in, dtype=np.float64)
        if dt_in.shape != (self.nsteps,):
            raise ValueError(f"dt_in must hold {self.nsteps} time steps")
        g = Grid2DGPU.from_config(self.cfg)
        tmp, nxt = Grid2DGPU.from_config(self.cfg), Grid2DGPU.from_config(self.cfg)
        try:
            g.upload_from_aos(U)
            apply_boundary_gpu(g, self.bc)
            for dt in dt_in:
                advance_second_order_gpu(g, tmp, nxt, None, float(dt), self.adj.solver, self.bc)
                g.swap(nxt)
            UN = g.download_to_aos()
        finally:
            for x in (g, tmp, nxt):
                x.close()
        return float(np.sum(self.weights[..., 0] * UN[..., 0]))

    def final_state(self):
        return self.UN.download_to_aos()

    def gradient(self):
        """(dJ/drho_b (ny, nx), dJ/dR) at the last objective() call made
        without dt_in.  Raises if the adjoint workspace ran another forward
        since (its stored trajectory is then not that call's)."""
        if self._dU is None:
            raise RuntimeError("gradient(): call objective() without dt_in first")
        if self.adj.generation != self._gen:
            raise RuntimeError("gradient(): the adjoint workspace has run another forward since objective()")
        self.adj.backward(self.lamN, self.lam0)
        lam0 = self.lam0.download_to_aos()[2:-2, 2:-2]
        dUr, dUR = self._dU
        g_rho = np.sum(lam0 * dUr[2:-2, 2:-2], axis=-1)
        g_R = float(np.sum(lam0 * dUR[2:-2, 2:-2]))
        return g_rho, g_R

    def state_adjoint(self):
        """dJ/dU^0 over the interior, (ny, nx, 4), from the last gradient()."""
        return self.lam0.download_to_aos()[2:-2, 2:-2]

    def close(self):
        for g in (self.U0, self.UN, self.lamN, self.lam0):
            g.close()
        self.adj.close()


# ---------------------------------------------------------------------------
# PyTorch binding
# ---------------------------------------------------------------------------
def _torch():
    import torch

    return torch


def smooth_bubble_state_torch(cfg, rho_b, R, params=None):
    """Interior (4, ny, nx) conserved state of the smoothed shock-bubble as a
    differentiable function of the per-cell bubble density rho_b (ny, nx) and
    the radius R (0-d tensor), both float64 on one device."""
    torch = _torch()
    p = params or _cases.SHOCK_BUBBLE_ND
    rho2, u2, p2 = _cases.post_shock_state(p.rho0, p.p0, p.mach)
    dev = rho_b.device
    x = cfg.x_min + (torch.arange(cfg.nx, dtype=torch.float64, device=dev) + 0.5) * cfg.dx
    y = cfg.y_min + (torch.arange(cfg.ny, dtype=torch.float64, device=dev) + 0.5) * cfg.dy
    Y, X = torch.meshgrid(y, x, indexing="ij")
    h = 0.5 * (cfg.dx + cfg.dy)
    behind = X < p.shock_x
    rho_a = torch.where(behind, rho2, p.rho0)
    u = torch.where(behind, u2, 0.0)
    pr = torch.where(behind, p2, p.p0)
    r = torch.sqrt((X - p.bubble_x) ** 2 + (Y - p.bubble_y) ** 2)
    f = torch.clamp((R - r) / h + 0.5, 0.0, 1.0)
    rho = f * rho_b + (1.0 - f) * rho_a
    E = pr / (_cases.GAMMA - 1.0) + 0.5 * rho * (u * u)
    return torch.stack([rho, rho * u, torch.zeros_like(rho), E])


class EulerSolve:
    """torch.autograd.Function over the checkpointed forward/adjoint.

    EulerSolve.apply(U0, adj, nsteps, cfl, dt_in) -> UN, where U0 and UN are
    (4, ny, nx) float64 CUDA tensors of interior conserved states.  The dt
    sequence of the forward call is kept on ``adj``.  One Adjoint holds one
    trajectory: a second apply() on the same Adjoint before the first one's
    backward makes that backward raise instead of reversing the wrong run.
    """

    _fn = None

    @classmethod
    def fn(cls):
        if cls._fn is None:
            torch = _torch()

            class _F(torch.autograd.Function):
                @staticmethod
                def forward(ctx, U0, adj, nsteps, cfl, dt_in):
                    cfg = adj.cfg
                    U0 = U0.detach().contiguous()
                    g0, gN = Grid2DGPU.from_config(cfg), Grid2DGPU.from_config(cfg)
                    stream = torch.cuda.current_stream()
                    torch.cuda.synchronize()
                    g0.copy_interior_from_tensor(U0, stream)
                    adj.forward(g0, nsteps, cfl, dt_in, gN, stream)
                    UN = torch.empty_like(U0)
                    gN.copy_interior_to_tensor(UN, stream)
                    torch.cuda.synchronize()
                    ctx.adj = adj
                    ctx.generation = adj.generation
                    g0.close()
                    gN.close()
                    return UN

                @staticmethod
                def backward(ctx, gUN):
                    adj = ctx.adj
                    if adj.generation != ctx.generation:
                        raise RuntimeError("EulerSolve.backward: the Adjoint workspace ran another forward after this "
                                           "one; its stored trajectory is not this call's (use one Adjoint per "
                                           "pending backward)")
                    cfg = adj.cfg
                    gUN = gUN.detach().contiguous()
                    lamN, lam0 = Grid2DGPU.from_config(cfg), Grid2DGPU.from_config(cfg)
                    stream = torch.cuda.current_stream()
                    torch.cuda.synchronize()
                    zero_grid(lamN, stream)
                    lamN.copy_interior_from_tensor(gUN, stream)
                    adj.backward(lamN, lam0, stream)
                    g0 = torch.empty_like(gUN)
                    lam0.copy_interior_to_tensor(g0, stream)
                    torch.cuda.synchronize()
                    lamN.close()
                    lam0.close()
                    return g0, None, None, None, None

            cls._fn = _F
        return cls._fn

    @classmethod
    def apply(cls, U0, adj, nsteps, cfl, dt_in=None):
        return cls.fn().apply(U0, adj, nsteps, cfl, dt_in)
