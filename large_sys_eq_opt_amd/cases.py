"""Test-case setup: the reference's case API plus the benchmark cases.

Mirrors head/test_cases.hpp:8-36 and src/test_cases.cpp:11-199
(``CaseConfig``, ``CaseId``, ``parse_case_id``, ``get_case_config``,
``get_n_case_config``, ``initial_state_at``) and head/init.hpp / src/init.cpp
(``make_initial_grid``, ``make_n_grid``, ``initialise_grid``: point sampling
at every cell centre, ghosts included).  Host arrays are numpy float64 of
shape (ny+4, nx+4, 4) -- the reference's AoS ``Conserved`` layout.

Cases
  shock_bubble      reference air/helium shock-bubble, SI units, 500x197
  blast_wave        reference blast, 500x500 on [0,1]^2
  cyl_blast         cylindrical blast on [0,2]^2: rho=1, p=1 for r <= 0.4 of
                    (1,1), rho=0.125, p=0.1 outside; 256^2, CFL 0.8, t_end 0.25
  shock_bubble_nd   nondimensional shock-bubble: quiescent rho=1, p=1; Mach
                    1.22 shock at x=0.1 with the Rankine-Hugoniot post-shock
                    state; bubble rho=0.138, r=0.2 at (0.5,0.5); [0,2]x[0,1],
                    1024x512, inflow (post-shock state) on the left, t_end 0.4
  shock_bubble_adj  shock_bubble_nd at 512x256 (the adjoint case)
"""
from __future__ import annotations

import dataclasses
import enum
import math

import numpy as np

GAMMA = 1.4  # head/physics.hpp:11


class CaseId(enum.IntEnum):
    ShockBubble = 0
    BlastWave = 1
    CylBlast = 2
    ShockBubbleND = 3
    ShockBubbleAdj = 4


_NAMES = {
    "shock_bubble": CaseId.ShockBubble,
    "blast_wave": CaseId.BlastWave,
    "cyl_blast": CaseId.CylBlast,
    "shock_bubble_nd": CaseId.ShockBubbleND,
    "shock_bubble_adj": CaseId.ShockBubbleAdj,
}


@dataclasses.dataclass
class CaseConfig:  # head/test_cases.hpp:8-18
    nx: int
    ny: int
    ng: int
    x_min: float
    x_max: float
    y_min: float
    y_max: float
    cfl: float
    t_end: float

    @property
    def dx(self) -> float:
        return (self.x_max - self.x_min) / float(self.nx)

    @property
    def dy(self) -> float:
        return (self.y_max - self.y_min) / float(self.ny)

    @property
    def bounds(self):
        return (self.x_min, self.x_max, self.y_min, self.y_max)


def parse_case_id(name: str) -> CaseId:  # src/test_cases.cpp:105-118
    if isinstance(name, CaseId):
        return name
    try:
        return _NAMES[name]
    except KeyError:
        raise ValueError(f"Unknown case name: {name}. Supported cases are: {', '.join(_NAMES)}.") from None


def case_id_to_string(case_id: CaseId) -> str:  # src/test_cases.cpp:120-130
    for k, v in _NAMES.items():
        if v == case_id:
            return k
    raise ValueError(f"Unhandled CaseId {case_id}")


def get_case_config(case) -> CaseConfig:  # src/test_cases.cpp:132-162
    c = parse_case_id(case)
    if c == CaseId.ShockBubble:
        return CaseConfig(500, 197, 2, 0.0, 0.225, 0.0, 0.089, 0.4, 0.0011741)
    if c == CaseId.BlastWave:
        return CaseConfig(500, 500, 2, 0.0, 1.0, 0.0, 1.0, 0.4, 0.2)
    if c == CaseId.CylBlast:
        return CaseConfig(256, 256, 2, 0.0, 2.0, 0.0, 2.0, 0.8, 0.25)
    if c == CaseId.ShockBubbleND:
        return CaseConfig(1024, 512, 2, 0.0, 2.0, 0.0, 1.0, 0.8, 0.4)
    return CaseConfig(512, 256, 2, 0.0, 2.0, 0.0, 1.0, 0.8, 0.0)


def get_n_case_config(case, n: int) -> CaseConfig:  # src/test_cases.cpp:184-195
    if n < 1:
        raise ValueError("get_n_case_config: n must be >= 1.")
    cfg = get_case_config(case)
    cfg.nx *= n
    cfg.ny *= n
    return cfg


# ---------------------------------------------------------------------------
# case geometry as parameters
# ---------------------------------------------------------------------------
@dataclasses.dataclass
class BlastParams:
    x0: float
    y0: float
    r0: float
    rho_in: float
    p_in: float
    rho_out: float
    p_out: float

    def array(self):
        return np.array([self.x0, self.y0, self.r0, self.rho_in, self.p_in, self.rho_out, self.p_out])


@dataclasses.dataclass
class ShockBubbleParams:
    rho0: float
    p0: float
    mach: float
    shock_x: float
    bubble_rho: float
    bubble_x: float
    bubble_y: float
    bubble_r: float

    def array(self):
        return np.array([self.rho0, self.p0, self.mach, self.shock_x, self.bubble_rho, self.bubble_x,
                         self.bubble_y, self.bubble_r])


REFERENCE_SHOCK_BUBBLE = ShockBubbleParams(1.29, 1.01325e5, 1.22, 0.005, 0.214, 0.035, 0.0445, 0.025)
REFERENCE_BLAST = BlastParams(0.5, 0.5, 0.1, 1.0, 100.0, 1.0, 1.0)
CYL_BLAST = BlastParams(1.0, 1.0, 0.4, 1.0, 1.0, 0.125, 0.1)
SHOCK_BUBBLE_ND = ShockBubbleParams(1.0, 1.0, 1.22, 0.1, 0.138, 0.5, 0.5, 0.2)


def case_params(case):
    """(kind, params) of a case for mh_init_state: kind 2 blast, 3 shock-bubble."""
    c = parse_case_id(case)
    if c == CaseId.ShockBubble:
        return 3, REFERENCE_SHOCK_BUBBLE
    if c == CaseId.BlastWave:
        return 2, REFERENCE_BLAST
    if c == CaseId.CylBlast:
        return 2, CYL_BLAST
    return 3, SHOCK_BUBBLE_ND


def post_shock_state(rho0: float, p0: float, mach: float):
    """Rankine-Hugoniot post-shock (rho2, u2, p2) into gas at rest
    (src/test_cases.cpp:35-47, same operation order)."""
    a0 = math.sqrt(GAMMA * p0 / rho0)
    M2 = mach * mach
    rho2 = rho0 * ((GAMMA + 1.0) * M2) / ((GAMMA - 1.0) * M2 + 2.0)
    p2 = p0 * (1.0 + 2.0 * GAMMA / (GAMMA + 1.0) * (M2 - 1.0))
    u2 = 2.0 * a0 / (GAMMA + 1.0) * (mach - 1.0 / mach)
    return rho2, u2, p2


def make_conserved(rho, u, v, p):  # src/test_cases.cpp:11-18
    rho, u, v, p = (np.asarray(a, dtype=np.float64) for a in (rho, u, v, p))
    return np.stack([rho, rho * u, rho * v, p / (GAMMA - 1.0) + 0.5 * rho * (u * u + v * v)], axis=-1)


def blast_state(b: BlastParams, x, y):  # src/test_cases.cpp:78-101
    dx = x - b.x0
    dy = y - b.y0
    inside = (dx * dx + dy * dy) <= b.r0 * b.r0
    rho = np.where(inside, b.rho_in, b.rho_out)
    p = np.where(inside, b.p_in, b.p_out)
    z = np.zeros_like(rho)
    return make_conserved(rho, z, z, p)


def shock_bubble_state(s: ShockBubbleParams, x, y):  # src/test_cases.cpp:24-69
    rho2, u2, p2 = post_shock_state(s.rho0, s.p0, s.mach)
    x = np.asarray(x, dtype=np.float64)
    y = np.asarray(y, dtype=np.float64)
    behind = x < s.shock_x
    rho = np.where(behind, rho2, s.rho0)
    u = np.where(behind, u2, 0.0)
    p = np.where(behind, p2, s.p0)
    dx = x - s.bubble_x
    dy = y - s.bubble_y
    rho = np.where(dx * dx + dy * dy <= s.bubble_r * s.bubble_r, s.bubble_rho, rho)
    return make_conserved(rho, u, np.zeros_like(u), p)


def initial_state_at(case, x, y):  # src/test_cases.cpp:168-182
    kind, p = case_params(case)
    return blast_state(p, x, y) if kind == 2 else shock_bubble_state(p, x, y)


def cell_centers(cfg: CaseConfig):
    """Cell-centre coordinates of every cell, ghosts included (src/init.cpp:8-14)."""
    i = np.arange(cfg.nx + 2 * cfg.ng)
    j = np.arange(cfg.ny + 2 * cfg.ng)
    x = cfg.x_min + ((i - cfg.ng).astype(np.float64) + 0.5) * cfg.dx
    y = cfg.y_min + ((j - cfg.ng).astype(np.float64) + 0.5) * cfg.dy
    return np.meshgrid(x, y)


def initialise_grid(case, cfg: CaseConfig):  # src/init.cpp:29-38
    X, Y = cell_centers(cfg)
    return initial_state_at(case, X, Y)


def make_initial_grid(case):
    cfg = get_case_config(case)
    return initialise_grid(case, cfg), cfg


def make_n_grid(case, n: int):
    cfg = get_n_case_config(case, n)
    return initialise_grid(case, cfg), cfg


def case_boundary(case):
    """Boundary spec of a case: shock_bubble_nd / _adj have a Dirichlet
    inflow of the post-shock state on the left; everything else is
    transmissive."""
    from .solver import Boundary

    c = parse_case_id(case)
    if c in (CaseId.ShockBubbleND, CaseId.ShockBubbleAdj):
        rho2, u2, p2 = post_shock_state(SHOCK_BUBBLE_ND.rho0, SHOCK_BUBBLE_ND.p0, SHOCK_BUBBLE_ND.mach)
        return Boundary.inflow_left(make_conserved(rho2, u2, 0.0, p2))
    return Boundary.transmissive()
