"""MI355X-native MUSCL-Hancock Euler solver (FP64, gfx950 HIP kernels).

The numerical path runs in libmh_euler.so (built from csrc/ by
``large_sys_eq_opt_amd.build``); this package is its Python face:
``solver`` mirrors the reference's GPU solver API, ``cases`` its test-case
API, ``driver`` its time loop and snapshot output, ``matrix`` its benchmark
matrix.
"""
from ._lib import MHError, load as load_library  # noqa: F401
from .cases import (CaseConfig, CaseId, case_boundary, get_case_config, get_n_case_config,  # noqa: F401
                    initial_state_at, make_initial_grid, make_n_grid, parse_case_id)
from .solver import (Boundary, GpuWorkspace, Grid2DGPU, RiemannSolver, Snapshot, Stepper,  # noqa: F401
                     advance_first_order_gpu, advance_second_order_gpu, apply_boundary_gpu,
                     apply_transmissive_boundary_gpu, compute_dt_gpu, parse_riemann_solver, sweep_gpu, synchronize)

__version__ = "1.0.0"
