"""Error behaviour through the C ABI on live device grids: what the
reference's exits and exceptions become (src/gpu/solver_gpu.cu:41-49,
src/cpu/solver_cpu.cpp:304-306) -- a status code and a message, the device
left usable."""
from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _still(nx, ny):
    """At rest with p = 0: every signal speed is 0 (no finite dt)."""
    U = np.zeros((ny + 4, nx + 4, 4))
    U[..., 0] = 1.0
    return U


def _state(nx, ny):
    U = np.zeros((ny + 4, nx + 4, 4))
    U[..., 0] = 1.0
    U[..., 3] = 2.5
    return U


def test_compute_dt_zero_wavespeed(hip):
    from large_sys_eq_opt_amd import MHError, _lib, compute_dt_gpu

    g = hip.Grid2DGPU(32, 24)
    g.upload_from_aos(_still(32, 24))
    with pytest.raises(MHError) as e:
        compute_dt_gpu(g, None, 0.8)
    assert e.value.code == _lib.MH_ERR_WAVESPEED and "speed" in str(e.value)
    g.upload_from_aos(_state(32, 24))   # the grid and the device stay usable
    assert compute_dt_gpu(g, None, 0.8) > 0.0


def test_stepper_reports_zero_wavespeed(hip):
    from large_sys_eq_opt_amd import MHError, _lib

    g = hip.Grid2DGPU(32, 24)
    g.upload_from_aos(_still(32, 24))
    st = hip.Stepper(g, 1, 0.8, 0.0, None, log_capacity=4)
    st.advance(3)
    with pytest.raises(MHError) as e:
        st.clock()
    assert e.value.code == _lib.MH_ERR_WAVESPEED
    st.close()
    g.upload_from_aos(_state(32, 24))
    st = hip.Stepper(g, 1, 0.8, 0.0, None, log_capacity=4)
    st.advance(2)
    _, n, dts = st.clock()
    assert n == 2 and (dts > 0).all()
    st.close()


def test_bad_arguments_on_live_grids(hip):
    """Status codes straight from the C ABI (the Python layer would raise
    ValueError on an unknown solver name or number before calling it)."""
    import ctypes as C

    from large_sys_eq_opt_amd import _lib, advance_second_order_gpu

    lib = _lib.load()
    U, T, N = (hip.Grid2DGPU(32, 24) for _ in range(3))
    small = hip.Grid2DGPU(30, 24)
    for g in (U, T, N):
        g.upload_from_aos(_state(32, 24))
    h = C.c_void_p()
    statuses = [
        lib.mh_advance_second_order(U.handle, T.handle, small.handle, 1e-3, 1, None, None),   # shape mismatch
        lib.mh_advance_second_order(U.handle, U.handle, N.handle, 1e-3, 1, None, None),       # Utmp aliases Uold
        lib.mh_advance_second_order(U.handle, T.handle, N.handle, 1e-3, 7, None, None),       # no such solver
        lib.mh_advance_first_order(U.handle, small.handle, 1e-3, 1, None, None),
        lib.mh_sweep(U.handle, T.handle, 1e-3, 2, 1, None),                                  # no such direction
        lib.mh_stepper_create(U.handle, 9, 0.8, 0.0, None, 0, C.byref(h), None),
        lib.mh_stepper_create(U.handle, 1, -0.5, 0.0, None, 0, C.byref(h), None),             # cfl <= 0
    ]
    assert statuses == [_lib.MH_ERR_INVALID] * len(statuses), statuses
    assert lib.mh_last_error()
    advance_second_order_gpu(U, T, N, None, 1e-3, 1)   # still usable
    hip.synchronize()
    assert np.isfinite(N.download_to_aos()).all()
