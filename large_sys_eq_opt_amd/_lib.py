"""ctypes binding of ``libmh_euler.so`` (the C ABI declared in include/mh_euler.h).

There is no fallback: if the HIP library is missing or fails to load, every
entry point raises ``MHError``.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmh_euler.so")

MH_OK, MH_ERR_INVALID, MH_ERR_OOM, MH_ERR_LAUNCH, MH_ERR_WAVESPEED, MH_ERR_NODEVICE = range(6)
_P = C.c_void_p
_I = C.c_int
_D = C.c_double
_I64 = C.c_int64


class MHError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[mh status {code}] {msg}")
        self.code = code


class GridDesc(C.Structure):
    _fields_ = [("nx", _I), ("ny", _I), ("x_min", _D), ("x_max", _D), ("y_min", _D), ("y_max", _D)]


class LaunchShape(C.Structure):
    _fields_ = [("seg_x", _I), ("seg_y", _I), ("y_waves", _I), ("x_layout", _I), ("x_wpb", _I), ("y_wpb", _I),
                ("min_tiles", _I), ("fused", _I), ("fused_tail", _I)]


class BcDesc(C.Structure):
    _fields_ = [("kind", _I * 4), ("state", (_D * 4) * 4)]


# Every exported symbol with its argtypes (restype int unless noted).
SIGNATURES = {
    "mh_last_error": ([], C.c_char_p),
    "mh_version": ([], C.c_char_p),
    "mh_set_device": ([_I], _I),
    "mh_grid_create": ([C.POINTER(GridDesc), C.POINTER(_P)], _I),
    "mh_grid_destroy": ([_P], _I),
    "mh_grid_info": ([_P, C.POINTER(GridDesc), C.POINTER(_D), C.POINTER(_D), C.POINTER(_I64), C.POINTER(_I64)], _I),
    "mh_grid_plane_ptr": ([_P, _I, C.POINTER(_P)], _I),
    "mh_grid_upload_aos": ([_P, _P], _I),
    "mh_grid_download_aos": ([_P, _P], _I),
    "mh_grid_upload_soa": ([_P, _P * 4], _I),
    "mh_grid_download_soa": ([_P, _P * 4], _I),
    "mh_grid_copy_interior_from_device": ([_P, _P, _P], _I),
    "mh_grid_copy_interior_to_device": ([_P, _P, _P], _I),
    "mh_grid_swap": ([_P, _P], _I),
    "mh_fill_ghosts": ([_P, C.POINTER(BcDesc), _P], _I),
    "mh_compute_dt": ([_P, _D, C.POINTER(_D), C.POINTER(_D), _P], _I),
    "mh_advance_second_order": ([_P, _P, _P, _D, _I, C.POINTER(BcDesc), _P], _I),
    "mh_advance_first_order": ([_P, _P, _D, _I, C.POINTER(BcDesc), _P], _I),
    "mh_sweep": ([_P, _P, _D, _I, _I, _P], _I),
    "mh_stepper_create": ([_P, _I, _D, _D, C.POINTER(BcDesc), _I, C.POINTER(_P), _P], _I),
    "mh_stepper_destroy": ([_P], _I),
    "mh_stepper_advance": ([_P, _I, _P], _I),
    "mh_stepper_clock": ([_P, C.POINTER(_D), C.POINTER(_I), _P, _P], _I),
    "mh_stepper_finish": ([_P, _P], _I),
    "mh_stepper_set_stop": ([_P, _D, _P], _I),
    "mh_init_state": ([_P, _I, _P, _P], _I),
    "mh_stepper_advance_profiled": ([_P, _I, _P, C.POINTER(_D), C.POINTER(_D)], _I),
    "mh_set_launch_shape": ([C.POINTER(LaunchShape)], _I),
    "mh_get_launch_shape": ([C.POINTER(LaunchShape)], _I),
    "mh_batch_create": ([_I, _P, _I, _D, _P, C.POINTER(_P)], _I),
    "mh_batch_destroy": ([_P], _I),
    "mh_batch_grid": ([_P, _I, C.POINTER(_P)], _I),
    "mh_batch_prepare": ([_P, _P], _I),
    "mh_batch_set_schedule": ([_P, _I, _I, _I, _I, _I], _I),
    "mh_batch_set_size_class": ([_P, _I, _I, _I], _I),
    "mh_batch_step": ([_P, _I, _P], _I),
    "mh_batch_capture": ([_P, _I, _P], _I),
    "mh_batch_clock": ([_P, _P, _P, _P, _P, _P], _I),
    "mh_batch_sync": ([_P, _P], _I),
    "mh_batch_info": ([_P, C.POINTER(_I), C.POINTER(_I), C.POINTER(_I)], _I),
    "mh_adjoint_create": ([C.POINTER(GridDesc), _I, C.POINTER(BcDesc), _I, _I, C.POINTER(_P)], _I),
    "mh_adjoint_destroy": ([_P], _I),
    "mh_adjoint_step_vjp": ([_P, _P, _D, _P, _P, _P], _I),
    "mh_adjoint_sweep_vjp": ([_P, _P, _D, _I, _P, _P, _P], _I),
    "mh_adjoint_forward": ([_P, _P, _I, _D, _P, _P, _P, _P], _I),
    "mh_adjoint_backward": ([_P, _P, _P, _P], _I),
    "mh_adjoint_dt": ([_P, _P, C.POINTER(_I)], _I),
    "mh_snapshot_create": ([_P, _I, C.POINTER(_P)], _I),
    "mh_snapshot_destroy": ([_P], _I),
    "mh_snapshot_capture": ([_P, _P, _I, _P], _I),
    "mh_snapshot_wait": ([_P, _I, _P * 4], _I),
    "mh_snapshot_write_csv": ([_P, _I, C.c_char_p, _I, _I], _I),
    "mh_write_csv_aos": ([_P, C.POINTER(GridDesc), _I, _I, C.c_char_p], _I),
    "mh_grid_zero": ([_P, _P], _I),
    "mh_grid_copy": ([_P, _P, _P], _I),
    "mh_copy": ([_P, _P, _I64, _P], _I),
    "mh_device_alloc": ([_I64, C.POINTER(_P)], _I),
    "mh_device_free": ([_P], _I),
    "mh_synchronize": ([_P], _I),
    "mh_fast_ops": ([_I, _P, _P, _P, _I, _P, _P, _P], _I),
    "mh_riemann_flux": ([_I, _P, _P, _I, _I, _P, _P], _I),
    "mh_riemann_flux_noalloc": ([_I, _P, _P, _I, _I, _P, _P], _I),
    "mh_reconstruct": ([_I, _P, _P, _P, _D, _I, _P, _P], _I),
}

_lib = None


def load(path: str | None = None):
    """Loads the HIP library once (``MH_EULER_LIB`` names another build of
    it); raises MHError if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    path = path or os.environ.get("MH_EULER_LIB") or LIB_PATH
    if not os.path.exists(path):
        raise MHError(MH_ERR_NODEVICE, f"HIP library not built: {path} (run python -m large_sys_eq_opt_amd.build)")
    # One HIP runtime per process: torch ships its own libamdhip64 (soname
    # libamdhip64.so.7, loaded by the name libamdhip64.so).  Loaded first, it
    # is the runtime this library binds to; loaded after this library it
    # would come in as a second runtime that finds no GPU.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = C.CDLL(path)
    for name, (args, res) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    _lib = lib
    return lib


def call(name: str, *args) -> None:
    """Calls a C-ABI entry point and raises MHError on a non-zero status."""
    lib = load()
    rc = getattr(lib, name)(*args)
    if rc != MH_OK:
        raise MHError(rc, lib.mh_last_error().decode())


def exported_symbols():
    return list(SIGNATURES)
