"""The C-ABI library loads and exports every symbol include/mh_euler.h declares.

No compute call is made here (this runs without a GPU); only argument
validation that returns before touching the device is exercised.
"""
from __future__ import annotations

import ctypes as C
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "mh_euler.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(mh_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_boundary():
    names = declared_functions()
    for must in ("mh_grid_create", "mh_compute_dt", "mh_advance_second_order", "mh_fill_ghosts",
                 "mh_stepper_advance", "mh_riemann_flux"):
        assert must in names


def test_library_exports_every_declared_symbol(hip):
    from large_sys_eq_opt_amd import _lib

    lib = _lib.load()
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
    # the dynamic symbol table agrees (extern "C": unmangled names)
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r"\b(mh_[a-z0-9_]+)\b", out))
    assert set(declared_functions()) <= exported


def test_binding_covers_header(hip):
    from large_sys_eq_opt_amd import _lib

    assert set(declared_functions()) <= set(_lib.exported_symbols())


def test_invalid_arguments_return_status_not_crash(hip):
    from large_sys_eq_opt_amd import _lib

    lib = _lib.load()
    h = C.c_void_p()
    bad = _lib.GridDesc(0, 10, 0.0, 1.0, 0.0, 1.0)
    assert lib.mh_grid_create(C.byref(bad), C.byref(h)) == _lib.MH_ERR_INVALID
    assert b"nx" in lib.mh_last_error()
    assert lib.mh_grid_create(None, C.byref(h)) == _lib.MH_ERR_INVALID
    assert lib.mh_advance_second_order(None, None, None, 0.1, 1, None, None) == _lib.MH_ERR_INVALID
    assert lib.mh_riemann_flux(-1, None, None, 0, 0, None, None) == _lib.MH_ERR_INVALID
    assert lib.mh_riemann_flux(1, None, None, 0, 9, None, None) == _lib.MH_ERR_INVALID
    for bad_shape in ((-1, 0, 1, 0, 1, 1, 0), (0, 0, 3, 0, 1, 1, 0), (0, 0, 1, 3, 1, 1, 0), (0, 0, 1, 0, 2, 1, 0),
                      (0, 0, 1, 0, 1, 1, -1), (0, 0, 1, 0, 1, 1, 0, 3), (0, 0, 1, 0, 1, 1, 0, -1),
                      (0, 0, 1, 0, 1, 1, 0, 1, -1)):
        assert lib.mh_set_launch_shape(C.byref(_lib.LaunchShape(*bad_shape))) == _lib.MH_ERR_INVALID
    assert lib.mh_set_launch_shape(None) == _lib.MH_ERR_INVALID
    x = (C.c_double * 4)()
    assert lib.mh_fast_ops(1, x, x, x, 3, x, x, x) == _lib.MH_ERR_INVALID
    assert lib.mh_fast_ops(1, None, x, x, 1, x, x, x) == _lib.MH_ERR_INVALID
    assert lib.mh_version().startswith(b"mh_euler")


def test_python_raises_mherror(hip):
    from large_sys_eq_opt_amd import MHError, _lib

    with pytest.raises(MHError) as e:
        _lib.call("mh_grid_create", None, None)
    assert e.value.code == _lib.MH_ERR_INVALID
