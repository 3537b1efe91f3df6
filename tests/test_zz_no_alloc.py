"""mh_riemann_flux_noalloc (staged through the library's fixed device array,
no allocation) against mh_riemann_flux (per-call device buffers) through the
C ABI: bit for bit for every solver and direction, on two calls in a row with
the second smaller (the array keeps the first call's contents), at the
maximum size, and sizes beyond it refused."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from conftest import random_states, same_bits

pytestmark = pytest.mark.gpu

MAX_FACES = 65536


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _flux(name, UL, UR, d, s):
    from large_sys_eq_opt_amd._lib import call

    F = np.full_like(UL, np.nan)
    it = np.full(len(UL), -1, np.int32)
    call(name, len(UL), _p(UL), _p(UR), d, s, _p(F), _p(it))
    return F, it


def _pairs(oracle, rng, n):
    UL = oracle.prim_to_cons(random_states(rng, n))
    UR = oracle.prim_to_cons(random_states(rng, n))
    UL[: n // 50] = UR[: n // 50]               # equal states
    UR[n // 50: n // 25, 0] = -1.0              # unphysical right states: the HLL fallback
    return np.ascontiguousarray(UL), np.ascontiguousarray(UR)


@pytest.mark.parametrize("solver", [0, 1, 2, 3])
@pytest.mark.parametrize("direction", [0, 1])
def test_noalloc_flux_matches_allocating_entry_point(hip, oracle, solver, direction):
    rng = np.random.default_rng(100 + 10 * solver + direction)
    for n in (MAX_FACES, 5000, 37):             # the array keeps the previous call's contents
        UL, UR = _pairs(oracle, rng, n)
        Fa, ia = _flux("mh_riemann_flux", UL, UR, direction, solver)
        Fb, ib = _flux("mh_riemann_flux_noalloc", UL, UR, direction, solver)
        assert same_bits(Fa, Fb), (n, solver, direction)
        assert np.array_equal(ia, ib)


def test_noalloc_flux_limits(hip, oracle):
    from large_sys_eq_opt_amd._lib import MHError, call, load

    rng = np.random.default_rng(7)
    UL, UR = _pairs(oracle, rng, MAX_FACES + 1)
    F = np.empty_like(UL)
    with pytest.raises(MHError, match="MH_NOALLOC_MAX_FACES"):
        call("mh_riemann_flux_noalloc", MAX_FACES + 1, _p(UL), _p(UR), 0, 1, _p(F), None)
    assert load().mh_riemann_flux_noalloc(0, None, None, 0, 1, None, None) == 0
    # iteration counts are optional
    UL, UR = _pairs(oracle, rng, 1000)
    Fa, _ = _flux("mh_riemann_flux", UL, UR, 1, 2)
    call("mh_riemann_flux_noalloc", 1000, _p(UL), _p(UR), 1, 2, _p(F[:1000]), None)
    assert same_bits(F[:1000], Fa)
