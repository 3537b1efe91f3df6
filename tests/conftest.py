"""Shared fixtures.  GPU tests are marked ``gpu``; everything else runs on the CPU.

The oracle (oracle/_build/liboracle.so) and, where the reference tree was
present at build time, the compiled reference (oracle/_ref/libmuscl_ref.so)
are the checkers.  The product under test is libmh_euler.so.
"""
from __future__ import annotations

import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import pyoracle  # noqa: E402


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and libmh_euler.so")
    config.addinivalue_line("markers", "slow: long CPU run (full reference step counts)")


@pytest.fixture(scope="session")
def oracle():
    if not os.path.exists(pyoracle.ORACLE_SO):
        pyoracle.build(reference=False)
    return pyoracle.Oracle()


@pytest.fixture(scope="session")
def reference():
    if not pyoracle.reference_available():
        pytest.skip("reference harness not built (needs the reference tree at build time)")
    return pyoracle.Reference()


@pytest.fixture(scope="session")
def hip():
    """The HIP library; fails (does not skip) when it cannot be loaded."""
    import large_sys_eq_opt_amd as m

    m.load_library()
    return m


def interior(U):
    return U[2:-2, 2:-2]


def same_bits(a, b) -> bool:
    """Bit-for-bit equality of two float64 arrays (so +0 and -0 differ).
    NaN payloads are not portable between the GPU and the host libm, so
    NaN entries only have to be NaN on both sides."""
    a, b = np.ascontiguousarray(a, dtype=np.float64), np.ascontiguousarray(b, dtype=np.float64)
    if a.shape != b.shape:
        return False
    na, nb = np.isnan(a), np.isnan(b)
    if not np.array_equal(na, nb):
        return False
    return bool(np.array_equal(a[~na].view(np.uint64), b[~nb].view(np.uint64)))


def rel_linf(a, b):
    """max |a-b| / max |b| per conserved component, maximised."""
    a, b = interior(a), interior(b)
    out = 0.0
    for c in range(4):
        den = np.abs(b[..., c]).max()
        num = np.abs(a[..., c] - b[..., c]).max()
        out = max(out, num / den if den > 0 else num)
    return out


def random_states(rng, n, rho=(0.1, 2.0), vel=(-1.5, 1.5), p=(0.05, 5.0)):
    """n random physical primitive states (rho, u, v, p)."""
    V = np.empty((n, 4))
    V[:, 0] = rng.uniform(*rho, n)
    V[:, 1] = rng.uniform(*vel, n)
    V[:, 2] = rng.uniform(*vel, n)
    V[:, 3] = rng.uniform(*p, n)
    return V


def smooth_state(oracle, nx, ny, seed=0):
    """A smooth, everywhere non-uniform state: no limiter slope is exactly
    zero, so the scheme is differentiable there (on piecewise-constant data
    minmod sits at its kink and central differences see a one-sided slope)."""
    rng = np.random.default_rng(seed)
    y, x = (np.mgrid[0:ny + 4, 0:nx + 4] + 0.5) / np.array([ny + 4, nx + 4])[:, None, None]
    ph = rng.uniform(0, 2 * np.pi, 6)
    rho = 1.0 + 0.3 * np.sin(2 * np.pi * x + ph[0]) * np.cos(2 * np.pi * y + ph[1])
    u = 0.3 * np.sin(2 * np.pi * (y + 0.37 * x) + ph[2])
    v = 0.2 * np.cos(2 * np.pi * (x - 0.41 * y) + ph[3])
    p = 1.0 + 0.4 * np.cos(2 * np.pi * (x + 0.83 * y) + ph[4])
    V = np.stack([rho, u, v, p], axis=-1).reshape(-1, 4)
    return oracle.prim_to_cons(V).reshape(ny + 4, nx + 4, 4)
