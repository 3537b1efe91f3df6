"""The C++ code the integration notes give a maintainer (the reference's Grid2DGPU /
compute_dt_gpu / advance_*_gpu API over libmh_euler.so, and the
device-resident loop) compiles against include/mh_euler.h and the
reference's own headers (head/types.hpp, head/riemann.hpp) and links
against the library; the driver built on it (tests/integration_driver.py)
is what tests/test_gpu_integration.py runs on the GPU.  Nothing runs here
(no GPU); skipped where the reference tree is absent."""
from __future__ import annotations

import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_HEAD = "/root/reference/head"


def _blocks():
    with open(os.path.join(ROOT, "INTEGRATION.md")) as f:
        return re.findall(r"```cpp\n(.*?)```", f.read(), re.S)


@pytest.mark.skipif(not os.path.isdir(REF_HEAD) or shutil.which("g++") is None,
                    reason="needs the reference headers and g++")
def test_integration_cpp_compiles_and_links(tmp_path):
    lib = os.path.join(ROOT, "large_sys_eq_opt_amd", "libmh_euler.so")
    if not os.path.exists(lib):
        pytest.skip("library not built")
    backend, loop = _blocks()[:2]
    (tmp_path / "mh_backend.hpp").write_text(backend)
    (tmp_path / "main.cpp").write_text(f"""#include <cstdio>
#include <vector>
#include "mh_backend.hpp"
static void device_loop(Grid2DGPU& U, double cfl, double t_end, void* stream) {{
{loop}
}}
int main() {{
    Grid2DGPU U(64, 32, 2, 0.0, 2.0, 0.0, 1.0), T(64, 32, 2, 0.0, 2.0, 0.0, 1.0), N(64, 32, 2, 0.0, 2.0, 0.0, 1.0);
    std::vector<Conserved> h((64 + 4) * (32 + 4));
    U.upload_from_aos(h);
    GpuWorkspace ws;
    init_gpu_workspace(ws, U);
    apply_transmissive_boundary_gpu(U);
    const double dt = compute_dt_gpu(U, ws, 0.8);
    advance_second_order_gpu(U, T, N, ws, dt, RiemannSolver::HLLC);
    advance_first_order_gpu(U, N, dt, RiemannSolver::HLL);
    U.swap(N);
    U.download_to_aos(h);
    free_gpu_workspace(ws);
    device_loop(U, 0.8, 0.25, nullptr);
    std::puts("ok");
    return 0;
}}
""")
    cmd = ["g++", "-std=c++17", "-O1", str(tmp_path / "main.cpp"), "-I", os.path.join(ROOT, "include"),
           "-I", REF_HEAD, "-L", os.path.dirname(lib), "-lmh_euler", "-Wl,-rpath," + os.path.dirname(lib),
           "-o", str(tmp_path / "main")]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]


@pytest.mark.skipif(not os.path.isdir(REF_HEAD) or shutil.which("g++") is None,
                    reason="needs the reference headers and g++")
def test_integration_driver_builds():
    import integration_driver as drv

    if not os.path.exists(os.path.join(ROOT, "large_sys_eq_opt_amd", "libmh_euler.so")):
        pytest.skip("library not built")
    exe = drv.build()
    assert exe and os.access(exe, os.X_OK)
