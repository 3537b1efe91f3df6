"""Builds the reference-side C++ driver of the integration notes (the ```cpp
blocks of the notes file at the repository root): the binding header given
there (mh_backend.hpp: the reference's Grid2DGPU / compute_dt_gpu /
advance_*_gpu API over libmh_euler.so) and a main() with the loop of the
reference's GPU driver (scripts/gpu/main_gpu.cu:225-299: compute_dt_gpu,
clamp to t_end, advance_second_order_gpu, swap, t += dt, a snapshot download
at the first step with t >= k t_end / nsnap), followed by the
device-resident loop block of the notes from the same initial state.

The header includes the reference's own head/types.hpp and head/riemann.hpp,
so the driver is built where the reference tree is (g++, no HIP compiler)
into tests/_build/mh_main_gpu, which travels with the tree; the GPU test
runs it.

Driver I/O (raw little-endian float64 in the reference AoS layout, ghosts
included): argv = in.bin nx ny x_min x_max y_min y_max cfl t_end solver nsnap
out_prefix; writes <out>_snap<k>.bin, <out>_final.bin, <out>_dt.bin and
<out>_device_final.bin, and prints "steps t device_steps device_t".
"""
from __future__ import annotations

import os
import re
import shutil
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_HEAD = "/root/reference/head"
OUT_DIR = os.path.join(ROOT, "tests", "_build")
DRIVER = os.path.join(OUT_DIR, "mh_main_gpu")

MAIN = r"""#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>
#include "mh_backend.hpp"

static bool write_bin(const std::string& path, const void* p, size_t bytes) {
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) return false;
    const bool ok = std::fwrite(p, 1, bytes, f) == bytes;
    return std::fclose(f) == 0 && ok;
}

static void device_loop(Grid2DGPU& U, double cfl, double t_end, void* stream, double& t_out, int& steps_out) {
@@DEVICE_LOOP@@
    t_out = t;
    steps_out = steps;
}

int main(int argc, char** argv) {
    if (argc != 13) { std::fprintf(stderr, "usage: in.bin nx ny xmin xmax ymin ymax cfl t_end solver nsnap out\n"); return 2; }
    const int nx = std::atoi(argv[2]), ny = std::atoi(argv[3]);
    const double x_min = std::atof(argv[4]), x_max = std::atof(argv[5]);
    const double y_min = std::atof(argv[6]), y_max = std::atof(argv[7]);
    const double cfl = std::atof(argv[8]), t_end = std::atof(argv[9]);
    const RiemannSolver solver = static_cast<RiemannSolver>(std::atoi(argv[10]));
    const int num_snapshots = std::atoi(argv[11]);
    const std::string out = argv[12];
    std::vector<Conserved> host_init(size_t(nx + 4) * (ny + 4));
    FILE* f = std::fopen(argv[1], "rb");
    if (!f || std::fread(host_init.data(), sizeof(Conserved), host_init.size(), f) != host_init.size()) return 3;
    std::fclose(f);
    try {
        // scripts/gpu/main_gpu.cu:199-223
        Grid2DGPU d_U(nx, ny, 2, x_min, x_max, y_min, y_max);
        Grid2DGPU d_U_mid(nx, ny, 2, x_min, x_max, y_min, y_max);
        Grid2DGPU d_U_next(nx, ny, 2, x_min, x_max, y_min, y_max);
        GpuWorkspace ws;
        init_gpu_workspace(ws, d_U);
        d_U.upload_from_aos(host_init);
        double t = 0.0;
        int step = 0;
        std::vector<double> snapshot_times, dts;
        for (int k = 1; k <= num_snapshots; ++k)
            snapshot_times.push_back(t_end * static_cast<double>(k) / static_cast<double>(num_snapshots));
        int next_snapshot = 0;
        // scripts/gpu/main_gpu.cu:243-299
        while (t < t_end) {
            double dt = compute_dt_gpu(d_U, ws, cfl);
            if (t + dt > t_end) dt = t_end - t;
            advance_second_order_gpu(d_U, d_U_mid, d_U_next, ws, dt, solver);
            d_U.swap(d_U_next);
            t += dt;
            ++step;
            dts.push_back(dt);
            while (next_snapshot < num_snapshots && t >= snapshot_times[next_snapshot]) {
                std::vector<Conserved> host_snapshot;
                d_U.download_to_aos(host_snapshot);
                if (!write_bin(out + "_snap" + std::to_string(next_snapshot + 1) + ".bin", host_snapshot.data(),
                               host_snapshot.size() * sizeof(Conserved)))
                    return 4;
                ++next_snapshot;
            }
        }
        std::vector<Conserved> fin;
        d_U.download_to_aos(fin);
        if (!write_bin(out + "_final.bin", fin.data(), fin.size() * sizeof(Conserved))) return 4;
        if (!write_bin(out + "_dt.bin", dts.data(), dts.size() * sizeof(double))) return 4;
        free_gpu_workspace(ws);
        // the device-resident loop from the same initial state
        Grid2DGPU U(nx, ny, 2, x_min, x_max, y_min, y_max);
        U.upload_from_aos(host_init);
        double td = 0.0;
        int sd = 0;
        device_loop(U, cfl, t_end, nullptr, td, sd);
        U.download_to_aos(fin);
        if (!write_bin(out + "_device_final.bin", fin.data(), fin.size() * sizeof(Conserved))) return 4;
        std::printf("%d %.17g %d %.17g\n", step, t, sd, td);
    } catch (const std::exception& e) {
        std::fprintf(stderr, "[GPU] Error: %s\n", e.what());
        return 1;
    }
    return 0;
}
"""


def blocks():
    """The ```cpp blocks of the integration notes: the binding header, then
    the device-resident loop."""
    with open(os.path.join(ROOT, "INTEGRATION.md")) as f:
        return re.findall(r"```cpp\n(.*?)```", f.read(), re.S)


def build(out_dir: str = OUT_DIR) -> str | None:
    """Compiles the driver; returns its path, or None where the reference
    headers or g++ are absent (then nothing is built)."""
    lib = os.path.join(ROOT, "large_sys_eq_opt_amd", "libmh_euler.so")
    if not os.path.isdir(REF_HEAD) or shutil.which("g++") is None or not os.path.exists(lib):
        return None
    backend, loop = blocks()[:2]
    os.makedirs(out_dir, exist_ok=True)
    with open(os.path.join(out_dir, "mh_backend.hpp"), "w") as f:
        f.write(backend)
    with open(os.path.join(out_dir, "main_gpu.cpp"), "w") as f:
        f.write(MAIN.replace("@@DEVICE_LOOP@@", loop))
    exe = os.path.join(out_dir, "mh_main_gpu")
    # the library is found next to the tree at run time (rpath relative to
    # the executable), so the built driver works from any copy of the tree
    cmd = ["g++", "-std=c++17", "-O2", os.path.join(out_dir, "main_gpu.cpp"), "-I", os.path.join(ROOT, "include"),
           "-I", REF_HEAD, "-I", out_dir, "-L", os.path.dirname(lib), "-lmh_euler",
           "-Wl,-rpath,$ORIGIN/../../large_sys_eq_opt_amd", "-o", exe + ".tmp"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("integration driver build failed:\n" + r.stderr[-3000:])
    os.replace(exe + ".tmp", exe)
    return exe
