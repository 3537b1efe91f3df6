// grid.h -- device layout of a grid and the host-side handle shared by the
// HIP translation units of the library.
//
// A grid is four FP64 planes (rho, rho*u, rho*v, E) in one allocation.  Each
// row holds the nx interior cells and two ghost cells per side; rows are
// padded to a pitch that is a multiple of 16 doubles with a 16-double lead,
// so interior column 0 of every row starts on a 128-byte boundary and a
// wavefront's 64 consecutive cells are one aligned 512-byte segment.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "mh_euler.h"

namespace mh {

constexpr int kNg = 2;       // ghost layers (src/test_cases.cpp:138, :151)
constexpr int kLead = 16;    // doubles before interior column 0 in each row

// Non-owning device view; `o` points at interior cell (0, 0) of plane 0.
struct View {
    double* o;
    int64_t ps;   // plane stride (doubles)
    int64_t rp;   // row pitch (doubles)
    int nx, ny;
};

inline int64_t round_up(int64_t a, int64_t b) { return (a + b - 1) / b * b; }

}  // namespace mh

struct mh_grid_s {
    mh_grid_desc d;
    double dx, dy;
    int64_t rp, ps;      // row pitch, plane stride (doubles)
    double* base;        // allocation start
    unsigned long long* aux;  // 8 device words of per-grid scratch (max-speed slot), or null
    // y-sweep redo list of tiles whose states left the fast-path range:
    // redo[0] count, redo[1] workgroups done, redo[2..] tile indices (null
    // for the grids of a batch)
    int* redo = nullptr;
    int64_t redo_cap = 0;
    int device;
    int owns;            // 1: base/aux were allocated for this grid (0: a view into a batch arena)
    mh::View view() const {
        mh::View v;
        v.o = base + (int64_t)mh::kNg * rp + mh::kLead;
        v.ps = ps;
        v.rp = rp;
        v.nx = d.nx;
        v.ny = d.ny;
        return v;
    }
};

namespace mh {
// Thread-local error string behind mh_last_error().
void set_error(const std::string& msg);
int fail(int code, const std::string& msg);
int check_hip(hipError_t e, const char* what);
}  // namespace mh
