// fused.hip -- the fused split step (fused_tile.h): its kernels and their
// launch.  Built with the machine scheduler's max-ilp strategy: the kernel
// holds four wavefronts per SIMD either way, and the FP64 chains of its x and
// y rows interleave better (0.7 % faster than iterative-minreg at 8192^2).
#include <hip/hip_runtime.h>

#include <algorithm>

#include "clock.h"
#include "fused_tile.h"
#include "grid.h"

namespace mh {

mh_launch_shape shape();   // solver.hip

// ---------------------------------------------------------------------------
// fused split step (fused_tile.h): one workgroup of two wavefronts per
// 64-column strip and chunk of rows; the x wavefront feeds the y wavefront
// through LDS.  A tile whose states leave the fast-path range (or a step
// whose dt/dx or dt/dy is outside it) goes on the output grid's redo list,
// and k_step_fused_redo runs its IEEE pass; the fast pass's speed maximum of
// such a tile is dropped.
// ---------------------------------------------------------------------------
struct FusedArgs {
    DtSrc dts;
    double dx, dy;
    int seg;           // rows per chunk
    int ngroups;       // 64-column strips
    int ntiles;
    int ghosts;
    Bc bc;
    unsigned long long* speed;
    int clock_speed;
    int* redo;
    int pp;            // views g0/g1 alternate as input/output with the clock's cur
    int big_tiles;     // tiles [0, big_tiles): chunks of seg rows from row 0
    int rbig;          // rows [rbig, ny): chunks of seg2 rows (tiles from big_tiles on)
    int seg2;
};

__device__ __forceinline__ unsigned long long* speed_slot(const FusedArgs& a) {
    return a.clock_speed ? &a.dts.clock->smax[a.dts.clock->slot ^ 1] : a.speed;
}

// Strip and rows [j0, j1) of a tile.
__device__ __forceinline__ void tile_span(const FusedArgs& a, int ny, int tile, int& cg, int& j0, int& j1) {
    const bool tail = tile >= a.big_tiles;
    const int t = tail ? tile - a.big_tiles : tile;
    cg = t % a.ngroups;
    j0 = tail ? a.rbig + (t / a.ngroups) * a.seg2 : (t / a.ngroups) * a.seg;
    j1 = min(j0 + (tail ? a.seg2 : a.seg), ny);
}

// The pair over rows [j0, j1) of strip cg.
template <int S, bool F>
__device__ __forceinline__ void fused_rows(const View& in, const View& out, const FusedArgs& a, int cg, int j0, int j1,
                                           double dt, PairLds& X, int w, int lane, bool& ok, double& m) {
    const double dtdx = dt / a.dx, dtdy = dt / a.dy;
    const int want = (a.speed || a.clock_speed) ? 1 : 0;
    ok = pair_side<S, F>(w, in, out, cg * 64, j0, j1, dtdx, dtdy, a.bc, a.ghosts, want, X, lane, m);
    if (w == 1) ok = ok && X.xok != 0u;
}

template <int S>
__global__ __launch_bounds__(128) __attribute__((amdgpu_waves_per_eu(4, 8))) void k_step_fused(View g0, View g1,
                                                                                             FusedArgs a) {
    double dt;
    if (!dt_of(a.dts, dt)) return;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    const int tile = blockIdx.x;
    if (tile >= a.ntiles) return;
    const bool swap = a.pp && a.dts.clock->cur;
    const View in = swap ? g1 : g0, out = swap ? g0 : g1;
    __shared__ PairLds X;
    if (!(dt_in_fast_range(dt / a.dx) && dt_in_fast_range(dt / a.dy))) {
        if (w == 1 && lane == 0) a.redo[2 + atomicAdd(a.redo, 1)] = tile;
        return;
    }
    bool ok = true;
    double m = 0.0;
    int cg, j0, j1;
    tile_span(a, in.ny, tile, cg, j0, j1);
    fused_rows<S, true>(in, out, a, cg, j0, j1, dt, X, w, lane, ok, m);
    if (w == 1) {
        if (!ok) {
            if (lane == 0) a.redo[2 + atomicAdd(a.redo, 1)] = tile;
        } else if (a.speed || a.clock_speed) {
            m = wave_max(m);
            if (lane == 0) atomic_max_pos(speed_slot(a), m);
        }
    }
}

// IEEE pass over the tiles k_step_fused put on the redo list.  A list too
// short to fill the chip splits each tile into pieces of at least kRedoRows
// rows (with their own warm-up rows) that run side by side, so a few redone
// tiles cost a few pieces' chains of rows rather than a whole tile's each
// (the exact solver's redo on the 8192^2 blast, about 180 tiles: 1.20 ms as
// whole tiles, 0.38 ms in pieces); a long list keeps whole tiles (no extra
// warm-up rows).  Both
// wavefronts of a workgroup walk the same pieces, with a barrier between
// pieces so the ring is free.  The list is emptied by the next step's clock
// kernel, before the next k_step_fused appends to it.
constexpr int kRedoRows = 16, kRedoPieces = 2048;
template <int S>
__global__ __launch_bounds__(128) void k_step_fused_redo(View g0, View g1, FusedArgs a) {
    __shared__ PairLds X;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    const int n = __hip_atomic_load(a.redo, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    double dt;
    if (n > 0 && dt_of(a.dts, dt)) {
        const bool swap = a.pp && a.dts.clock->cur;
        const View in = swap ? g1 : g0, out = swap ? g0 : g1;
        // pieces per tile, and rows per piece (every piece but a tile's
        // last has `rows` rows; a tile of fewer rows has fewer pieces)
        const int longest = max(a.seg, a.seg2);
        const int pieces = max(1, min((longest + kRedoRows - 1) / kRedoRows, (kRedoPieces + n - 1) / n));
        const int rows = (longest + pieces - 1) / pieces;
        // The exact solver's tiles are on the list for a non-uniform face,
        // which says nothing of their range: a piece first takes the full
        // solver with the fast division sequences (kExactFull, bit-identical
        // where its verdict holds) and the IEEE pass only if it fails.
        const bool fast = S == kExact && dt_in_fast_range(dt / a.dx) && dt_in_fast_range(dt / a.dy);
        for (int q = blockIdx.x; q < n * pieces; q += gridDim.x) {
            int cg, j0, j1;
            tile_span(a, in.ny, a.redo[2 + q / pieces], cg, j0, j1);
            const int p0 = j0 + (q % pieces) * rows;
            if (p0 < j1) {
                bool ok = false;
                double m = 0.0;
                if constexpr (S == kExact) {
                    if (fast) {
                        fused_rows<kExactFull, true>(in, out, a, cg, p0, min(p0 + rows, j1), dt, X, w, lane, ok, m);
                        if (w == 1 && lane == 0) X.verdict = ok ? 1u : 0u;
                        pair_barrier();
                        ok = X.verdict != 0u;
                    }
                }
                if (!ok) fused_rows<S, false>(in, out, a, cg, p0, min(p0 + rows, j1), dt, X, w, lane, ok, m);
                if (w == 1 && (a.speed || a.clock_speed)) {
                    m = wave_max(m);
                    if (lane == 0) atomic_max_pos(speed_slot(a), m);
                }
            }
            pair_barrier();
        }
    }
}

// The fused step applies with every solver (the exact solver's fast pass
// solves the uniform wavefronts and leaves the rest to the redo pass).
bool fused_applies(const mh_grid_s*, int) { return true; }

// Rows per chunk of the fused step: the configured y chunk, shortened (to
// no fewer than 16 rows: each chunk recomputes 4 warm-up rows) while the
// grid gives fewer pairs than min_tiles / 2 (the pairs hold two wavefronts
// each; by default 4096, two rounds of the eight a CU holds).
int seg_fused_for(const mh_grid_s* g, const mh_launch_shape& sh) {
    const int ngroups = (g->d.nx + 63) / 64, ny = g->d.ny;
    int seg = sh.seg_y > 0 ? sh.seg_y : ny;
    const int want = sh.min_tiles / 2;
    if ((int64_t)ngroups * ((ny + seg - 1) / seg) < want && seg > 16) {
        const int chunks = (want + ngroups - 1) / ngroups;
        seg = std::min(seg, std::max(16, (ny + chunks - 1) / chunks));
    }
    return seg;
}

// Tiles of the fused step: chunks of seg rows and, where there is room,
// the last rows of every strip in chunks of the shape's fused_tail rows,
// dispatched last.  The pairs run at very different speeds while they share
// a CU (the older wavefronts issue first), so with equal chunks the last
// round of a launch leaves CUs idle for up to a chunk's time; short chunks
// at the end let the CUs that finish first take the remaining rows.  The
// tail holds about min_tiles / 2 pairs (two rounds of the eight a CU holds)
// and at most half the rows: 8192^2 HLLC 1.750 -> 1.690 ms per launch with
// 512 rows of 16-row chunks after 60 chunks of 128 rows.
struct FusedPlan {
    int seg, ngroups, big_tiles, rbig, seg2, ntiles;
};
FusedPlan fused_plan(const mh_grid_s* g, const mh_launch_shape& sh) {
    FusedPlan p;
    const int ny = g->d.ny;
    p.seg = seg_fused_for(g, sh);
    p.ngroups = (g->d.nx + 63) / 64;
    p.big_tiles = p.ngroups * ((ny + p.seg - 1) / p.seg);
    p.rbig = ny;
    p.seg2 = p.seg;
    p.ntiles = p.big_tiles;
    if (sh.fused_tail > 0 && sh.fused_tail < p.seg) {
        const int64_t rows2 = (int64_t)((sh.min_tiles / 2 + p.ngroups - 1) / p.ngroups) * sh.fused_tail;
        if (rows2 > 0 && 2 * rows2 <= ny) {
            p.rbig = (int)((ny - rows2) / p.seg * p.seg);
            p.big_tiles = p.ngroups * (p.rbig / p.seg);
            p.seg2 = sh.fused_tail;
            p.ntiles = p.big_tiles + p.ngroups * ((ny - p.rbig + p.seg2 - 1) / p.seg2);
        }
    }
    return p;
}

int fused_tiles(const mh_grid_s* g) { return fused_plan(g, shape()).ntiles; }

// The redo pass: 256 workgroups walk the list (a few tiles outside the fast
// range); for the exact solver, whose fast pass leaves every tile with a
// non-uniform face to it, one workgroup per four tiles.
template <int S>
void go_fused(hipStream_t st, const View& v0, const View& v1, const FusedArgs& a) {
    hipLaunchKernelGGL(k_step_fused<S>, dim3((unsigned)a.ntiles), dim3(128), 0, st, v0, v1, a);
    const unsigned nredo = S == kExact ? (unsigned)std::max(256, (a.ntiles + 3) / 4) : 256u;
    hipLaunchKernelGGL(k_step_fused_redo<S>, dim3(nredo), dim3(128), 0, st, v0, v1, a);
}

// One fused step g0 -> g1 (pp = 0), or between g0 and g1 as the clock's cur
// says (pp = 1).  `redo` is a redo list with room for every tile, empty when
// the step starts (the stepper's clock kernel empties it).
int launch_step_fused(const mh_grid_s* g0, const mh_grid_s* g1, int solver, DtSrc dts, const Bc& bc, int ghosts,
                      unsigned long long* speed, int clock_speed, int pp, int* redo, hipStream_t st) {
    FusedArgs a;
    a.dts = dts;
    a.dx = g0->dx;
    a.dy = g0->dy;
    const FusedPlan p = fused_plan(g0, shape());
    a.seg = p.seg;
    a.ngroups = p.ngroups;
    a.ntiles = p.ntiles;
    a.big_tiles = p.big_tiles;
    a.rbig = p.rbig;
    a.seg2 = p.seg2;
    a.ghosts = ghosts;
    a.bc = bc;
    a.speed = speed;
    a.clock_speed = clock_speed;
    a.redo = redo;
    a.pp = pp;
    const View v0 = g0->view(), v1 = g1->view();
    switch (solver) {
        case kHLL: go_fused<kHLL>(st, v0, v1, a); break;
        case kHLLC: go_fused<kHLLC>(st, v0, v1, a); break;
        case kForce: go_fused<kForce>(st, v0, v1, a); break;
        case kExact: go_fused<kExact>(st, v0, v1, a); break;
        default: return fail(MH_ERR_INVALID, "fused step: unknown solver");
    }
    return check_hip(hipGetLastError(), "fused step launch");
}

}  // namespace mh
